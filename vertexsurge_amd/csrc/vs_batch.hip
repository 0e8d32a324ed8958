// Batched VExpand: many independent queries packed into shared bit-matrix
// tiles, expanded by one persistent kernel per hop, the whole launch
// sequence of a wave of tiles captured once into a hipGraph and replayed for
// every wave.  Also the tile-width / push-threshold autotuner.
//
// The reference has no batched path: each VExpand instance expands its own
// source list (src/objects/operator/vexpand.cpp:166-284).  Source rows are
// independent through every hop, the dedup and the union, so rows of
// different queries can share a 64-bit word as long as every row keeps its own
// hop range; that is what the per-word masks do:
//   live[i][w]  rows of word w with k_max >= i      (expand at hop i)
//   racc[i][w]  rows of word w with k_min <= i      (layer i joins the result)
// A Shortest row whose query has k_min <= 1 reads its result from the visited
// matrix (minus its source bit when k_min == 1), as the reference does for
// k_min == 0 (vexpand.cpp:306-310); otherwise layers accumulate into R.
//
// Per wave: init (slot -> tile, counters, zero visited/result, L0) ->
// for each hop: decide (push / pull per tile, on the device, heuristic or
// MLP) -> build push items -> plan work units -> persistent expand (pull
// vertex blocks and push items of every tile in one launch) -> push epilogue
// -> clear the consumed frontier; then per-row counts / digests into the
// caller's arrays and the wave counter advances.  No step reads the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "vs_kernels.hh"
#include "vs_sched.hh"

namespace vs {

constexpr uint32_t kMaxSlots = 64;

// Counters and front words are accumulated by every wave of the expand and
// epilogue kernels.  Thousands of waves adding into one 80-byte HopCounters
// (one L2 line) and 64 front words serialise on that line, so each wave
// adds into one of kParts partial copies (picked by its block), 128-byte
// apart; the next hop's decision kernel (or the finishing kernel for the
// last hop) folds the partials into the primary copy and zeroes them.
constexpr uint32_t kParts = 16;
constexpr size_t kPartStride = 128;  // bytes between partial HopCounters
static_assert(sizeof(HopCounters) <= kPartStride, "HopCounters outgrew its partial slot");

struct SlotDesc {
  uint64_t *F[2];
  uint64_t *Fm[2];   // word masks of F[0] / F[1] (V each)
  uint64_t *bm[2];
  uint64_t *N;
  uint64_t *touched;
  ull *items;
  uint64_t *Vis;
  uint64_t *R;
  HopCounters *ctr;  // [H + 1]
  uint64_t *front;   // [H + 1][64]: per-word OR of the rows of L_h
  char *part;        // [H + 1][kParts] partial HopCounters, kPartStride apart
  uint64_t *fpart;   // [H + 1][kParts][64] partial front words
  ull *wdeg;         // [H + 1][64]: per-word push degree of L_h
  ull *wpart;        // [H + 1][kWParts][64] partial per-word push degrees
  uint64_t pw;       // this hop: the words the word push scatters (pull slots)
  uint32_t tile;     // global tile of the current wave, UINT32_MAX if idle
  int32_t mode;      // this hop: 0 skip, 1 push, 2 pull
  uint32_t dense;    // this hop: pull without the active-bitmap lookup
  uint32_t masked;   // this hop: pull reads the frontier's word masks
};

struct TileInfo {
  uint32_t row0;
  uint32_t rows;
  uint32_t hmax;
  uint32_t word0;
};

struct BatchCtl {
  uint32_t wave;
  uint32_t pad;
  ull unit_prefix[kMaxSlots + 1];    // units of push slots and dense pull slots
  ull unit_prefix_s[kMaxSlots + 1];  // units of sparse pull slots (word masks)
  ull unit_prefix_w[kMaxSlots + 1];  // word-push items of pull slots
  ull te;
  ull push_dec;
  ull pull_dec;
  ull bytes;
  ull bytes_sparse;  // of the hops run by the sparse (word-mask) expansion
  ull wpush_words;   // words scattered by the word push
};

struct BatchConst {
  uint32_t V;
  uint32_t TW;
  uint32_t H;          // hops of the group (max k_max)
  uint32_t nslots;
  uint32_t ntiles;
  uint32_t Wpad;       // words per hop row of the mask arrays (ntiles * TW)
  uint32_t S;          // rows of the group
  int32_t sched;
  uint32_t alpha;
  uint32_t sparse_pct;  // word-mask threshold (percent of words, by degree)
  uint32_t use_r;
  uint64_t E;
  uint64_t wpush_budget;  // word push: push-degree budget per tile and hop (0: off)
  const uint64_t *poff;
  const uint32_t *padj;
  const uint64_t *xoff;
  const uint32_t *xadj;
  const uint32_t *src;    // [S]
  const uint8_t *sub_src; // [S] result = visited minus the source bit
  const uint64_t *live;   // [(H + 2) * Wpad]
  const uint64_t *racc;   // [(H + 1) * Wpad] or nullptr
  SlotDesc *slots;
  const TileInfo *tiles;
  BatchCtl *ctl;
  ull *out_counts;        // [S] (group order), zeroed before the first wave
  uint32_t want_rows;     // per-row counts / digests wanted: the row pass runs
  ull *out_digests;       // [S] or nullptr
  const float *mlp;       // scheduler parameters or nullptr
  const uint32_t *hubcnt; // hub neighbours at the head of each pull list
  const ull *pitems;      // static pull items of the pull CSR
  uint64_t npitems;
  uint32_t has_heavy;     // some in-list spans several pull items
  uint32_t nhwords;       // bitmap words holding such vertices
  const uint32_t *hwords;
  // Result matrices (vs_vexpand_batch_results), all nullptr otherwise:
  //  res_base  a one-query group expands straight into its result: tile t's
  //            result words (R, or Vis when no R is kept) live at
  //            res_base + t * V * TW
  //  wdest     a multi-query group copies each result word of global word w
  //            to query wdest[w].x's matrix qbase[x], at its word wdest[w].y
  //            (x == UINT32_MAX: a padding word)
  //  tile_front  per tile, hop and word: the OR of the rows of L_h (a row's
  //            bit is set when the row gained a vertex at hop h)
  uint64_t *res_base;
  const uint2 *wdest;
  uint64_t *const *qbase;
  uint64_t *tile_front;   // [ntiles][H + 1][TW]
  uint32_t res_sub;       // results: source bits of sub_src rows are cleared
                          // in the matrix, so counts and digests need no correction
};

constexpr uint32_t kNoRow = 0xFFFFFFFFu;  // src of a padding row (no source)

// The slot table is read at a wave-uniform slot index, but the compiler
// cannot tell (the index comes from a search) and would keep every field in
// vector registers; readfirstlane moves them to scalar registers.
__device__ __forceinline__ uint32_t rfl32(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}
__device__ __forceinline__ uint64_t rfl64(uint64_t x) {
  return static_cast<uint64_t>(rfl32(static_cast<uint32_t>(x))) |
         (static_cast<uint64_t>(rfl32(static_cast<uint32_t>(x >> 32))) << 32);
}

// A pointer field of the slot table, read as a wave-uniform global-memory
// pointer: every slot buffer is a device allocation, and saying so lets the
// compiler use global (not flat) loads, stores and atomics through it.
template <class T>
__device__ __forceinline__ T *gp(T *const &field) {
  using G = __attribute__((address_space(1))) T *;
  const G g = *reinterpret_cast<const G *>(reinterpret_cast<const void *>(&field));
  return (T *)(G)rfl64((uint64_t)g);
}
// The same for a field read at a per-thread index (no readfirstlane).
template <class T>
__device__ __forceinline__ T *gpn(T *const &field) {
  using G = __attribute__((address_space(1))) T *;
  return (T *)*reinterpret_cast<const G *>(reinterpret_cast<const void *>(&field));
}

__device__ __forceinline__ HopArgs slot_args(const BatchConst &c, const SlotDesc &sd,
                                             uint32_t hop) {
  HopArgs a;
  const uint32_t tile = rfl32(sd.tile);
  const uint32_t word0 = rfl32(c.tiles[tile].word0), hmax = rfl32(c.tiles[tile].hmax);
  a.V = c.V;
  a.S = c.S;
  a.word0 = word0;
  a.dense = rfl32(sd.dense);
  a.poff = c.poff;
  a.padj = c.padj;
  a.xoff = c.xoff;
  a.xadj = c.xadj;
  a.F = gp(sd.F[(hop - 1) & 1]);
  a.Fm = rfl32(sd.masked) ? gp(sd.Fm[(hop - 1) & 1]) : nullptr;
  a.actbm = gp(sd.bm[(hop - 1) & 1]);
  // the tile's last hop: its frontier is never read, so it is not written
  // (and needs no clearing)
  const bool last = hop >= hmax;
  a.Fn = last ? nullptr : gp(sd.F[hop & 1]);
  a.Fnm = last ? nullptr : gp(sd.Fm[hop & 1]);
  a.actbm_n = last ? nullptr : gp(sd.bm[hop & 1]);
  a.N = gp(sd.N);
  a.touched = gp(sd.touched);
  a.Vis = gp(sd.Vis);
  a.R = c.use_r ? gp(sd.R) : nullptr;
  a.items = gp(sd.items);
  a.nitems = &gp(sd.ctr)[hop - 1].items;
  const size_t part = static_cast<size_t>(hop) * kParts + blockIdx.x % kParts;
  a.ctr = reinterpret_cast<HopCounters *>(gp(sd.part) + part * kPartStride);
  a.live = c.live + static_cast<size_t>(hop) * c.Wpad + word0;
  a.live_next = c.live + static_cast<size_t>(hop + 1) * c.Wpad + word0;
  a.racc = c.racc ? c.racc + static_cast<size_t>(hop) * c.Wpad + word0 : nullptr;
  a.pitems = c.pitems;
  a.hubcnt = c.hubcnt;
  a.npitems = c.npitems;
  a.ewords = nullptr;
  a.newords = 0;
  a.front = gp(sd.front) + static_cast<size_t>(hop - 1) * 64;
  a.front_next = gp(sd.fpart) + part * 64;
  a.wdeg = gp(sd.wpart) + static_cast<size_t>(hop) * kWParts * 64;
  a.pw = rfl32(sd.mode) == 2 ? rfl64(sd.pw) : 0;
  a.Fwm = gp(sd.Fm[(hop - 1) & 1]);
  return a;
}

// Adds the accumulated fields of b into a (not items / how).
__device__ __forceinline__ void add_ctr(HopCounters &a, const HopCounters &b) {
  a.pop += b.pop;
  a.nz += b.nz;
  a.act += b.act;
  a.actdeg += b.actdeg;
  a.te += b.te;
  a.visnew += b.visnew;
  a.rnew += b.rnew;
  a.nzdeg += b.nzdeg;
}

__device__ __forceinline__ HopCounters *part_ctr(const SlotDesc &sd, uint32_t h, uint32_t p) {
  return reinterpret_cast<HopCounters *>(sd.part + (static_cast<size_t>(h) * kParts + p) *
                                                       kPartStride);
}

// One wave: folds hop h's partial counters, front words and per-word push
// degrees into the primary copy, zeroes the partials, and returns the folded
// counters to every lane (and lane j word j's push degree in wd).
__device__ HopCounters fold_parts(const SlotDesc &sd, uint32_t h, uint64_t &wd) {
  const int lane = lane_id();
  HopCounters q{};
  if (lane < static_cast<int>(kParts)) {
    HopCounters *pp = part_ctr(sd, h, lane);
    q = *pp;
    *pp = HopCounters{};
  }
  HopCounters t = sd.ctr[h];
  HopCounters w{};
  w.pop = wave_sum(q.pop);
  w.nz = wave_sum(q.nz);
  w.act = wave_sum(q.act);
  w.actdeg = wave_sum(q.actdeg);
  w.te = wave_sum(q.te);
  w.visnew = wave_sum(q.visnew);
  w.rnew = wave_sum(q.rnew);
  w.nzdeg = wave_sum(q.nzdeg);
  add_ctr(t, w);
  uint64_t f = 0;
  uint64_t *fp = sd.fpart + static_cast<size_t>(h) * kParts * 64 + lane;
  for (uint32_t p = 0; p < kParts; ++p) {
    const uint64_t x = fp[p * 64];
    if (x) {
      f |= x;
      fp[p * 64] = 0;
    }
  }
  if (f) sd.front[static_cast<size_t>(h) * 64 + lane] |= f;
  uint64_t d = 0;
  ull *wp = sd.wpart + static_cast<size_t>(h) * kWParts * 64 + lane;
  for (uint32_t p = 0; p < kWParts; ++p) {
    const uint64_t x = wp[p * 64];
    if (x) {
      d += x;
      wp[p * 64] = 0;
    }
  }
  wd = sd.wdeg[static_cast<size_t>(h) * 64 + lane] + d;
  if (d) sd.wdeg[static_cast<size_t>(h) * 64 + lane] = wd;
  if (lane == 0) add_ctr(sd.ctr[h], w);
  return t;
}

// ---- wave init: slot -> tile, zero counters (one block)
__global__ void k_bw_assign(BatchConst c) {
  const uint32_t wave = c.ctl->wave;
  for (uint32_t s = threadIdx.x; s < c.nslots; s += blockDim.x) {
    const uint32_t t = wave * c.nslots + s;
    c.slots[s].tile = t < c.ntiles ? t : UINT32_MAX;
    c.slots[s].mode = 0;
    if (c.res_base && t < c.ntiles) {
      uint64_t *m = c.res_base + static_cast<size_t>(t) * c.V * c.TW;
      if (c.use_r) c.slots[s].R = m;
      else c.slots[s].Vis = m;
    }
  }
  for (uint32_t i = threadIdx.x; i < c.nslots * (c.H + 1); i += blockDim.x) {
    HopCounters *h = &c.slots[i / (c.H + 1)].ctr[i % (c.H + 1)];
    *h = HopCounters{};
  }
  for (uint32_t i = threadIdx.x; i < c.nslots * (c.H + 1) * 64; i += blockDim.x) {
    c.slots[i / ((c.H + 1) * 64)].front[i % ((c.H + 1) * 64)] = 0;
    c.slots[i / ((c.H + 1) * 64)].wdeg[i % ((c.H + 1) * 64)] = 0;
  }
}

// ---- zero visited / result matrices of active slots (fill kernel: no
// memset node inside the captured graph)
// Each block zeroes chunks of kZeroChunk 16-byte words (the launch covers
// the matrix with one chunk per block where it can).
constexpr size_t kZeroChunk = static_cast<size_t>(kBlock) * 2;

__global__ __launch_bounds__(kBlock) void k_bw_zero(BatchConst c) {
  const size_t words = static_cast<size_t>(c.V) * c.TW;
  const size_t n4 = words / 2;
  for (uint32_t s = 0; s < c.nslots; ++s) {
    const SlotDesc &sd = c.slots[s];
    if (sd.tile == UINT32_MAX) continue;
    for (int m = 0; m < 2; ++m) {
      uint64_t *M = m == 0 ? gpn(sd.Vis) : (c.use_r ? gpn(sd.R) : nullptr);
      if (!M) continue;
      // 16-byte stores (the matrices are hipMalloc'd, 256-byte aligned)
      uint4 *M4 = reinterpret_cast<uint4 *>(M);
      for (size_t b0 = blockIdx.x * kZeroChunk; b0 < n4; b0 += gridDim.x * kZeroChunk) {
        const size_t e = b0 + kZeroChunk < n4 ? b0 + kZeroChunk : n4;
        for (size_t i = b0 + threadIdx.x; i < e; i += kBlock) M4[i] = make_uint4(0, 0, 0, 0);
      }
      if ((words & 1) && blockIdx.x == 0 && threadIdx.x == 0) M[words - 1] = 0;
    }
  }
}

// ---- L0 of every active slot, with its counters
__global__ void k_bw_l0(BatchConst c) {
  const uint32_t rows_per_tile = c.TW * 64;
  const uint32_t total = c.nslots * rows_per_tile;
  const int lane = lane_id();
  // The loop is wave-uniform (total and the stride are multiples of 64), and
  // a wave's 64 rows share one slot and one word j: the counters take one
  // atomic per wave each, not one per row (thousands of same-address atomics
  // otherwise serialise in L2).
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += gridDim.x * blockDim.x) {
    const uint32_t s = i / rows_per_tile, lr = i % rows_per_tile;
    const SlotDesc &sd = c.slots[s];
    if (sd.tile == UINT32_MAX) continue;  // uniform: the wave's slot
    const TileInfo ti = c.tiles[sd.tile];
    const uint32_t j = lr / 64;
    const ull bit = 1ull << (lr % 64);
    const uint32_t v = lr < ti.rows ? c.src[ti.row0 + lr] : kNoRow;  // kNoRow: padding
    ull nz = 0, nzdeg = 0, visnew = 0, rnew = 0, act = 0, actdeg = 0, pop = 0, te = 0;
    uint64_t fb = 0;
    if (v != kNoRow) {
      const size_t wi = static_cast<size_t>(v) * c.TW + j;
      const uint64_t w_live1 = c.live[static_cast<size_t>(1) * c.Wpad + ti.word0 + j];
      const uint64_t deg = c.xoff[v + 1] - c.xoff[v];
      const ull oldF = atomicOr(reinterpret_cast<ull *>(&gpn(sd.F[0])[wi]), bit);
      if (oldF == 0) {
        nz = 1;
        nzdeg = deg;
      }
      atomicOr(reinterpret_cast<ull *>(&gpn(sd.Fm[0])[v]), 1ull << j);
      if (sd.Vis) {
        const ull oldV = atomicOr(reinterpret_cast<ull *>(&gpn(sd.Vis)[wi]), bit);
        visnew = oldV == 0;
      }
      if (c.use_r && c.racc && (c.racc[ti.word0 + j] & bit)) {
        const ull oldR = atomicOr(reinterpret_cast<ull *>(&gpn(sd.R)[wi]), bit);
        rnew = oldR == 0;
      }
      const ull oldB =
          atomicOr(reinterpret_cast<ull *>(&gpn(sd.bm[0])[v / 64]), 1ull << (v % 64));
      if (!(oldB & (1ull << (v % 64)))) {
        act = 1;
        actdeg = deg;
      }
      pop = 1;
      fb = bit;
      if (w_live1 & bit) te = deg;
    }
    nz = wave_sum(nz);
    nzdeg = wave_sum(nzdeg);
    visnew = wave_sum(visnew);
    rnew = wave_sum(rnew);
    act = wave_sum(act);
    actdeg = wave_sum(actdeg);
    pop = wave_sum(pop);
    te = wave_sum(te);
    fb = wave_or(fb);
    if (lane == 0 && pop) {
      HopCounters *h = &sd.ctr[0];
      if (nz) atomicAdd(&h->nz, nz);
      if (nzdeg) {
        atomicAdd(&h->nzdeg, nzdeg);
        atomicAdd(&sd.wdeg[j], nzdeg);
      }
      if (visnew) atomicAdd(&h->visnew, visnew);
      if (rnew) atomicAdd(&h->rnew, rnew);
      if (act) atomicAdd(&h->act, act);
      if (actdeg) atomicAdd(&h->actdeg, actdeg);
      atomicAdd(&h->pop, pop);
      atomicOr(reinterpret_cast<ull *>(&sd.front[j]), static_cast<ull>(fb));
      if (te) atomicAdd(&h->te, te);
    }
  }
}

// ---- per-hop decision, one wave per slot
__global__ __launch_bounds__(64) void k_bw_decide(BatchConst c, uint32_t hop) {
  __shared__ float s_x[kIn];
  __shared__ float s_h[2 * kHid];
  const uint32_t s = blockIdx.x;
  SlotDesc &sd = c.slots[s];
  const int lane = lane_id();
  int mode = 0;
  uint64_t wd = 0;  // lane j: push degree of word j of L_{hop-1}
  const HopCounters pc = fold_parts(sd, hop - 1, wd);
  bool wpush_ok = false;  // the tile writes a frontier this hop
  if (sd.tile != UINT32_MAX) {
    const TileInfo ti = c.tiles[sd.tile];
    wpush_ok = hop < ti.hmax;
    if (pc.pop > 0 && hop <= ti.hmax) {
      if (c.sched == VS_SCHED_PUSH) mode = 1;
      else if (c.sched == VS_SCHED_PULL) mode = 2;
      else {
        const int heur = (pc.actdeg * static_cast<uint64_t>(c.alpha) < c.E) ? 1 : 2;
        mode = heur;
        if (c.sched == VS_SCHED_LEARNED && c.mlp) {
          uint64_t visnz = pc.visnew;
          for (uint32_t h = 0; h + 1 < hop; ++h) visnz += sd.ctr[h].visnew;
          if (lane == 0) {
            s_x[0] = log2f(1.f + static_cast<float>(pc.pop)) / 32.f;
            s_x[1] = log2f(1.f + static_cast<float>(pc.act)) / 32.f;
            s_x[2] = log2f(1.f + static_cast<float>(pc.actdeg)) / 40.f;
            s_x[3] = static_cast<float>(hop) / 8.f;
            s_x[4] = static_cast<float>(static_cast<double>(pc.pop) /
                                        (static_cast<double>(ti.rows) * c.V));
            s_x[5] = log2f(1.f + static_cast<float>(c.E)) / 40.f;
            s_x[6] = log2f(1.f + static_cast<float>(c.V)) / 32.f;
            s_x[7] = sd.Vis ? static_cast<float>(static_cast<double>(visnz) /
                                                (static_cast<double>(c.V) * c.TW))
                           : 0.f;
          }
          wave_sync();
          if (mlp_in_domain_wave(c.mlp, s_x)) {  // else the heuristic decides
            float h1, h2, y;
            mlp_forward_wave(c.mlp, s_x, s_h, h1, h2, y);
            const float tp = __shfl(y, kOptPush), tl = __shfl(y, kOptPull);
            if (isfinite(tp) && isfinite(tl) && fabsf(tp - tl) >= 0.1f) mode = tp < tl ? 1 : 2;
          }
        }
      }
    }
  }
  const bool masked = pc.nzdeg * 100 < pc.actdeg * c.TW * c.sparse_pct;
  // word push (pull over whole rows, frontier written this hop): the words
  // with push degree > 0, cheapest first (ties by word), within the budget
  uint64_t pw = 0;
  if (mode == 2 && !masked && wpush_ok && c.wpush_budget) {
    const uint64_t w = lane < static_cast<int>(c.TW) ? wd : 0;
    uint64_t cum = 0;
    for (int k = 0; k < 64; ++k) {
      const uint64_t wk = shfl64(w, k);
      if (wk && (wk < w || (wk == w && k <= lane))) cum += wk;
    }
    pw = __ballot(w != 0 && cum <= c.wpush_budget);
  }
  if (lane == 0) {
    sd.mode = mode;
    sd.pw = pw;
    if (pw) atomicAdd(&c.ctl->wpush_words, static_cast<ull>(__popcll(pw)));
    // most in-edges come from active vertices: pull skips the bitmap lookup
    sd.dense = (mode == 2 && pc.actdeg * 4 > 3 * c.E) ? 1u : 0u;
    // rows sparse in words: gather only the words their masks name
    sd.masked = masked ? 1u : 0u;
    sd.ctr[hop].how = (mode == 2 && sd.masked) ? 3u : static_cast<ull>(mode);
    if (mode == 1) atomicAdd(&c.ctl->push_dec, 1ull);
    if (mode == 2) atomicAdd(&c.ctl->pull_dec, 1ull);
  }
}

// ---- push items for push slots (whole blocks per bitmap slice)
__global__ __launch_bounds__(kBlock) void k_bw_items(BatchConst c, uint32_t hop) {
  __shared__ ull s_wsum[kWavesPerBlock];
  __shared__ ull s_base;
  const uint32_t nvb = (c.V + 63) / 64;
  const uint32_t nslice = (nvb + kBlock - 1) / kBlock;
  for (uint32_t u = blockIdx.x; u < c.nslots * nslice; u += gridDim.x) {
    const uint32_t s = u / nslice;
    const SlotDesc &sd = c.slots[s];
    // uniform across the block: push slots, and pull slots with a word push
    // (only the vertices holding bits in a pushed word)
    const bool wp = sd.mode == 2 && sd.pw;
    if (sd.mode != 1 && !wp) continue;
    items_slice(gpn(sd.bm[(hop - 1) & 1]), nvb, (u % nslice) * kBlock, c.xoff, gpn(sd.items),
                &gpn(sd.ctr)[hop - 1].items, s_wsum, &s_base, wp ? gpn(sd.Fm[(hop - 1) & 1]) : nullptr,
                wp ? sd.pw : 0);
  }
}

// ---- unit prefix over slots (one thread)
__global__ void k_bw_plan(BatchConst c, uint32_t hop) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t nvb = (c.V + 63) / 64;
  ull acc = 0, acc_s = 0, acc_w = 0;
  for (uint32_t s = 0; s < c.nslots; ++s) {
    c.ctl->unit_prefix[s] = acc;
    c.ctl->unit_prefix_s[s] = acc_s;
    c.ctl->unit_prefix_w[s] = acc_w;
    const SlotDesc &sd = c.slots[s];
    if (sd.mode == 2 && sd.masked) acc_s += c.npitems;
    else if (sd.mode == 2) acc += c.npitems;
    else if (sd.mode == 1) acc += sd.ctr[hop - 1].items;
    if (sd.mode == 2 && sd.pw) acc_w += sd.ctr[hop - 1].items;
  }
  c.ctl->unit_prefix[c.nslots] = acc;
  c.ctl->unit_prefix_s[c.nslots] = acc_s;
  c.ctl->unit_prefix_w[c.nslots] = acc_w;
}

// ---- the persistent expansion: pull vertex-blocks and push items of all
// slots; work unit u belongs to the slot whose prefix range holds it.  Two
// launches per hop: SPARSE takes the pull slots whose frontier rows are read
// through their word masks, the other one the push and dense pull slots.
template <int TW, bool SHORTEST, bool SPARSE>
__global__ __launch_bounds__(kBlock)
void k_bw_expand(BatchConst c, uint32_t hop) {
  __shared__ uint32_t s_nbr[kWavesPerBlock][64];
  __shared__ uint64_t s_nm[kWavesPerBlock][SPARSE ? 64 : 1];
  __shared__ uint16_t s_pairs[kWavesPerBlock][SPARSE ? kPairCap : 1];
  __shared__ uint64_t s_acc[kWavesPerBlock][64];
  __shared__ uint8_t s_wl[kWavesPerBlock][64];
  const ull *prefix = SPARSE ? c.ctl->unit_prefix_s : c.ctl->unit_prefix;
  const int wid = wave_id();
  const ull total = prefix[c.nslots];
  const ull nwaves = static_cast<ull>(gridDim.x) * kWavesPerBlock;
  const int j = lane_id() % TW;
  // Slots in order: a slot's units are a contiguous range [sb, se) of the
  // unit space and this wave's units (stride nwaves) only grow, so each slot
  // is visited once, with its arguments fixed for the inner loop.
  ull u = blockIdx.x * kWavesPerBlock + wid;
  for (uint32_t s = 0; s < c.nslots && u < total; ++s) {
    const ull sb = prefix[s], se = prefix[s + 1];
    if (u >= se) continue;
    const HopArgs a = slot_args(c, c.slots[s], hop);
    const int mode = __builtin_amdgcn_readfirstlane(c.slots[s].mode);
    if (mode == 2) {
      const LaneMasks mk = lane_masks(a, j);
      LocalCtr lc;
      // Items in flight per wave: the head (visited word, list bounds) of a
      // later item, the first list ids of the next one -- and at tile width
      // 64 their active-bitmap words (word-mask frontiers and bitmap-tested
      // dense ones) one item ahead, the ids two ahead, the head three ahead.
      // Narrower tiles keep the shallower pipeline: their wider lane groups
      // leave no registers for it at four waves per SIMD.
      constexpr bool kDeep = TW == 64;
      const bool bmp = kDeep && (SPARSE || a.dense == 0);
      constexpr ull D = kDeep ? 3 : 2;  // items from a head load to its body
      PullHead h0 = pull_head<TW, SHORTEST>(a, a.pitems[u - sb], j);
      PullHead h1{}, h2{};
      if (u + nwaves < se) h1 = pull_head<TW, SHORTEST>(a, a.pitems[u + nwaves - sb], j);
      if (kDeep && u + 2 * nwaves < se)
        h2 = pull_head<TW, SHORTEST>(a, a.pitems[u + 2 * nwaves - sb], j);
      uint32_t u0 = 0, u1 = 0, bw0 = 0;
      bool p0 = pull_prefetch<TW, SPARSE>(a, h0, mk, j, u0);
      if (p0 && bmp) bw0 = pull_prefetch_bm(a, h0, u0);
      bool p1 = kDeep && u + nwaves < se && pull_prefetch<TW, SPARSE>(a, h1, mk, j, u1);
      for (; u < se; u += nwaves) {
        const ull nh = u + D * nwaves;
        PullHead hl{};
        if (nh < se) hl = pull_head<TW, SHORTEST>(a, a.pitems[nh - sb], j);
        uint32_t un = 0, bw1 = 0;
        bool pn;
        if constexpr (kDeep) {
          pn = u + 2 * nwaves < se && pull_prefetch<TW, SPARSE>(a, h2, mk, j, un);
          if (p1 && bmp) bw1 = pull_prefetch_bm(a, h1, u1);
        } else {
          pn = u + nwaves < se && pull_prefetch<TW, SPARSE>(a, h1, mk, j, un);
        }
        // word-mask pulls scan every in-edge: their next group's ids load
        // during the current group's gathers
        pull_body<TW, SHORTEST, SPARSE, SPARSE && TW == 64>(
            a, h0, s_nbr[wid], s_nm[wid], s_pairs[wid], s_acc[wid], s_wl[wid], mk, lc, u0, p0,
            bw0, p0 && bmp);
        h0 = h1;
        if constexpr (kDeep) {
          h1 = h2;
          h2 = hl;
          u0 = u1;
          u1 = un;
          p0 = p1;
          p1 = pn;
          bw0 = bw1;
        } else {
          h1 = hl;
          u0 = un;
          p0 = pn;
        }
      }
      flush_ctr(lc, a.ctr, a.front_next, a.wdeg, j);
    } else {
      for (; u < se; u += nwaves)
        if (!SPARSE) push_item<TW>(a, a.items[u - sb]);
    }
  }
}

// ---- word push of the pull slots that have one: persistent over their
// items, before the pull
template <int TW>
__global__ __launch_bounds__(kBlock) void k_bw_wpush(BatchConst c, uint32_t hop) {
  const ull *prefix = c.ctl->unit_prefix_w;
  const ull total = prefix[c.nslots];
  const ull nwaves = static_cast<ull>(gridDim.x) * kWavesPerBlock;
  uint32_t cur = UINT32_MAX;
  HopArgs a;
  ull sb = 0, se = 0;
  for (ull u = blockIdx.x * kWavesPerBlock + wave_id(); u < total; u += nwaves) {
    if (cur == UINT32_MAX || u < sb || u >= se) {
      uint32_t s = cur == UINT32_MAX ? 0 : cur;
      while (s + 1 <= c.nslots && prefix[s + 1] <= u) ++s;
      while (prefix[s] > u) --s;
      sb = prefix[s];
      se = prefix[s + 1];
      cur = s;
      a = slot_args(c, c.slots[s], hop);
    }
    wpush_item<TW>(a, a.items[u - sb]);
  }
}

// ---- push epilogue and frontier clear, grid-stride over (slot, vertex block)
template <int TW, bool SHORTEST>
__global__ __launch_bounds__(kBlock) void k_bw_epi(BatchConst c, uint32_t hop) {
  const int wid = wave_id();
  const uint32_t nvb = (c.V + 63) / 64;
  const ull total = static_cast<ull>(c.nslots) * nvb;
  const ull nwaves = static_cast<ull>(gridDim.x) * kWavesPerBlock;
  uint32_t cur = UINT32_MAX;
  HopArgs a;
  LaneMasks mk;
  LocalCtr lc;
  const ull w = blockIdx.x * kWavesPerBlock + wid;
  // units only grow along a wave's rounds, so slots are still visited in order
  for (ull k = 0; k * nwaves < total; ++k) {
    const ull u = skewed_unit(k, w, nwaves);
    if (u >= total) continue;
    const uint32_t s = static_cast<uint32_t>(u / nvb);
    // push slots (every bitmap word), and pull slots whose chunked vertices
    // left partial words in N (only the words that hold such vertices)
    const int md = c.slots[s].mode;
    if (md != 1 && !(md == 2 && c.has_heavy)) continue;
    const uint32_t i = static_cast<uint32_t>(u % nvb);
    if (md == 2 && i >= c.nhwords) continue;
    if (s != cur) {
      if (cur != UINT32_MAX) flush_ctr(lc, a.ctr, a.front_next, a.wdeg, lane_id() % TW);
      cur = s;
      a = slot_args(c, c.slots[s], hop);
      mk = lane_masks(a, lane_id() % TW);
    }
    epi_block<TW, SHORTEST>(a, md == 2 ? c.hwords[i] : i, mk, lc);
  }
  if (cur != UINT32_MAX) flush_ctr(lc, a.ctr, a.front_next, a.wdeg, lane_id() % TW);
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_bw_clear(BatchConst c, uint32_t which) {
  const int wid = wave_id();
  const uint32_t nvb = (c.V + 63) / 64;
  const ull total = static_cast<ull>(c.nslots) * nvb;
  const ull nwaves = static_cast<ull>(gridDim.x) * kWavesPerBlock;
  const ull w = blockIdx.x * kWavesPerBlock + wid;
  for (ull k = 0; k * nwaves < total; ++k) {
    const ull u = skewed_unit(k, w, nwaves);
    if (u >= total) continue;
    const SlotDesc &sd = c.slots[u / nvb];
    if (sd.tile == UINT32_MAX) continue;
    clear_block_masked<TW>(gpn(sd.F[which & 1]), gpn(sd.Fm[which & 1]), gpn(sd.bm[which & 1]),
                           static_cast<uint32_t>(u % nvb));
  }
}

// Carry-save adder: h:l = a + b + c, bitwise.
__device__ __forceinline__ void csa(uint64_t &h, uint64_t &l, uint64_t a, uint64_t b,
                                    uint64_t c) {
  const uint64_t u = a ^ b;
  h = (a & b) | (u & c);
  l = u ^ c;
}

// Adds x, weighted 2^K, into a bit-sliced counter (plane k = bit k of every
// bit position's count).
template <int NB, int K>
__device__ __forceinline__ void vadd(uint64_t (&cnt)[NB], uint64_t x) {
#pragma unroll
  for (int k = K; k < NB; ++k) {
    const uint64_t t = cnt[k] & x;
    cnt[k] ^= x;
    x = t;
  }
}

// ---- per-row popcount and digest of each active slot's result
// Each lane counts, per bit position, the set bits of its word over the
// chunk's vertices: words are combined 16 at a time by a carry-save tree
// (ones / twos / fours / eights carried across groups, one sixteens word per
// group added into the bit-sliced counter), about 6 logic ops per word
// instead of 3 per counter plane.  The four waves of a block work on chunks
// of one slot at a time and add their per-row counts into LDS; the block
// adds the LDS counts to the output once per slot, not once per chunk.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_bw_rows(BatchConst c) {
  constexpr int G = 64 / TW;
  constexpr uint32_t CV = 2048;
  constexpr int NB = 12;  // counts up to CV = 2^11
  // counts of the block's current slot: row j * 64 + b at [b * 64 + j], so
  // the lanes of one bit b add to consecutive banks
  __shared__ uint32_t s_cnt[64 * 64];
  const int lane = lane_id(), wid = wave_id();
  const int g = lane / TW, j = lane % TW;
  const uint32_t nch = (c.V + CV - 1) / CV;
  const uint32_t nchp = (nch + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
  const ull total = static_cast<ull>(c.nslots) * nchp;  // chunks, padded per slot
  for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
  uint32_t cur = UINT32_MAX;  // slot whose counts s_cnt holds (block-uniform)
  auto flush = [&](uint32_t s) {
    __syncthreads();
    const SlotDesc &sd = c.slots[s];
    const TileInfo ti = c.tiles[sd.tile];
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
      const uint32_t n = s_cnt[i];
      if (n) {
        const uint32_t lr = (i % 64) * 64 + i / 64;
        if (lr < ti.rows) atomicAdd(&c.out_counts[ti.row0 + lr], static_cast<ull>(n));
        s_cnt[i] = 0;
      }
    }
    __syncthreads();
  };
  for (ull q = static_cast<ull>(blockIdx.x) * kWavesPerBlock; q < total;
       q += static_cast<ull>(gridDim.x) * kWavesPerBlock) {
    const uint32_t s = static_cast<uint32_t>(q / nchp);  // same slot for the block
    const SlotDesc &sd = c.slots[s];
    if (sd.tile == UINT32_MAX) continue;
    if (s != cur) {
      if (cur != UINT32_MAX) flush(cur);
      cur = s;
    }
    const uint32_t ch = static_cast<uint32_t>((q + wid) % nchp);
    if (ch >= nch) continue;
    const uint64_t *M = c.use_r ? gpn(sd.R) : gpn(sd.Vis);
    const uint32_t v0 = ch * CV, v1 = min(c.V, v0 + CV);
    uint64_t cnt[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) cnt[k] = 0;
    uint64_t ones = 0, twos = 0, fours = 0, eights = 0;
    uint32_t v = v0 + g;
    for (; v + 15 * G < v1; v += 16 * G) {
      uint64_t d[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) d[q] = M[static_cast<size_t>(v + q * G) * TW + j];
      uint64_t twosA, twosB, foursA, foursB, eightsA, eightsB, sixteens;
      csa(twosA, ones, ones, d[0], d[1]);
      csa(twosB, ones, ones, d[2], d[3]);
      csa(foursA, twos, twos, twosA, twosB);
      csa(twosA, ones, ones, d[4], d[5]);
      csa(twosB, ones, ones, d[6], d[7]);
      csa(foursB, twos, twos, twosA, twosB);
      csa(eightsA, fours, fours, foursA, foursB);
      csa(twosA, ones, ones, d[8], d[9]);
      csa(twosB, ones, ones, d[10], d[11]);
      csa(foursA, twos, twos, twosA, twosB);
      csa(twosA, ones, ones, d[12], d[13]);
      csa(twosB, ones, ones, d[14], d[15]);
      csa(foursB, twos, twos, twosA, twosB);
      csa(eightsB, fours, fours, foursA, foursB);
      csa(sixteens, eights, eights, eightsA, eightsB);
      vadd<NB, 4>(cnt, sixteens);
    }
    for (; v < v1; v += G) vadd<NB, 0>(cnt, M[static_cast<size_t>(v) * TW + j]);
    vadd<NB, 0>(cnt, ones);
    vadd<NB, 1>(cnt, twos);
    vadd<NB, 2>(cnt, fours);
    vadd<NB, 3>(cnt, eights);
    for (int b = 0; b < 64; ++b) {
      uint32_t n = 0;
#pragma unroll
      for (int k = 0; k < NB; ++k) n |= static_cast<uint32_t>((cnt[k] >> b) & 1) << k;
#pragma unroll
      for (int m = TW; m < 64; m <<= 1) n += __shfl_xor(n, m);
      if (g == 0 && n) atomicAdd(&s_cnt[b * 64 + j], n);
    }
  }
  if (cur != UINT32_MAX) flush(cur);
  if (!c.out_digests) return;
  const ull nwaves = static_cast<ull>(gridDim.x) * kWavesPerBlock;
  // digests: lane = row bit, walk 64-vertex blocks of each word
  const uint32_t nvb = (c.V + 63) / 64;
  constexpr uint32_t BPW = 16;
  const uint32_t nitem_v = (nvb + BPW - 1) / BPW;
  const ull total2 = static_cast<ull>(c.nslots) * TW * nitem_v;
  for (ull it = blockIdx.x * kWavesPerBlock + wid; it < total2; it += nwaves) {
    const uint32_t iv = static_cast<uint32_t>(it % nitem_v);
    const uint32_t sj = static_cast<uint32_t>(it / nitem_v);
    const SlotDesc &sd = c.slots[sj / TW];
    if (sd.tile == UINT32_MAX) continue;
    const TileInfo ti = c.tiles[sd.tile];
    const uint32_t jj = sj % TW;
    if (jj * 64 >= ti.rows) continue;
    const uint64_t *M = (c.use_r ? gpn(sd.R) : gpn(sd.Vis)) + jj;
    uint64_t acc = 0;
    for (uint32_t vb = iv * BPW; vb < min(nvb, (iv + 1) * BPW); ++vb) {
      const uint32_t v = vb * 64 + lane;
      const uint64_t mine = v < c.V ? M[static_cast<size_t>(v) * TW] : 0;
      uint64_t any = __ballot(mine != 0);
      while (any) {
        const int l = __ffsll(static_cast<long long>(any)) - 1;
        any &= any - 1;
        const uint64_t x = shfl64(mine, l);
        if ((x >> lane) & 1) acc += digest_hash(vb * 64 + l);
      }
    }
    const uint32_t lr = jj * 64 + lane;
    if (acc && lr < ti.rows) atomicAdd(&c.out_digests[ti.row0 + lr], static_cast<ull>(acc));
  }
}

// ---- visited-trick correction, TE accumulation, advance the wave
__global__ void k_bw_finish(BatchConst c) {
  if (blockIdx.x == 0) {
    // fold the last hop's partials (earlier hops were folded by decide)
    for (uint32_t i = threadIdx.x; i < c.nslots * 64; i += blockDim.x) {
      const SlotDesc &sd = c.slots[i / 64];
      uint64_t *fp = sd.fpart + static_cast<size_t>(c.H) * kParts * 64 + i % 64;
      uint64_t f = 0;
      for (uint32_t p = 0; p < kParts; ++p) {
        f |= fp[p * 64];
        fp[p * 64] = 0;
      }
      sd.front[static_cast<size_t>(c.H) * 64 + i % 64] |= f;
      ull *wp = sd.wpart + static_cast<size_t>(c.H) * kWParts * 64 + i % 64;
      for (uint32_t p = 0; p < kWParts; ++p) wp[p * 64] = 0;
    }
    if (threadIdx.x == 0)
      for (uint32_t s = 0; s < c.nslots; ++s)
        for (uint32_t p = 0; p < kParts; ++p) {
          HopCounters *pp = part_ctr(c.slots[s], c.H, p);
          add_ctr(c.slots[s].ctr[c.H], *pp);
          *pp = HopCounters{};
        }
    __syncthreads();
    if (c.tile_front) {
      const uint32_t n = c.nslots * (c.H + 1) * c.TW;
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t sl = i / ((c.H + 1) * c.TW), rem = i % ((c.H + 1) * c.TW);
        const SlotDesc &sd = c.slots[sl];
        if (sd.tile == UINT32_MAX) continue;
        const uint32_t h = rem / c.TW, j = rem % c.TW;
        c.tile_front[(static_cast<size_t>(sd.tile) * (c.H + 1) + h) * c.TW + j] =
            sd.front[static_cast<size_t>(h) * 64 + j];
      }
    }
  }
  const uint32_t rows_per_tile = c.TW * 64;
  const uint32_t total = c.nslots * rows_per_tile;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += gridDim.x * blockDim.x) {
    const SlotDesc &sd = c.slots[i / rows_per_tile];
    if (sd.tile == UINT32_MAX) continue;
    const TileInfo ti = c.tiles[sd.tile];
    const uint32_t lr = i % rows_per_tile;
    if (lr >= ti.rows) continue;
    const uint32_t row = ti.row0 + lr;
    if (!c.use_r && !c.res_sub && c.sub_src[row]) {
      atomicAdd(&c.out_counts[row], ~0ull);  // minus one
      if (c.out_digests) atomicAdd(&c.out_digests[row], 0ull - digest_hash(c.src[row]));
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // traversed edges, and the compulsory bytes of the expanded hops:
    // 8 nz(F_i) + 8 nz(L_i) + [Shortest] 8 (nz(Vis_{i-1}) + nz(Vis_i))
    // + [Any] 8 (nz(R_{i-1}) + nz(R_i)) + 4 actdeg(F_i) + 16 act(F_i),
    // plus 8 nz of the final result
    ull te = 0, bytes = 0, bytes_sparse = 0;
    for (uint32_t s = 0; s < c.nslots; ++s) {
      const SlotDesc &sd = c.slots[s];
      if (sd.tile == UINT32_MAX) continue;
      const uint32_t hmax = c.tiles[sd.tile].hmax;
      ull vis_nz = sd.ctr[0].visnew, r_nz = sd.ctr[0].rnew;
      for (uint32_t h = 0; h < c.H; ++h) te += sd.ctr[h].te;
      for (uint32_t h = 1; h <= c.H; ++h) {
        const HopCounters &pc = sd.ctr[h - 1], &cc = sd.ctr[h];
        const ull vis_prev = vis_nz, r_prev = r_nz;
        vis_nz += cc.visnew;
        r_nz += cc.rnew;
        if (pc.pop == 0 || h > hmax) continue;
        ull b = 8 * pc.nz + 8 * cc.nz + 4 * pc.actdeg + 16 * pc.act;
        if (sd.Vis) b += 8 * (vis_prev + vis_nz);
        else b += 8 * (r_prev + r_nz);
        bytes += b;
        if (cc.how == 3) bytes_sparse += b;
      }
      bytes += 8 * (c.use_r ? r_nz : vis_nz);
    }
    c.ctl->te += te;
    c.ctl->bytes += bytes;
    c.ctl->bytes_sparse += bytes_sparse;
  }
}

// ---- results: the result of a Shortest row with k_min == 1 is its visited
// set minus its source (vexpand.cpp:306-310 keeps the source for k_min = 0)
__global__ void k_bw_subsrc(BatchConst c) {
  const uint32_t rows_per_tile = c.TW * 64;
  const uint32_t total = c.nslots * rows_per_tile;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += gridDim.x * blockDim.x) {
    const SlotDesc &sd = c.slots[i / rows_per_tile];
    if (sd.tile == UINT32_MAX) continue;
    const TileInfo ti = c.tiles[sd.tile];
    const uint32_t lr = i % rows_per_tile;
    if (lr >= ti.rows || !c.sub_src[ti.row0 + lr]) continue;
    const size_t wi = static_cast<size_t>(c.src[ti.row0 + lr]) * c.TW + lr / 64;
    atomicAnd(reinterpret_cast<ull *>(&gpn(sd.Vis)[wi]), ~(1ull << (lr % 64)));
  }
}

// ---- results of a multi-query group: every result word of the wave's tiles
// goes to its query's matrix (queries own whole words)
template <int TW>
__global__ __launch_bounds__(kBlock) void k_bw_export(BatchConst c) {
  const uint64_t per = static_cast<uint64_t>(c.V) * TW;
  const uint64_t total = static_cast<uint64_t>(c.nslots) * per;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const SlotDesc &sd = c.slots[i / per];
    if (sd.tile == UINT32_MAX) continue;
    const uint64_t k = i % per;  // vertex * TW + j
    const uint32_t v = static_cast<uint32_t>(k / TW), j = static_cast<uint32_t>(k % TW);
    const uint2 d = c.wdest[static_cast<size_t>(sd.tile) * TW + j];
    if (d.x == UINT32_MAX) continue;
    const uint64_t *M = c.use_r ? gpn(sd.R) : gpn(sd.Vis);
    gpn(c.qbase[d.x])[(static_cast<size_t>(d.y / TW) * c.V + v) * TW + d.y % TW] = M[k];
  }
}

__global__ void k_bw_advance(BatchCtl *ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->wave += 1;
}

#define VS_TW_BATCH(tw, FN, ...)          \
  switch (tw) {                           \
    case 1: FN<1>(__VA_ARGS__); break;    \
    case 2: FN<2>(__VA_ARGS__); break;    \
    case 4: FN<4>(__VA_ARGS__); break;    \
    case 8: FN<8>(__VA_ARGS__); break;    \
    case 16: FN<16>(__VA_ARGS__); break;  \
    case 32: FN<32>(__VA_ARGS__); break;  \
    default: FN<64>(__VA_ARGS__); break;  \
  }

// Optional HIP events around each persistent expand launch, per variant
// (timing runs, never inside a captured graph).
struct ExpandEvents {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[2];  // dense, sparse
  ~ExpandEvents() {
    for (auto &v : ev)
      for (auto &e : v) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
  }
  void record(hipStream_t s, bool start, bool sparse) {
    auto &v = ev[sparse ? 1 : 0];
    if (start) {
      hipEvent_t a = nullptr, b = nullptr;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        (void)hipGetLastError();
        return;
      }
      v.emplace_back(a, b);
      (void)hipEventRecord(a, s);
    } else if (!v.empty()) {
      (void)hipEventRecord(v.back().second, s);
    }
  }
  double total_ms(bool sparse) const {
    double t = 0;
    for (auto &e : ev[sparse ? 1 : 0]) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) t += ms;
      else (void)hipGetLastError();
    }
    return t;
  }
};

template <int TW>
static void enqueue_hop_tw(const BatchConst &c, bool shortest, uint32_t hop,
                           uint32_t grid, hipStream_t s, ExpandEvents *ev) {
  hipLaunchKernelGGL((k_bw_wpush<TW>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  if (ev) ev->record(s, true, false);
  if (shortest)
    hipLaunchKernelGGL((k_bw_expand<TW, true, false>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  else
    hipLaunchKernelGGL((k_bw_expand<TW, false, false>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  if (ev) ev->record(s, false, false);
  if (ev) ev->record(s, true, true);
  if (shortest)
    hipLaunchKernelGGL((k_bw_expand<TW, true, true>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  else
    hipLaunchKernelGGL((k_bw_expand<TW, false, true>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  if (ev) ev->record(s, false, true);
  if (shortest) hipLaunchKernelGGL((k_bw_epi<TW, true>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  else hipLaunchKernelGGL((k_bw_epi<TW, false>), dim3(grid), dim3(kBlock), 0, s, c, hop);
  hipLaunchKernelGGL((k_bw_clear<TW>), dim3(grid), dim3(kBlock), 0, s, c, hop - 1);
}

template <int TW>
static void enqueue_final_tw(const BatchConst &c, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_bw_clear<TW>), dim3(grid), dim3(kBlock), 0, s, c, c.H);
  if (c.res_sub)
    hipLaunchKernelGGL(k_bw_subsrc, dim3(std::max<uint32_t>(1, (c.nslots * TW * 64 + 255) / 256)),
                       dim3(256), 0, s, c);
  // half the usual grid: each block accumulates two chunk groups in LDS
  if (c.want_rows)
    hipLaunchKernelGGL((k_bw_rows<TW>), dim3(std::max<uint32_t>(1, grid / 2)), dim3(kBlock), 0,
                       s, c);
  if (c.wdest) hipLaunchKernelGGL((k_bw_export<TW>), dim3(grid), dim3(kBlock), 0, s, c);
}

// Enqueues the whole launch sequence of one wave; returns the number of
// kernel launches.
static uint32_t enqueue_wave(const BatchConst &c, bool shortest, uint32_t grid,
                             hipStream_t s, ExpandEvents *ev = nullptr) {
  uint32_t n = 0;
  hipLaunchKernelGGL(k_bw_assign, dim3(1), dim3(256), 0, s, c);
  {
    const size_t n4 = static_cast<size_t>(c.V) * c.TW / 2;
    const size_t zb = std::min<size_t>((n4 + kZeroChunk - 1) / kZeroChunk, size_t{1} << 22);
    hipLaunchKernelGGL(k_bw_zero, dim3(static_cast<uint32_t>(std::max<size_t>(zb, 1))),
                       dim3(kBlock), 0, s, c);
  }
  hipLaunchKernelGGL(k_bw_l0, dim3(std::max<uint32_t>(1, (c.nslots * c.TW * 64 + 255) / 256)),
                     dim3(256), 0, s, c);
  n += 3;
  const uint32_t nvb = (c.V + 63) / 64;
  const uint32_t item_grid =
      std::min<uint32_t>(c.nslots * ((nvb + kBlock - 1) / kBlock), grid);
  for (uint32_t hop = 1; hop <= c.H; ++hop) {
    hipLaunchKernelGGL(k_bw_decide, dim3(c.nslots), dim3(64), 0, s, c, hop);
    hipLaunchKernelGGL(k_bw_items, dim3(std::max<uint32_t>(1, item_grid)), dim3(kBlock), 0, s, c, hop);
    hipLaunchKernelGGL(k_bw_plan, dim3(1), dim3(64), 0, s, c, hop);
    VS_TW_BATCH(c.TW, enqueue_hop_tw, c, shortest, hop, grid, s, ev);
    n += 8;
  }
  VS_TW_BATCH(c.TW, enqueue_final_tw, c, grid, s);
  hipLaunchKernelGGL(k_bw_finish, dim3(std::max<uint32_t>(1, (c.nslots * c.TW * 64 + 255) / 256)),
                     dim3(256), 0, s, c);
  hipLaunchKernelGGL(k_bw_advance, dim3(1), dim3(64), 0, s, c.ctl);
  n += 3 + (c.want_rows ? 1 : 0) + (c.res_sub ? 1 : 0) + (c.wdest ? 1 : 0);
  return n;
}

// Owning device buffer.  Move-only: a container that reallocates must move
// the handle (leaving the source empty), never copy it and free the original.
struct DevVec {
  void *p = nullptr;
  DevVec() = default;
  DevVec(const DevVec &) = delete;
  DevVec &operator=(const DevVec &) = delete;
  DevVec(DevVec &&o) noexcept : p(o.p) { o.p = nullptr; }
  DevVec &operator=(DevVec &&o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevVec() { if (p) (void)hipFree(p); }
  vs_status alloc(size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      set_error("batch: device allocation of %zu bytes failed", bytes);
      return VS_ERR_NOMEM;
    }
    return VS_OK;
  }
};

// key[r] = sum over u in N(src[r]) of deg(u), one wave per source
__global__ __launch_bounds__(kBlock) void k_src_vol2(const uint64_t *__restrict__ off,
                                                     const uint32_t *__restrict__ adj,
                                                     const uint32_t *__restrict__ src,
                                                     uint32_t S, ull *__restrict__ key) {
  const int lane = lane_id(), wid = wave_id();
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  for (uint32_t r = blockIdx.x * kWavesPerBlock + wid; r < S; r += nwaves) {
    const uint32_t s = src[r];
    uint64_t k = 0;
    for (uint64_t e = off[s] + lane; e < off[s + 1]; e += 64) {
      const uint32_t u = adj[e];
      k += off[u + 1] - off[u];
    }
    k = wave_sum(k);
    if (lane == 0) key[r] = k;
  }
}

vs_status source_volumes(vs_graph_s *g, const DevCSR *push, const uint32_t *src, size_t S,
                         uint64_t *key) {
  if (S == 0) return VS_OK;
  void *ds = nullptr, *dk = nullptr;
  VS_TRY(g->pool.get(S * 4, false, g->stream, &ds));
  vs_status st = g->pool.get(S * 8, false, g->stream, &dk);
  hipError_t e = hipSuccess;
  if (st == VS_OK) e = hipMemcpyAsync(ds, src, S * 4, hipMemcpyHostToDevice, g->stream);
  if (st == VS_OK && e == hipSuccess) {
    hipLaunchKernelGGL(k_src_vol2, dim3(grid_for(S, g->num_cus)), dim3(kBlock), 0, g->stream,
                       push->off, push->adj, static_cast<const uint32_t *>(ds),
                       static_cast<uint32_t>(S), static_cast<ull *>(dk));
    e = hipGetLastError();
  }
  if (st == VS_OK && e == hipSuccess)
    e = hipMemcpyAsync(key, dk, S * 8, hipMemcpyDeviceToHost, g->stream);
  if (st == VS_OK && e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (st == VS_OK && e != hipSuccess) {
    set_error("source volumes: %s", hipGetErrorString(e));
    st = VS_ERR_HIP;
  }
  g->pool.put(ds, S * 4, false);
  if (dk) g->pool.put(dk, S * 8, false);
  return st;
}

// Result matrices of a group's queries (vs_vexpand_batch_results), owned
// until handed out: returned to the pool otherwise.
struct GroupResults {
  vs_graph_s *g = nullptr;
  std::vector<vs_result_s *> res;  // per query of the group, in group order
  std::vector<Mat> mats;           // their matrices while not handed out
  ~GroupResults() {
    for (auto &m : mats)
      if (m.d) g->pool.put(m.d, m.bytes, false);
    for (auto *r : res) delete r;
  }
};

// One group = queries of one path type.  With ro, every query also gets its
// result matrix (rows packed per query; see the packing comment).
static vs_status run_group(vs_graph_s *g, const std::vector<const vs_query *> &qs,
                           const vs_batch_params &p, bool shortest,
                           std::vector<uint64_t> &counts, std::vector<uint64_t> &digests,
                           bool want_digests, vs_batch_stats &st, GroupResults *ro = nullptr,
                           bool want_counts = true) {
  const DevCSR *push = nullptr, *pull = nullptr;
  VS_TRY(graph_csrs(g, p.direction, &push, &pull));
  const uint32_t V = g->V, nvb = (V + 63) / 64;
  uint64_t S64 = 0;
  uint32_t H = 0;
  bool use_r = !shortest;
  for (auto *q : qs) {
    S64 += q->S;
    H = std::max(H, q->k_max);
    if (q->k_min >= 2) use_r = true;
  }
  const bool results = ro != nullptr;
  const uint32_t nq = static_cast<uint32_t>(qs.size());
  if (results) {  // empty queries get empty results even when no row runs
    ro->g = g;
    ro->res.assign(nq, nullptr);
    ro->mats.assign(nq, Mat{});
  }
  // packed row slots: with results every query starts at a word boundary,
  // so a query owns whole words (the rest of its last word is padding)
  uint64_t W64 = 0;
  for (auto *q : qs) W64 += results ? (q->S + 63) / 64 : 0;
  if (!results) W64 = (S64 + 63) / 64;
  if (S64 > 0xFFFFFFF0ull || W64 * 64 > 0xFFFFFFF0ull) {
    set_error("batch: too many rows");
    return VS_ERR_INVALID;
  }
  const uint32_t S = static_cast<uint32_t>(S64);
  const uint32_t W = static_cast<uint32_t>(W64);
  const uint32_t Sp = results ? W * 64 : S;  // >= S
  // 2-hop volumes of the group's sources (group row order): the packing key,
  // and with the learned scheduler the features of the tile-width choice
  std::vector<uint64_t> key;
  const bool learn_tw = p.sched == VS_SCHED_LEARNED && g->sched && p.tile_words == 0 &&
                        g->tuning.tile_words == 0;
  if (S && (learn_tw || !(p.flags & VS_BATCH_KEEP_ORDER))) {
    std::vector<uint32_t> all(S);
    uint32_t r = 0;
    for (auto *q : qs)
      for (uint32_t i = 0; i < q->S; ++i, ++r) {
        if (q->sources[i] >= V) {
          set_error("batch: source %u out of range (V=%u)", q->sources[i], V);
          return VS_ERR_INVALID;
        }
        all[r] = q->sources[i];
      }
    key.resize(S);
    VS_TRY(source_volumes(g, push, all.data(), S, key.data()));
  }
  uint32_t TW = choose_tile_words(p.tile_words ? p.tile_words : g->tuning.tile_words, W);
  if (learn_tw && S) {
    QueryFeatures qf;
    qf.S = S;
    qf.k_max = H;
    qf.V = V;
    qf.E = push->E;
    uint32_t r = 0;
    for (auto *q : qs)
      for (uint32_t i = 0; i < q->S; ++i, ++r) {
        qf.vol2 += key[r];
        qf.degsum += push->hoff[q->sources[i] + 1] - push->hoff[q->sources[i]];
      }
    TW = sched_tile_words(g->sched, qf, choose_tile_words(64, W), TW);
  }
  const uint32_t T = W ? (W + TW - 1) / TW : 0;
  const uint32_t Wpad = T * TW;
  const uint32_t RT = 64 * TW;
  const size_t mat = static_cast<size_t>(V) * TW * 8;
  // one query: its result tiles are the batch's tiles, expanded in place
  const bool zero_copy = results && nq == 1;
  std::vector<std::vector<uint32_t>> stored;  // per query: caller row -> stored row
  std::vector<uint32_t> qword0;               // per query: its first packed word
  if (results) {
    stored.resize(nq);
    qword0.assign(nq, 0);
    for (uint32_t qi = 0; qi < nq; ++qi) {
      const vs_query *q = qs[qi];
      const uint32_t tq = (q->S + RT - 1) / RT;
      ro->res[qi] = new vs_result_s;
      vs_result_s *r = ro->res[qi];
      r->g = g;
      r->S = q->S;
      r->V = V;
      r->tw = TW;
      r->ntiles = tq;
      r->k_actual_max = 0;
      if (q->S) r->row_map.assign(q->sources, q->sources + q->S);
      if (q->S == 0 || zero_copy) continue;
      Mat &m = ro->mats[qi];
      m.bytes = static_cast<size_t>(tq) * mat;
      // words past the query's own rows stay zero (no batch word feeds them)
      VS_TRY(g->pool.get(m.bytes, true, g->stream, reinterpret_cast<void **>(&m.d)));
    }
    if (zero_copy && S) {
      Mat &m = ro->mats[0];
      m.bytes = static_cast<size_t>(T) * mat;
      VS_TRY(g->pool.get(m.bytes, false, g->stream, reinterpret_cast<void **>(&m.d)));
    }
  }
  if (S64 == 0) return VS_OK;
  // host tables
  std::vector<uint32_t> src(S);
  std::vector<uint8_t> sub(S, 0);
  std::vector<uint32_t> kmin(S), kmax(S);
  {
    uint32_t r = 0;
    for (auto *q : qs) {
      for (uint32_t i = 0; i < q->S; ++i, ++r) {
        if (q->sources[i] >= V) {
          set_error("batch: source %u out of range (V=%u)", q->sources[i], V);
          return VS_ERR_INVALID;
        }
        src[r] = q->sources[i];
        kmin[r] = q->k_min;
        kmax[r] = q->k_max;
        sub[r] = (!use_r && q->k_min == 1) ? 1 : 0;
      }
    }
  }
  // Row packing: rows are independent, so they may share words in any order.
  // Rows are grouped by k_max first (a tile runs as many hops as its deepest
  // row), then ordered by the source's 2-hop volume sum_{u in N(s)} deg(u)
  // (descending; isolated sources last), which puts rows that finish at
  // similar hops into the same words: words of isolated sources are dead
  // after L0 (their lanes never load again), and words of central sources
  // become fully visited early (their lanes stop early).  The keys come from
  // one small kernel over the sources' lists.  pos[r] is the packed position
  // of group row r.
  //
  // With results, the packing is per query instead: queries in order of
  // k_max (descending), each starting at a word boundary, its rows sorted
  // by 2-hop volume within each tile of its own (64 * TW rows of its caller
  // order), the plain engine's packing.  A query's words then hold only its
  // rows, so its result is a word-for-word copy -- or, for a one-query
  // group, the batch's own tiles -- and the result's stored-row map is the
  // within-tile order.
  std::vector<uint32_t> pos(S);
  {
    if (p.flags & VS_BATCH_KEEP_ORDER) key.clear();  // computed for the tile width only
    if (!results) {
      std::vector<uint32_t> order(S);
      for (uint32_t r = 0; r < S; ++r) order[r] = r;
      if (!key.empty())
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
          if (kmax[x] != kmax[y]) return kmax[x] > kmax[y];
          return key[x] > key[y];
        });
      for (uint32_t i = 0; i < S; ++i) pos[order[i]] = i;
    } else {
      std::vector<uint32_t> qorder(nq), qrow0(nq, 0);
      for (uint32_t qi = 0; qi < nq; ++qi) {
        qorder[qi] = qi;
        if (qi) qrow0[qi] = qrow0[qi - 1] + qs[qi - 1]->S;
      }
      std::stable_sort(qorder.begin(), qorder.end(),
                       [&](uint32_t x, uint32_t y) { return qs[x]->k_max > qs[y]->k_max; });
      uint32_t w0 = 0;
      for (uint32_t qi : qorder) {
        qword0[qi] = w0;
        const uint32_t n = qs[qi]->S, r0 = qrow0[qi];
        w0 += (n + 63) / 64;
        std::vector<uint32_t> ord(n);
        for (uint32_t i = 0; i < n; ++i) ord[i] = i;
        if (!key.empty())
          for (uint32_t t0 = 0; t0 < n; t0 += RT)
            std::stable_sort(ord.begin() + t0, ord.begin() + std::min(n, t0 + RT),
                             [&](uint32_t x, uint32_t y) { return key[r0 + x] > key[r0 + y]; });
        stored[qi].resize(n);
        bool ident = true;
        for (uint32_t k = 0; k < n; ++k) {
          stored[qi][ord[k]] = k;
          ident &= ord[k] == k;
          pos[r0 + ord[k]] = qword0[qi] * 64 + k;
        }
        if (ident) stored[qi].clear();
      }
    }
    std::vector<uint32_t> s2(Sp, kNoRow), kn(Sp, 0), kx(Sp, 0);
    std::vector<uint8_t> sb(Sp, 0);
    for (uint32_t r = 0; r < S; ++r) {
      const uint32_t i = pos[r];
      s2[i] = src[r];
      kn[i] = kmin[r];
      kx[i] = kmax[r];
      sb[i] = sub[r];
    }
    src.swap(s2);
    kmin.swap(kn);
    kmax.swap(kx);
    sub.swap(sb);
  }
  std::vector<uint64_t> live(static_cast<size_t>(H + 2) * Wpad, 0);
  std::vector<uint64_t> racc;
  if (use_r) racc.assign(static_cast<size_t>(H + 1) * Wpad, 0);
  for (uint32_t r = 0; r < Sp; ++r) {
    if (src[r] == kNoRow) continue;
    const uint32_t w = r / 64;
    const uint64_t bit = 1ull << (r % 64);
    for (uint32_t i = 0; i <= std::min(kmax[r], H + 1); ++i)
      live[static_cast<size_t>(i) * Wpad + w] |= bit;
    if (use_r)
      for (uint32_t i = kmin[r]; i <= kmax[r]; ++i)
        racc[static_cast<size_t>(i) * Wpad + w] |= bit;
  }
  std::vector<TileInfo> tiles(T);
  for (uint32_t t = 0; t < T; ++t) {
    tiles[t].row0 = t * TW * 64;
    tiles[t].rows = std::min<uint32_t>(Sp - tiles[t].row0, TW * 64);
    tiles[t].word0 = t * TW;
    uint32_t h = 0;
    for (uint32_t r = tiles[t].row0; r < tiles[t].row0 + tiles[t].rows; ++r)
      h = std::max(h, kmax[r]);
    tiles[t].hmax = h;
  }
  // slots that fit the memory budget
  const size_t items_cap = static_cast<size_t>(V) + push->E / kPushChunk + 64;
  const size_t per_slot = mat * (shortest ? 4 : 3) + (use_r ? mat : 0) + 3 * nvb * 8 +
                          2 * static_cast<size_t>(V) * 8 +
                          items_cap * 8 + (H + 1) * sizeof(HopCounters) +
                          (H + 1) * kParts * (kPartStride + 64 * 8) +
                          (H + 1) * (kWParts + 1) * 64 * 8;
  size_t freeb = 0, totalb = 0;
  VS_HIP(hipMemGetInfo(&freeb, &totalb));
  freeb += g->pool.cached();  // the slot buffers come from the graph's pool
  const size_t budget = p.mem_budget ? p.mem_budget : freeb / 10 * 7;
  uint32_t nslots = static_cast<uint32_t>(std::min<size_t>(budget / per_slot, T));
  nslots = std::min<uint32_t>(nslots, kMaxSlots);
  if (nslots == 0 && per_slot <= freeb) nslots = 1;  // the budget is a soft cap
  if (nslots == 0) {
    set_error("batch: one tile needs %zu bytes, budget %zu", per_slot, budget);
    return VS_ERR_NOMEM;
  }
  const uint32_t nwaves = (T + nslots - 1) / nslots;
  hipStream_t s = g->stream;
  // device arrays
  DevVec d_src, d_sub, d_live, d_racc, d_tiles, d_slots, d_ctl, d_counts, d_dig;
  DevVec d_wdest, d_qbase, d_front;
  VS_TRY(d_src.alloc(Sp * 4ull));
  VS_TRY(d_sub.alloc(Sp));
  VS_TRY(d_live.alloc(live.size() * 8));
  if (use_r) VS_TRY(d_racc.alloc(racc.size() * 8));
  VS_TRY(d_tiles.alloc(T * sizeof(TileInfo)));
  VS_TRY(d_slots.alloc(nslots * sizeof(SlotDesc)));
  VS_TRY(d_ctl.alloc(sizeof(BatchCtl)));
  VS_TRY(d_counts.alloc(Sp * 8ull));
  if (want_digests) VS_TRY(d_dig.alloc(Sp * 8ull));
  const size_t front_words = static_cast<size_t>(T) * (H + 1) * TW;
  if (results) {
    VS_TRY(d_front.alloc(front_words * 8));
    if (!zero_copy) {
      std::vector<uint2> wdest(Wpad, make_uint2(UINT32_MAX, 0));
      std::vector<uint64_t *> qbase(nq, nullptr);
      for (uint32_t qi = 0; qi < nq; ++qi) {
        qbase[qi] = ro->mats[qi].d;
        for (uint32_t w = 0; w < (qs[qi]->S + 63) / 64; ++w)
          wdest[qword0[qi] + w] = make_uint2(qi, w);
      }
      VS_TRY(d_wdest.alloc(Wpad * sizeof(uint2)));
      VS_TRY(d_qbase.alloc(nq * sizeof(uint64_t *)));
      VS_HIP(hipMemcpyAsync(d_wdest.p, wdest.data(), Wpad * sizeof(uint2), hipMemcpyHostToDevice, s));
      VS_HIP(hipMemcpyAsync(d_qbase.p, qbase.data(), nq * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
      VS_HIP(hipStreamSynchronize(s));  // wdest / qbase leave scope
    }
  }
  VS_HIP(hipMemcpyAsync(d_src.p, src.data(), Sp * 4ull, hipMemcpyHostToDevice, s));
  VS_HIP(hipMemcpyAsync(d_sub.p, sub.data(), Sp, hipMemcpyHostToDevice, s));
  VS_HIP(hipMemcpyAsync(d_live.p, live.data(), live.size() * 8, hipMemcpyHostToDevice, s));
  if (use_r) VS_HIP(hipMemcpyAsync(d_racc.p, racc.data(), racc.size() * 8, hipMemcpyHostToDevice, s));
  VS_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), T * sizeof(TileInfo), hipMemcpyHostToDevice, s));
  // Slot buffers come from the graph's pool (zeroed once per call, kept
  // clean by the kernels during it), so repeated calls -- and the plain
  // engine, whose matrices have the same size -- reuse the same device
  // memory; they go back to the pool when the call ends.
  struct SlotBufs {
    Pool &pool;
    std::vector<std::pair<void *, size_t>> v;
    ~SlotBufs() {
      for (auto &x : v) pool.put(x.first, x.second, false);
    }
  } bufs{g->pool, {}};
  std::vector<SlotDesc> sd(nslots);
  auto get = [&](size_t bytes, bool zero, void **out) -> vs_status {
    VS_TRY(g->pool.get(bytes, zero, s, out));
    bufs.v.emplace_back(*out, bytes);
    return VS_OK;
  };
  for (uint32_t i = 0; i < nslots; ++i) {
    void *q;
    VS_TRY(get(mat, true, &q)); sd[i].F[0] = static_cast<uint64_t *>(q);
    VS_TRY(get(mat, true, &q)); sd[i].F[1] = static_cast<uint64_t *>(q);
    VS_TRY(get(V * 8ull, true, &q)); sd[i].Fm[0] = static_cast<uint64_t *>(q);
    VS_TRY(get(V * 8ull, true, &q)); sd[i].Fm[1] = static_cast<uint64_t *>(q);
    VS_TRY(get(nvb * 8ull, true, &q)); sd[i].bm[0] = static_cast<uint64_t *>(q);
    VS_TRY(get(nvb * 8ull, true, &q)); sd[i].bm[1] = static_cast<uint64_t *>(q);
    VS_TRY(get(mat, true, &q)); sd[i].N = static_cast<uint64_t *>(q);
    VS_TRY(get(nvb * 8ull, true, &q)); sd[i].touched = static_cast<uint64_t *>(q);
    VS_TRY(get(items_cap * 8, false, &q)); sd[i].items = static_cast<ull *>(q);
    sd[i].Vis = nullptr;
    sd[i].R = nullptr;
    // with a one-query group's results the result matrix (R, or Vis when no
    // R is kept) is the query's own tile, set per wave by k_bw_assign
    if (shortest && !(zero_copy && !use_r)) {
      VS_TRY(get(mat, false, &q));
      sd[i].Vis = static_cast<uint64_t *>(q);
    }
    if (use_r && !zero_copy) {
      VS_TRY(get(mat, false, &q));
      sd[i].R = static_cast<uint64_t *>(q);
    }
    VS_TRY(get((H + 1) * 64 * 8ull, true, &q));
    sd[i].front = static_cast<uint64_t *>(q);
    VS_TRY(get((H + 1) * sizeof(HopCounters), true, &q));
    sd[i].ctr = static_cast<HopCounters *>(q);
    VS_TRY(get((H + 1) * kParts * kPartStride, true, &q));
    sd[i].part = static_cast<char *>(q);
    VS_TRY(get((H + 1) * kParts * 64 * 8ull, true, &q));
    sd[i].fpart = static_cast<uint64_t *>(q);
    VS_TRY(get((H + 1) * 64 * 8ull, true, &q));
    sd[i].wdeg = static_cast<ull *>(q);
    VS_TRY(get((H + 1) * kWParts * 64 * 8ull, true, &q));
    sd[i].wpart = static_cast<ull *>(q);
    sd[i].pw = 0;
    sd[i].tile = UINT32_MAX;
    sd[i].mode = 0;
  }
  VS_HIP(hipMemcpyAsync(d_slots.p, sd.data(), nslots * sizeof(SlotDesc), hipMemcpyHostToDevice, s));
  BatchCtl ctl0{};
  VS_HIP(hipMemcpyAsync(d_ctl.p, &ctl0, sizeof(ctl0), hipMemcpyHostToDevice, s));
  VS_HIP(hipMemsetAsync(d_counts.p, 0, Sp * 8ull, s));
  if (want_digests) VS_HIP(hipMemsetAsync(d_dig.p, 0, Sp * 8ull, s));

  BatchConst c{};
  c.V = V;
  c.TW = TW;
  c.H = H;
  c.nslots = nslots;
  c.ntiles = T;
  c.Wpad = Wpad;
  c.S = Sp;
  c.sched = p.sched;
  c.alpha = g->tuning.push_alpha ? g->tuning.push_alpha : 16;
  c.sparse_pct = g->tuning.sparse_pct ? g->tuning.sparse_pct : kSparsePct;
  c.use_r = use_r ? 1 : 0;
  c.E = push->E;
  c.wpush_budget = word_push_budget(g->tuning, push->E);
  c.poff = pull->off;
  c.padj = pull->adj;
  c.xoff = push->off;
  c.xadj = push->adj;
  c.src = static_cast<const uint32_t *>(d_src.p);
  c.sub_src = static_cast<const uint8_t *>(d_sub.p);
  c.live = static_cast<const uint64_t *>(d_live.p);
  c.racc = use_r ? static_cast<const uint64_t *>(d_racc.p) : nullptr;
  c.slots = static_cast<SlotDesc *>(d_slots.p);
  c.tiles = static_cast<const TileInfo *>(d_tiles.p);
  c.ctl = static_cast<BatchCtl *>(d_ctl.p);
  c.out_counts = static_cast<ull *>(d_counts.p);
  c.want_rows = (want_counts || want_digests) ? 1u : 0u;
  c.out_digests = want_digests ? static_cast<ull *>(d_dig.p) : nullptr;
  c.mlp = (p.sched == VS_SCHED_LEARNED && g->sched) ? sched_params(g->sched) : nullptr;
  VS_TRY(ensure_pull_items(*const_cast<DevCSR *>(pull), V, g->tuning.pull_hubs));
  c.pitems = pull->pitems;
  c.padj = pull->hadj;  // hubs-first lists (pull only; push uses xadj)
  c.hubcnt = pull->hubcnt;
  c.npitems = pull->npitems;
  c.has_heavy = pull->has_heavy ? 1 : 0;
  c.nhwords = pull->nhwords;
  c.hwords = pull->hwords;
  if (results) {
    c.res_base = zero_copy ? ro->mats[0].d : nullptr;
    c.wdest = zero_copy ? nullptr : static_cast<const uint2 *>(d_wdest.p);
    c.qbase = zero_copy ? nullptr : static_cast<uint64_t *const *>(d_qbase.p);
    c.tile_front = static_cast<uint64_t *>(d_front.p);
    c.res_sub = use_r ? 0u : 1u;
  }
  const uint32_t grid = static_cast<uint32_t>(g->num_cus) * 8;

  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  struct EvG { hipEvent_t a, b; ~EvG() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } evg{e0, e1};
  VS_HIP(hipStreamSynchronize(s));
  uint32_t launches = 0;
  double ms = 0.0;
  const uint32_t replays = std::max<uint32_t>(1, p.replays);
  const bool timing = (p.flags & VS_VX_TIMING) != 0;  // per-launch events: no capture
  ExpandEvents xev;
  double replay_ms = 0.0, replay_wall = 0.0;
  uint32_t replay_n = 0;
  if (p.use_graph && !timing) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const auto c0 = std::chrono::steady_clock::now();
    VS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launches = enqueue_wave(c, shortest, grid, s);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) {
      set_error("batch: graph capture failed: %s", hipGetErrorString(ce));
      return VS_ERR_HIP;
    }
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ce != hipSuccess) {
      (void)hipGraphDestroy(graph);
      set_error("batch: graph instantiate failed: %s", hipGetErrorString(ce));
      return VS_ERR_HIP;
    }
    st.ms_capture += std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - c0).count();
    hipError_t re = hipSuccess;
    for (uint32_t rep = 0; rep < replays && re == hipSuccess; ++rep) {
      const auto w0 = std::chrono::steady_clock::now();
      if (rep > 0) {  // timing replays: restart from wave 0 with clean outputs
        re = hipMemcpyAsync(d_ctl.p, &ctl0, sizeof(ctl0), hipMemcpyHostToDevice, s);
        if (re == hipSuccess) re = hipMemsetAsync(d_counts.p, 0, Sp * 8ull, s);
        if (re == hipSuccess && want_digests) re = hipMemsetAsync(d_dig.p, 0, Sp * 8ull, s);
      }
      if (re == hipSuccess) re = hipEventRecord(e0, s);
      for (uint32_t w = 0; w < nwaves && re == hipSuccess; ++w) re = hipGraphLaunch(exec, s);
      if (re == hipSuccess) re = hipEventRecord(e1, s);
      if (re == hipSuccess) re = hipStreamSynchronize(s);
      float rms = 0.f;
      if (re == hipSuccess) re = hipEventElapsedTime(&rms, e0, e1);
      // the first replay warms caches and the instantiated graph: it counts
      // only when it is the only one
      if (re == hipSuccess && (rep > 0 || replays == 1)) {
        replay_ms += rms;
        replay_wall += std::chrono::duration<double, std::milli>(
                           std::chrono::steady_clock::now() - w0).count();
        replay_n += 1;
      }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (re != hipSuccess) {
      set_error("batch: graph replay failed: %s", hipGetErrorString(re));
      return VS_ERR_HIP;
    }
  } else {
    for (uint32_t rep = 0; rep < replays; ++rep) {
      if (rep > 0) {
        VS_HIP(hipMemcpyAsync(d_ctl.p, &ctl0, sizeof(ctl0), hipMemcpyHostToDevice, s));
        VS_HIP(hipMemsetAsync(d_counts.p, 0, Sp * 8ull, s));
        if (want_digests) VS_HIP(hipMemsetAsync(d_dig.p, 0, Sp * 8ull, s));
      }
      const bool last = rep + 1 == replays;
      VS_HIP(hipEventRecord(e0, s));
      for (uint32_t w = 0; w < nwaves; ++w)
        launches = enqueue_wave(c, shortest, grid, s, (timing && last) ? &xev : nullptr);
      VS_HIP(hipGetLastError());
      VS_HIP(hipEventRecord(e1, s));
      VS_HIP(hipStreamSynchronize(s));
    }
    if (timing) {
      st.ms_expand += xev.total_ms(false);
      st.n_expand += xev.ev[0].size();
      st.ms_expand_sparse += xev.total_ms(true);
      st.n_expand_sparse += xev.ev[1].size();
    }
  }
  float fms = 0.f;
  VS_HIP(hipEventElapsedTime(&fms, e0, e1));
  ms = replay_n ? replay_ms / replay_n : fms;
  st.ms_wall += replay_n ? replay_wall / replay_n : 0.0;
  BatchCtl ctl{};
  VS_HIP(hipMemcpy(&ctl, d_ctl.p, sizeof(ctl), hipMemcpyDeviceToHost));
  std::vector<uint64_t> hc(Sp), hd;
  VS_HIP(hipMemcpy(hc.data(), d_counts.p, Sp * 8ull, hipMemcpyDeviceToHost));
  if (want_digests) {
    hd.resize(Sp);
    VS_HIP(hipMemcpy(hd.data(), d_dig.p, Sp * 8ull, hipMemcpyDeviceToHost));
  }
  if (results) {
    // k_actual_max per query as VExpand::after_child sets it
    // (vexpand.cpp:222-233): Shortest stops at the first hop that adds no
    // vertex to any of its rows (that hop counts), Any runs to k_max
    std::vector<uint64_t> front(front_words);
    VS_HIP(hipMemcpy(front.data(), d_front.p, front_words * 8, hipMemcpyDeviceToHost));
    for (uint32_t qi = 0; qi < nq; ++qi) {
      const vs_query *q = qs[qi];
      vs_result_s *r = ro->res[qi];
      if (q->S == 0) continue;
      uint32_t ka = q->k_max;
      if (shortest)
        for (uint32_t h = 1; h <= q->k_max; ++h) {
          uint64_t any = 0;
          for (uint32_t w = 0; w < (q->S + 63) / 64; ++w) {
            const uint32_t gw = qword0[qi] + w;
            const uint32_t n = std::min<uint32_t>(64, q->S - w * 64);
            const uint64_t mask = n == 64 ? ~0ull : ((1ull << n) - 1);
            any |= front[(static_cast<size_t>(gw / TW) * (H + 1) + h) * TW + gw % TW] & mask;
          }
          if (!any) {
            ka = h;
            break;
          }
        }
      r->k_actual_max = ka;
      r->stored.swap(stored[qi]);
    }
  }
  // back from packed positions to group row order
  for (uint32_t r = 0; r < S; ++r) counts.push_back(hc[pos[r]]);
  if (want_digests)
    for (uint32_t r = 0; r < S; ++r) digests.push_back(hd[pos[r]]);
  st.te += ctl.te;
  st.ms_total += ms;
  st.num_tiles += T;
  st.num_waves += nwaves;
  st.launches += launches;
  st.push_decisions += ctl.push_dec;  // of the last replay (counters reset per replay)
  st.pull_decisions += ctl.pull_dec;
  st.bytes_alg += ctl.bytes;
  st.bytes_sparse += ctl.bytes_sparse;
  st.wpush_words += ctl.wpush_words;
  return VS_OK;
}

} // namespace vs

using namespace vs;

extern "C" {

static vs_status batch_impl(const char *fn, vs_graph_t g, const vs_query *queries, uint32_t nq,
                            const vs_batch_params *p, uint64_t *row_counts,
                            uint64_t *row_digests, vs_result_t *results,
                            vs_batch_stats *stats) {
  if (!g || !p || (nq > 0 && !queries)) {
    set_error("%s: null argument", fn);
    return VS_ERR_INVALID;
  }
  if (p->tile_words > 64 || (p->tile_words & (p->tile_words - 1))) {
    set_error("%s: tile_words %u is not a power of two <= 64", fn, p->tile_words);
    return VS_ERR_INVALID;
  }
  if (p->sched < VS_SCHED_HEURISTIC || p->sched > VS_SCHED_LEARNED) {
    set_error("%s: unknown sched %d", fn, p->sched);
    return VS_ERR_INVALID;
  }
  for (uint32_t q = 0; q < nq; ++q) {
    const vs_query &Q = queries[q];
    if (Q.path_type != VS_PATH_ANY && Q.path_type != VS_PATH_SHORTEST) {
      set_error("%s: query %u: path_type %d not batchable", fn, q, Q.path_type);
      return VS_ERR_INVALID;
    }
    if (Q.k_max >= VS_MAX_HOPS || (Q.S > 0 && !Q.sources)) {
      set_error("%s: query %u: bad k_max or null sources", fn, q);
      return VS_ERR_INVALID;
    }
  }
  VS_HIP(hipSetDevice(g->device));
  vs_batch_stats st{};
  std::vector<uint64_t> out_c, out_d;
  std::vector<uint64_t> row_off(nq + 1, 0);
  for (uint32_t q = 0; q < nq; ++q) row_off[q + 1] = row_off[q] + queries[q].S;
  const bool want_d = row_digests != nullptr;
  // results are handed out only when every group has succeeded
  std::vector<std::unique_ptr<GroupResults>> held;
  std::vector<std::vector<uint32_t>> held_ids;
  for (int path : {VS_PATH_SHORTEST, VS_PATH_ANY}) {
    std::vector<const vs_query *> qs;
    std::vector<uint32_t> ids;
    for (uint32_t q = 0; q < nq; ++q)
      if (queries[q].path_type == path) { qs.push_back(&queries[q]); ids.push_back(q); }
    if (qs.empty()) continue;
    out_c.clear();
    out_d.clear();
    GroupResults *ro = nullptr;
    if (results) {
      held.emplace_back(new GroupResults);
      held_ids.push_back(ids);
      ro = held.back().get();
    }
    VS_TRY(run_group(g, qs, *p, path == VS_PATH_SHORTEST, out_c, out_d, want_d, st, ro,
                     row_counts != nullptr));
    uint64_t r = 0;
    for (uint32_t id : ids) {
      for (uint32_t i = 0; i < queries[id].S; ++i, ++r) {
        if (row_counts) row_counts[row_off[id] + i] = out_c[r];
        if (want_d) row_digests[row_off[id] + i] = out_d[r];
      }
    }
  }
  for (size_t k = 0; k < held.size(); ++k) {
    GroupResults &gr = *held[k];
    for (size_t qi = 0; qi < held_ids[k].size(); ++qi) {
      vs_result_s *res = gr.res[qi];
      res->mats.push_back(gr.mats[qi]);
      gr.mats[qi] = Mat{};
      gr.res[qi] = nullptr;
      g->refs++;
      results[held_ids[k][qi]] = res;
    }
  }
  if (stats) *stats = st;
  return VS_OK;
}

vs_status vs_vexpand_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                           const vs_batch_params *p, uint64_t *row_counts,
                           uint64_t *row_digests, vs_batch_stats *stats) try {
  clear_error();
  return batch_impl("vs_vexpand_batch", g, queries, nq, p, row_counts, row_digests, nullptr,
                    stats);
} VS_CATCH("vs_vexpand_batch")

vs_status vs_vexpand_batch_results(vs_graph_t g, const vs_query *queries, uint32_t nq,
                                   const vs_batch_params *p, vs_result_t *results,
                                   vs_batch_stats *stats) try {
  clear_error();
  if (nq > 0 && !results) {
    set_error("vs_vexpand_batch_results: null results");
    return VS_ERR_INVALID;
  }
  for (uint32_t q = 0; q < nq; ++q) results[q] = nullptr;
  return batch_impl("vs_vexpand_batch_results", g, queries, nq, p, nullptr, nullptr, results,
                    stats);
} VS_CATCH("vs_vexpand_batch_results")

// Autotuner of the batched engine.  Coordinate descent over the graph's
// tuning: tile width, then the push threshold alpha, the word-mask threshold
// sparse_pct and the hub count pull_hubs, each knob over a few candidates
// with the others fixed; every candidate is timed by the batched engine on
// the caller's queries (device time: the mean of `reps` replays of the
// captured launch sequence after one warm replay).  The starting tuning is
// measured first and a candidate replaces the current choice only when it
// is at least 2% faster, so the result is never slower than the start by the
// tuner's own clock.  No choice changes a result bit.
constexpr double kTuneGain = 0.02;  // a candidate replaces the current choice only 2% faster

static vs_status autotune_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                                const vs_batch_params &p0, vs_tuning *best, double *best_ms,
                                double *start_ms) {
  uint64_t rows = 0;
  for (uint32_t q = 0; q < nq; ++q) rows += queries[q].S;
  if (rows == 0) {
    set_error("vs_autotune_batch: no source rows");
    return VS_ERR_INVALID;
  }
  vs_batch_params p = p0;
  p.use_graph = 1;
  p.replays = 3;  // one warm replay, the mean of the other two
  p.flags &= ~VS_VX_TIMING;
  const vs_tuning saved = g->tuning;
  vs_tuning cur = saved;
  const uint32_t tw_max = choose_tile_words(64, static_cast<uint32_t>((rows + 63) / 64));
  cur.tile_words = choose_tile_words(p.tile_words ? p.tile_words : cur.tile_words,
                                     static_cast<uint32_t>((rows + 63) / 64));
  p.tile_words = 0;  // the graph's tuning decides
  if (!cur.push_alpha) cur.push_alpha = 16;
  if (!cur.sparse_pct) cur.sparse_pct = kSparsePct;
  if (!cur.pull_hubs) cur.pull_hubs = kPullHubs;
  if (!cur.wpush_pml) cur.wpush_pml = kWPushPml;
  auto measure = [&](const vs_tuning &t, double *ms) -> vs_status {
    g->tuning = t;
    vs_batch_stats st{};
    const vs_status s = batch_impl("vs_autotune_batch", g, queries, nq, &p, nullptr, nullptr,
                                   nullptr, &st);
    *ms = st.ms_total;
    return s;
  };
  double bms = 0.0;
  vs_status s = measure(cur, &bms);
  if (s != VS_OK) { g->tuning = saved; return s; }
  if (start_ms) *start_ms = bms;
  auto try_knob = [&](uint32_t vs_tuning::*knob, std::vector<uint32_t> cands) -> vs_status {
    const uint32_t keep = cur.*knob;
    for (uint32_t v : cands) {
      if (v == keep) continue;
      vs_tuning t = cur;
      t.*knob = v;
      double ms = 0.0;
      VS_TRY(measure(t, &ms));
      // a candidate must be clearly faster: differences inside the
      // replay-to-replay spread would make the choice follow the noise
      if (ms < bms * (1.0 - kTuneGain)) {
        bms = ms;
        cur = t;
      }
    }
    return VS_OK;
  };
  std::vector<uint32_t> tws;
  for (uint32_t tw = tw_max; tw >= 1 && tws.size() < 3; tw /= 2) tws.push_back(tw);
  const uint32_t V = g->V;
  std::vector<uint32_t> hubs;
  for (uint32_t h : {32768u, 131072u, 524288u}) hubs.push_back(std::min(h, std::max(V, 1u)));
  s = try_knob(&vs_tuning::tile_words, tws);
  if (s == VS_OK) s = try_knob(&vs_tuning::push_alpha, {4u, 16u, 64u});
  if (s == VS_OK) s = try_knob(&vs_tuning::sparse_pct, {35u, 50u, 65u});
  if (s == VS_OK) s = try_knob(&vs_tuning::pull_hubs, hubs);
  // word push budget: default, larger (two or three sparse words), off
  if (s == VS_OK) s = try_knob(&vs_tuning::wpush_pml, {kWPushPml, 200u, 2000u});
  if (s != VS_OK) {
    g->tuning = saved;
    return s;
  }
  g->tuning = cur;
  // the last candidates rebuilt the hubs-first pull lists for their hub
  // counts: rebuild them for the kept one now, so the caller's next call
  // does not pay for it
  const DevCSR *push = nullptr, *pull = nullptr;
  VS_TRY(graph_csrs(g, p.direction, &push, &pull));
  VS_TRY(ensure_pull_items(*const_cast<DevCSR *>(pull), V, cur.pull_hubs));
  if (best) *best = cur;
  if (best_ms) *best_ms = bms;
  return VS_OK;
}

vs_status vs_autotune_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                            const vs_batch_params *p, vs_tuning *best, double *best_ms,
                            double *start_ms) try {
  clear_error();
  if (!g || !p || !queries || nq == 0) {
    set_error("vs_autotune_batch: need a graph, parameters and queries");
    return VS_ERR_INVALID;
  }
  return autotune_batch(g, queries, nq, *p, best, best_ms, start_ms);
} VS_CATCH("vs_autotune_batch")

// One query (hops 1..k_max) through the batched engine's tuner.
vs_status vs_autotune_vexpand(vs_graph_t g, const uint32_t *sources, uint32_t S,
                              uint32_t k_max, int32_t path_type, vs_tuning *best,
                              double *best_ms) try {
  clear_error();
  if (!g || !sources || S == 0 || k_max == 0) {
    set_error("vs_autotune_vexpand: need a graph, sources and k_max >= 1");
    return VS_ERR_INVALID;
  }
  vs_query q{sources, S, 1, k_max, path_type};
  vs_batch_params p{};
  p.sched = VS_SCHED_HEURISTIC;
  p.direction = VS_DIR_FORWARD;
  return autotune_batch(g, &q, 1, p, best, best_ms, nullptr);
} VS_CATCH("vs_autotune_vexpand")

} // extern "C"
