// Hop kernels of the bit-matrix multi-source expansion, shared by the
// single-query engine (vs_vexpand.hip) and the batched engine (vs_batch.hip).
//
// One hop of the reference is the Boolean-semiring SpMM
//   out[si][v] |= base[si][u]  for every edge (u, v) and stack si
// (Expand<DenseVarStackMatrix>::expand_dense, src/objects/operator/
// expand.cpp:263-311), followed for Shortest by
//   o &= ~vis;  vis |= o;  visit_count = popcount(vis)
// (DenseVarStackMatrix::dedup_and_update_visit_and_count,
// src/objects/storage/bitmap.cpp:1054-1096), and the layers are unioned over
// [k_min, k_max] (VExpand::merge_to_MT, src/objects/operator/vexpand.cpp:298-354).
//
// Here a hop is either
//   pull: each destination vertex ORs the frontier words of its active
//         in-neighbours (gather; no atomics), with the dedup / visit / union
//         epilogue fused, and an early exit once every still-needed row bit
//         is found; or
//   push: each active frontier vertex scatters its words into a zeroed
//         accumulator N with 64-bit atomicOr, then an epilogue walks the
//         touched-vertex bitmap and applies the same dedup / visit / union.
// Both produce identical bits.  Frontier buffers and their active bitmaps are
// all-zero outside their active rows, and are cleared through the bitmap
// after use, so no hop pays for a full-matrix memset.
//
// Word masks: with Fm / Fnm set, every frontier buffer carries one 64-bit
// mask per vertex whose bit j says that word j of the vertex's row is
// nonzero (written with the row, cleared with it).  A pull lane reads the
// mask of each compacted neighbour and loads only the words that hold bits:
// sparse frontier rows (hop 2 gathers rows of the sources' neighbours, a few
// words each) then cost a few cache lines instead of TW words.
//
// Word push: on a dense pull hop, the frontier words that are cheapest to
// scatter (the fewest out-edges of rows holding bits there, summed over the
// words from the cheapest up to a budget) are pushed instead of pulled.  A
// pull lane stops gathering once its word holds every row it still needs;
// words whose rows mostly lie farther away never get there and keep the
// wave reading every in-edge, so taking those few sparse words out of the
// pull lets the dense ones finish it early.  The scatter ORs them into N
// before the pull; the pull never gathers them: the first item of each
// destination takes them from N (a light vertex finishes them with its
// pulled words, a chunked one leaves them to its epilogue).
//
// Row masks: `live` bits are the rows that still expand at this hop (rows
// < S, and in a batch the rows whose query has k_max >= hop); `racc` bits are
// the rows whose result takes this hop's layer (k_min <= hop); `live_next`
// are the rows that will expand at the next hop (traversed-edge count).
#pragma once

#include "vs_device.hh"

namespace vs {

// Per (hop, tile) counters, filled by the kernels that write layer L_i.
struct HopCounters {
  ull pop;     // popcount of L_i
  ull nz;      // nonzero words of L_i
  ull act;     // vertices with a nonzero L_i row
  ull actdeg;  // sum of push-degree over those vertices
  ull te;      // sum over rows live at hop i+1 of push-degree * bits: TE of hop i+1
  ull visnew;  // visited words that went from zero to nonzero
  ull rnew;    // result words that went from zero to nonzero
  ull items;   // push work items built from L_i (written by the item builder)
  ull nzdeg;   // sum over active vertices of push-degree * nonzero words: the
               // words a full next-hop pull would gather, by frequency
  ull how;     // batched engine: how L_i was made (1 push, 2 pull, 3 pull
               // through word masks)
};

struct HopArgs {
  uint32_t V;
  uint32_t S;      // rows of the tile set (valid-bit mask when live == nullptr)
  uint32_t word0;  // global word index of this tile's word 0
  uint32_t dense;  // frontier covers most edges: pull skips the bitmap lookup
  const uint64_t *poff;  // pull CSR: edges arriving at v
  const uint32_t *padj;
  const uint64_t *xoff;  // push CSR: edges leaving u
  const uint32_t *xadj;
  const uint64_t *F;     // frontier L_{i-1}, V*TW words
  const uint64_t *actbm; // its active-vertex bitmap, ceil(V/64) words
  const uint64_t *Fm;    // its word masks (V) when pull should read them (sparse
                         // rows), or nullptr: gather every word
  uint64_t *Fn;          // L_i (all zero on entry)
  uint64_t *Fnm;         // word masks of L_i (all zero on entry), or nullptr
  uint64_t *actbm_n;     // its bitmap (all zero on entry); Fn, Fnm and actbm_n are
                         // nullptr on a hop whose frontier nothing reads (the
                         // last): the new bits still reach Vis / R and the counters
  uint64_t *N;           // push accumulator (all zero on entry and exit)
  uint64_t *touched;     // push touched bitmap (all zero on entry and exit)
  uint64_t *Vis;         // visited (Shortest) or nullptr (Any)
  uint64_t *R;           // result accumulator or nullptr
  const ull *items;      // push items built from actbm
  const ull *nitems;
  const ull *pitems;     // pull items (static per CSR)
  uint64_t npitems;
  const uint32_t *ewords; // bitmap words the epilogue visits, or nullptr: all
  uint32_t newords;
  const uint32_t *hubcnt; // hub neighbours at the head of each pull list, or nullptr
  HopCounters *ctr;      // counters of L_i
  const uint64_t *live;      // per-word rows expanding now, or nullptr (rows < S)
  const uint64_t *live_next; // per-word rows expanding next hop, or nullptr
  const uint64_t *racc;      // per-word rows accumulating into R, or nullptr (all)
  // Rows whose frontier F_{i-1} is non-empty anywhere (OR of all its rows,
  // per word), or nullptr.  A row outside it cannot gain a bit this hop, so
  // pull stops asking for it -- without this, rows of isolated sources would
  // keep every destination's early exit from firing.
  const uint64_t *front;
  uint64_t *front_next;      // the same OR for L_i, accumulated here
  // Per-word push degree of L_i: [kWParts][64] partial sums (a block adds
  // into partial blockIdx % kWParts), word j summing the push-degree of
  // every vertex whose row holds bits in word j; or nullptr.
  ull *wdeg;
  // Word push of this hop: the pushed words of the tile (0: none; N set
  // otherwise) and the frontier's word masks (read by the scatter).
  uint64_t pw;
  const uint64_t *Fwm;
};

// Grid-stride loops over bitmap words (epilogues, clears): the unit that
// wave w takes in round k (units [k * nwaves, (k + 1) * nwaves)), each round
// rotated by k * kSkew waves.  R-MAT puts its densest words at indices that
// agree in their low bits (multiples of large powers of two); a plain stride
// hands all of them to one wave, which then runs ~50x the mean.
constexpr uint64_t kSkew = 97;
__device__ __forceinline__ uint64_t skewed_unit(uint64_t k, uint64_t w, uint64_t nwaves) {
  return k * nwaves + (w + nwaves - (k * kSkew) % nwaves) % nwaves;
}

// Partial copies of the per-word push-degree sums.
constexpr uint32_t kWParts = 16;

// Word-push budget (push-degree sum of the pushed words) of a graph tuning.
inline uint64_t word_push_budget(const vs_tuning &t, uint64_t E) {
  const uint64_t pml = t.wpush_pml ? t.wpush_pml : kWPushPml;
  return pml > 1000 ? 0 : E / 1000 * pml + E % 1000 * pml / 1000;
}

// Word push choice of one tile from the push degrees of its frontier words:
// the words with degree > 0, cheapest first (ties by word index), while the
// degrees taken sum to at most budget.
inline uint64_t word_push_pick(const uint64_t *wdeg, uint32_t TW, uint64_t budget) {
  uint64_t pw = 0;
  for (uint32_t j = 0; j < TW; ++j) {
    const uint64_t w = wdeg[j];
    if (!w) continue;
    uint64_t cum = 0;
    for (uint32_t k = 0; k < TW; ++k)
      if (wdeg[k] && (wdeg[k] < w || (wdeg[k] == w && k <= j))) cum += wdeg[k];
    if (cum <= budget) pw |= 1ull << j;
  }
  return pw;
}

// Per-lane partial counters, summed over the wave at the end.
struct LocalCtr {
  uint64_t pop = 0, nz = 0, act = 0, actdeg = 0, te = 0, visnew = 0, rnew = 0;
  uint64_t wdeg = 0;   // push-degree of the rows whose word j (this lane's) became
                       // nonzero: summed over the lanes it is nzdeg
  uint64_t rowor = 0;  // OR of the words this lane emitted (its word index j)
};

__device__ __forceinline__ void flush_ctr(LocalCtr &c, HopCounters *ctr,
                                          uint64_t *front_next, ull *wdeg, int j) {
  if (front_next && c.rowor)
    atomicOr(reinterpret_cast<ull *>(&front_next[j]), static_cast<ull>(c.rowor));
  if (wdeg && c.wdeg)
    atomicAdd(&wdeg[(blockIdx.x % kWParts) * 64 + j], static_cast<ull>(c.wdeg));
  const uint64_t pop = wave_sum(c.pop), nz = wave_sum(c.nz), act = wave_sum(c.act),
                 actdeg = wave_sum(c.actdeg), te = wave_sum(c.te),
                 visnew = wave_sum(c.visnew), rnew = wave_sum(c.rnew),
                 nzdeg = wave_sum(c.wdeg);
  if (lane_id() == 0) {
    if (pop) atomicAdd(&ctr->pop, static_cast<ull>(pop));
    if (nz) atomicAdd(&ctr->nz, static_cast<ull>(nz));
    if (act) atomicAdd(&ctr->act, static_cast<ull>(act));
    if (actdeg) atomicAdd(&ctr->actdeg, static_cast<ull>(actdeg));
    if (te) atomicAdd(&ctr->te, static_cast<ull>(te));
    if (visnew) atomicAdd(&ctr->visnew, static_cast<ull>(visnew));
    if (rnew) atomicAdd(&ctr->rnew, static_cast<ull>(rnew));
    if (nzdeg) atomicAdd(&ctr->nzdeg, static_cast<ull>(nzdeg));
  }
  c = LocalCtr{};
}

// Lane masks for word j of the tile.
struct LaneMasks {
  uint64_t live, live_next, racc, front;
};

__device__ __forceinline__ LaneMasks lane_masks(const HopArgs &a, int j) {
  LaneMasks m;
  const uint64_t v = valid_mask(a.word0 + j, a.S);
  m.live = a.live ? a.live[j] : v;
  m.live_next = a.live_next ? a.live_next[j] : v;
  m.racc = a.racc ? a.racc[j] : ~0ull;
  m.front = a.front ? a.front[j] : ~0ull;
  return m;
}

// Writes one new word nb (nonzero) of L_i at index wi and updates Vis / R.
template <bool SHORTEST>
__device__ __forceinline__ void emit_word(const HopArgs &a, size_t wi, uint64_t nb,
                                          uint64_t vis, uint64_t racc, LocalCtr &c) {
  if (a.Fn) a.Fn[wi] = nb;
  if (SHORTEST) {
    a.Vis[wi] = vis | nb;
    c.visnew += (vis == 0);
  }
  if (a.R) {
    const uint64_t add = nb & racc;
    if (add) {
      const uint64_t r = a.R[wi];
      if ((r | add) != r) a.R[wi] = r | add;
      c.rnew += (r == 0);
    }
  }
  c.pop += __popcll(nb);
  c.nz += 1;
  c.rowor |= nb;
}

// ------------------------------------------------------------------- pull
// One work item = (destination vertex v, chunk of kPullChunk in-edges).
// Items are strided over the waves, so the hubs that R-MAT packs into the
// low vertex ids spread over the whole GPU.  Lanes = 64/TW groups x TW words;
// the groups split the chunk.  Neighbour ids are read 64 at a time, filtered
// by the frontier's active bitmap, compacted into LDS with ballot + mbcnt, and
// gathered with TW-word coalesced loads (four in flight per lane).  A lane
// stops loading once its word holds every needed row, the wave once all do.
// A vertex with one item finishes in place (dedup / visit / union); a hub's
// items OR their partial words into the push accumulator N and mark the
// touched bitmap, and the push epilogue finishes it.
// The loads that only depend on the item: issued one item ahead (software
// pipelining), so their latency overlaps the previous item's gathers.
struct PullHead {
  uint32_t v, ch;
  uint64_t vis, lb, le;
  uint32_t hubs;  // hub neighbours at the head of v's pull list
};

template <int TW, bool SHORTEST>
__device__ __forceinline__ PullHead pull_head(const HopArgs &a, ull item, int j) {
  PullHead h;
  h.v = static_cast<uint32_t>(item);
  h.ch = static_cast<uint32_t>(item >> 32);
  h.vis = SHORTEST ? a.Vis[static_cast<size_t>(h.v) * TW + j] : 0;
  h.lb = a.poff[h.v];
  h.le = a.poff[h.v + 1];
  h.hubs = a.hubcnt ? a.hubcnt[h.v] : 0xffffffffu;
  return h;
}

// The first 64 list ids of a pull item whose first group will be scanned:
// loaded one item ahead so that the scan does not start with a dependent
// load.  Returns whether they were loaded (wave-uniform); the item's body
// then takes them instead of loading its first group itself.  Later chunks
// of a chunked vertex are not prefetched (their needed rows depend on N).
template <int TW, bool SPARSE>
__device__ __forceinline__ bool pull_prefetch(const HopArgs &a, const PullHead &h,
                                              const LaneMasks &mk, int j, uint32_t &u0) {
  const bool pwl = !SPARSE && ((a.pw >> j) & 1);
  const bool want = h.ch == 0 && !pwl && (mk.live & mk.front & ~h.vis) != 0;
  if (__ballot(want) == 0) return false;
  const uint64_t idx = h.lb + lane_id();
  if (idx < h.le) u0 = a.padj[idx];
  return true;
}

// The bitmap halves of the ids pull_prefetch loaded (ahead of the body,
// like the ids; a dense frontier does not read the bitmap).
__device__ __forceinline__ uint32_t pull_prefetch_bm(const HopArgs &a, const PullHead &h,
                                                     uint32_t u0) {
  const uint32_t *bm = reinterpret_cast<const uint32_t *>(a.actbm);
  return h.lb + lane_id() < h.le ? bm[u0 >> 5] : 0u;
}

// Frontier word of neighbour u: hub rows stay cached, other rows stream.
template <int TW, bool CACHED>
__device__ __forceinline__ uint64_t gather(const HopArgs &a, uint32_t u, int j) {
  const uint64_t *p = a.F + static_cast<size_t>(u) * TW + j;
  if constexpr (CACHED) return *p;
  else return __builtin_nontemporal_load(p);
}

// ORs the frontier words of nbr[k], nbr[k + G], ... below end into acc, eight
// loads in flight, and stops as soon as this lane's partial word already
// holds every needed row (a lane-local test: the group OR can only add bits,
// so stopping on the partial word is exact for TW = 64 and conservative
// below).  Lanes stop independently; the hubs at the head of the list are
// gathered first, so most lanes that can stop do so before the streaming part.
template <int TW, bool CACHED>
__device__ __forceinline__ void gather_range(const HopArgs &a, const uint32_t *nbr, uint32_t &k,
                                             uint32_t end, int j, uint64_t need,
                                             uint64_t &acc) {
  constexpr int G = 64 / TW;
  for (; k + 7 * G < end; k += 8 * G) {
    if ((acc & need) == need) return;
    uint64_t x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = gather<TW, CACHED>(a, nbr[k + q * G], j);
    acc |= ((x[0] | x[1]) | (x[2] | x[3])) | ((x[4] | x[5]) | (x[6] | x[7]));
  }
  if (k + 3 * G < end && (acc & need) != need) {
    const uint64_t x0 = gather<TW, CACHED>(a, nbr[k], j);
    const uint64_t x1 = gather<TW, CACHED>(a, nbr[k + G], j);
    const uint64_t x2 = gather<TW, CACHED>(a, nbr[k + 2 * G], j);
    const uint64_t x3 = gather<TW, CACHED>(a, nbr[k + 3 * G], j);
    acc |= (x0 | x1) | (x2 | x3);
    k += 4 * G;
  }
  for (; k < end; k += G) {
    if ((acc & need) == need) return;
    acc |= gather<TW, CACHED>(a, nbr[k], j);
  }
}

// Sparse rows (word masks on): the (neighbour, word) pairs whose word holds
// bits and is still needed are laid out flat over the wave's lanes, so every
// load instruction carries up to 64 useful words instead of one per lane
// per neighbour; the words are ORed into the wave's accumulator accw[TW] in
// LDS.  Each lane first writes all of its neighbour's pairs (16-bit codes
// neighbour << 6 | word) into the wave's LDS list, which holds every pair of
// a 64-neighbour group (64 x 64); the gathers then read the list in rounds
// of kPairs (8 loads in flight per lane).
constexpr uint32_t kPairs = 512;
constexpr uint32_t kPairCap = 64 * 64;

template <int TW>
__device__ __forceinline__ void gather_pairs(const HopArgs &a, const uint32_t *nbr,
                                             const uint64_t *nm, uint32_t n, uint32_t nh,
                                             uint64_t needw, uint16_t *pairs, uint64_t *accw) {
  const int lane = lane_id();
  const uint64_t mine = lane < static_cast<int>(n) ? (nm[lane] & needw) : 0;
  const uint32_t cnt = __popcll(mine);
  uint32_t inc = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d);
    if (lane >= d) inc += y;
  }
  const uint32_t P = __shfl(inc, 63);
  {
    // every lane writes its pairs at once: as many steps as the longest list
    uint64_t m = mine;
    uint32_t p = inc - cnt;
    while (m) {
      const uint32_t jj = __ffsll(static_cast<long long>(m)) - 1;
      m &= m - 1;
      pairs[p++] = static_cast<uint16_t>((lane << 6) | jj);
    }
  }
  wave_sync();
  for (uint32_t base = 0; base < P; base += kPairs) {
    const uint32_t np = min(kPairs, P - base);
    uint64_t x[kPairs / 64];
#pragma unroll
    for (int r = 0; r < static_cast<int>(kPairs / 64); ++r) {
      const uint32_t q = r * 64 + lane;
      x[r] = 0;
      if (q < np) {
        const uint32_t pr = pairs[base + q];
        const uint32_t k = pr >> 6;
        const uint64_t *src = a.F + static_cast<size_t>(nbr[k]) * TW + (pr & 63);
        x[r] = k < nh ? *src : __builtin_nontemporal_load(src);
      }
    }
    // the word index is read back from LDS rather than held in registers
#pragma unroll
    for (int r = 0; r < static_cast<int>(kPairs / 64); ++r)
      if (x[r])
        atomicOr(reinterpret_cast<ull *>(&accw[pairs[base + r * 64 + lane] & 63]),
                 static_cast<ull>(x[r]));
    wave_sync();
  }
}

// Dense rows, few (neighbour, word) pairs still needed: the pairs (neighbour k, needed word)
// of all n active neighbours are laid flat over the lanes without reading
// word masks (neighbour-major, so a neighbour's words are neighbouring
// lanes), at most kCross of them: one round of loads instead of one load
// instruction per neighbour with only the needed words' lanes active.  The
// same words are read as by the per-neighbour gather.  wl is the wave's LDS
// list of the needed words.
constexpr uint32_t kCross = 256;

template <int TW>
__device__ __forceinline__ void gather_cross(const HopArgs &a, const uint32_t *nbr, uint32_t n,
                                             uint32_t nh, uint64_t needw, uint8_t *wl,
                                             uint64_t *accw) {
  const int lane = lane_id();
  if (lane < TW && ((needw >> lane) & 1))
    wl[__popcll(needw & ((1ull << lane) - 1))] = static_cast<uint8_t>(lane);
  wave_sync();
  const uint32_t nw = __popcll(needw);
  const uint32_t P = n * nw;           // <= kCross (checked by the caller)
  const uint32_t inv = (65536u + nw - 1) / nw;  // q / nw == (q * inv) >> 16 for q < 512
  uint64_t x[kCross / 64];
#pragma unroll
  for (int r = 0; r < static_cast<int>(kCross / 64); ++r) {
    const uint32_t q = r * 64 + lane;
    x[r] = 0;
    if (q < P) {
      const uint32_t k = (q * inv) >> 16;
      const uint64_t *src = a.F + static_cast<size_t>(nbr[k]) * TW + wl[q - k * nw];
      x[r] = k < nh ? *src : __builtin_nontemporal_load(src);
    }
  }
  // the word index is recomputed rather than held in registers
#pragma unroll
  for (int r = 0; r < static_cast<int>(kCross / 64); ++r) {
    const uint32_t q = r * 64 + lane;
    if (x[r]) {
      const uint32_t k = (q * inv) >> 16;
      atomicOr(reinterpret_cast<ull *>(&accw[wl[q - k * nw]]), static_cast<ull>(x[r]));
    }
  }
  wave_sync();
}

// SPARSE (the frontier's word masks are read, a.Fm set): pair gathers;
// otherwise every lane gathers its word of every active neighbour, or, when
// the (neighbour, needed word) cross product fits kCross pairs, that cross
// product is gathered flat over the lanes.
// GPF: a later group's ids are loaded while the group before it gathers.
template <int TW, bool SHORTEST, bool SPARSE, bool GPF = false>
__device__ __forceinline__ void pull_body(const HopArgs &a, const PullHead &h, uint32_t *nbr,
                                          uint64_t *nm, uint16_t *pairs, uint64_t *accw,
                                          uint8_t *wl, const LaneMasks &mk, LocalCtr &c,
                                          uint32_t u0 = 0, bool pre = false,
                                          uint32_t bw0 = 0, bool prebm = false) {
  constexpr int G = 64 / TW;
  const int lane = lane_id();
  const int g = lane / TW, j = lane % TW;
  const uint32_t v = h.v;
  const size_t wi = static_cast<size_t>(v) * TW + j;
  const uint64_t vis = h.vis;
  const uint64_t lb = h.lb, le = h.le;
  const bool heavy = le - lb > kPullChunk;
  // Pushed words (word push) are never gathered: the scatter left their
  // bits in N, and the vertex's first item takes them -- a light vertex
  // finishes them with its pulled words, a chunked one leaves them in N for
  // its epilogue.
  const bool pwl = !SPARSE && ((a.pw >> j) & 1);
  uint64_t acc = (pwl && h.ch == 0) ? a.N[wi] : 0;
  uint64_t need = mk.live & mk.front & ~vis;
  // A later chunk of a chunked vertex does not look again for what the
  // vertex's other items already put into N: the items list every first
  // chunk before any later one, so by now the first chunk (the hub
  // neighbours, the densest rows) has usually found most of it.
  if (heavy && h.ch > 0) need &= ~a.N[wi];
  bool todo = need != 0 && !pwl;
  const uint64_t b = lb + static_cast<uint64_t>(h.ch) * kPullChunk;
  const uint64_t e = __ballot(todo) ? min(b + kPullChunk, le) : b;
  const bool dense = a.dense != 0;
  const uint64_t hub_end = h.hubs == 0xffffffffu ? le : lb + h.hubs;
  if (lane < TW) accw[lane] = 0;
  uint32_t un = 0;  // GPF: the next group's ids
  for (uint64_t base = b; base < e; base += 64) {
    const uint64_t idx = base + lane;
    uint32_t u = 0;
    bool act = false;
    uint32_t unext = 0;
    if (GPF && idx + 64 < e) unext = a.padj[idx + 64];
    if (idx < e) {
      // the first group's ids may have been loaded while the previous item
      // ran (and with GPF a later group's during the group before it)
      const bool first = base == b;
      u = first ? (pre ? u0 : a.padj[idx]) : (GPF ? un : a.padj[idx]);
      // a dense frontier skips the bitmap lookup: inactive rows are zero;
      // the first group's bitmap words may also have been loaded ahead
      act = dense || (prebm && first ? ((bw0 >> (u & 31)) & 1)
                                     : ((a.actbm[u >> 6] >> (u & 63)) & 1));
    }
    un = unext;
    const uint64_t m = __ballot(act);
    const uint32_t n = __popcll(m);
    // hub neighbours precede the others in the list, so after compaction the
    // first nh slots are hubs
    const uint32_t nh = __popcll(__ballot(act && idx < hub_end));
    if (act) {
      const uint32_t r = lane_rank(m);
      nbr[r] = u;
      if (SPARSE) nm[r] = a.Fm[u];
    }
    wave_sync();
    if (SPARSE) {
      // words (group 0's lanes) that still miss rows
      const uint64_t needw = __ballot(g == 0 && todo);
      if (needw) gather_pairs<TW>(a, nbr, nm, n, nh, needw, pairs, accw);
      acc = accw[j];
    } else {
      const uint64_t needw = __ballot(g == 0 && todo);
      const uint32_t nw = __popcll(needw);
      if (n * nw <= kCross) {
        if (needw) {
          gather_cross<TW>(a, nbr, n, nh, needw, wl, accw);
          acc |= accw[j];
        }
      } else if (todo) {
        uint32_t k = g;
        gather_range<TW, true>(a, nbr, k, nh, j, need, acc);
        gather_range<TW, false>(a, nbr, k, n, j, need, acc);
      }
    }
    wave_sync();
    acc = group_or<TW>(acc);
    todo = !pwl && (acc & need) != need;
    if (__ballot(todo) == 0) break;  // nothing left to find
  }
  acc = group_or<TW>(acc);
  const uint64_t nb = acc & need;
  // bits 0..TW-1: the words of v's new row that hold bits (group 0's lanes)
  const uint64_t wmask = __ballot(g == 0 && nb != 0);
  const bool row_new = wmask != 0;
  if (heavy) {
    // the epilogue finishes v when it is touched: new bits, or pushed bits
    // in N (which it also clears)
    const bool take = __ballot(g == 0 && (nb != 0 || (pwl && acc != 0))) != 0;
    if (g == 0 && nb && !pwl)
      atomicOr(reinterpret_cast<ull *>(&a.N[wi]), static_cast<ull>(nb));
    if (take && lane == 0)
      atomicOr(reinterpret_cast<ull *>(&a.touched[v >> 6]), 1ull << (v & 63));
    return;
  }
  if (g == 0 && pwl && acc) a.N[wi] = 0;  // the pushed bits are taken
  if (g == 0 && nb) emit_word<SHORTEST>(a, wi, nb, vis, mk.racc, c);
  if (row_new) {
    const uint64_t deg = a.xoff == a.poff ? le - lb : a.xoff[v + 1] - a.xoff[v];
    if (lane == 0) {
      c.act += 1;
      c.actdeg += deg;
      if (a.actbm_n) atomicOr(reinterpret_cast<ull *>(&a.actbm_n[v >> 6]), 1ull << (v & 63));
      if (a.Fnm) a.Fnm[v] = wmask;
    }
    if (g == 0) {
      c.te += deg * __popcll(nb & mk.live_next);
      if (nb) c.wdeg += deg;
    }
  }
}

// ------------------------------------------------------------------- push
// One work item = (vertex, chunk of kPushChunk out-neighbours).
template <int TW>
__device__ __forceinline__ void push_item(const HopArgs &a, ull item) {
  constexpr int G = 64 / TW;
  const int lane = lane_id();
  const int g = lane / TW, j = lane % TW;
  const uint32_t u = static_cast<uint32_t>(item);
  const uint32_t ch = static_cast<uint32_t>(item >> 32);
  const uint64_t ue = a.xoff[u + 1];
  const uint64_t b = a.xoff[u] + static_cast<uint64_t>(ch) * kPushChunk;
  const uint64_t e = min(b + kPushChunk, ue);
  const uint64_t f = a.F[static_cast<size_t>(u) * TW + j];
  for (uint64_t base = b; base < e; base += 64) {
    const uint64_t idx = base + lane;
    const uint32_t vv = idx < e ? a.xadj[idx] : 0u;
    const uint32_t n = static_cast<uint32_t>(min<uint64_t>(64, e - base));
    for (uint32_t k0 = 0; k0 < n; k0 += G) {
      const uint32_t k = k0 + g;
      const uint32_t v = __shfl(vv, k < n ? static_cast<int>(k) : 0);
      if (k < n) {
        if (f) atomicOr(reinterpret_cast<ull *>(&a.N[static_cast<size_t>(v) * TW + j]),
                        static_cast<ull>(f));
        if (j == 0)
          atomicOr(reinterpret_cast<ull *>(&a.touched[v >> 6]), 1ull << (v & 63));
      }
    }
  }
}

// Word push, one item = (vertex u, chunk of kPushChunk out-neighbours), one
// lane per neighbour: ORs u's pushed words that hold bits into N.  Items are
// built only for vertices holding bits in a pushed word.
template <int TW>
__device__ __forceinline__ void wpush_item(const HopArgs &a, ull item) {
  const int lane = lane_id();
  const uint32_t u = static_cast<uint32_t>(item);
  const uint32_t ch = static_cast<uint32_t>(item >> 32);
  const uint64_t ue = a.xoff[u + 1];
  const uint64_t b = a.xoff[u] + static_cast<uint64_t>(ch) * kPushChunk;
  const uint64_t e = min(b + kPushChunk, ue);
  const uint64_t wm = a.Fwm[u] & a.pw;
  const uint64_t *fu = a.F + static_cast<size_t>(u) * TW;
  for (uint64_t base = b; base < e; base += 64) {
    const uint64_t idx = base + lane;
    if (idx < e) {
      const uint32_t v = a.xadj[idx];
      for (uint64_t m = wm; m; m &= m - 1) {
        const int jj = __ffsll(static_cast<long long>(m)) - 1;
        const uint64_t f = fu[jj];
        if (f) atomicOr(reinterpret_cast<ull *>(&a.N[static_cast<size_t>(v) * TW + jj]),
                        static_cast<ull>(f));
      }
    }
  }
}

// Push epilogue for one touched-bitmap word: consume N (re-zeroing it), apply
// dedup / visit / union, write L_i and its active bitmap word.
template <int TW, bool SHORTEST>
__device__ __forceinline__ void epi_block(const HopArgs &a, uint32_t vb,
                                          const LaneMasks &mk, LocalCtr &c) {
  constexpr int G = 64 / TW;
  constexpr int U = 4;  // vertices per group of lanes whose loads are in flight together
  const int lane = lane_id();
  const int g = lane / TW, j = lane % TW;
  uint64_t m = a.touched[vb];
  if (!m) return;
  // degrees of the block's touched vertices, one coalesced read (lane = vertex)
  const uint32_t vl = vb * 64 + lane;
  const uint64_t degl = ((m >> lane) & 1) ? a.xoff[vl + 1] - a.xoff[vl] : 0;
  uint64_t actmask = 0;
  while (m) {
    int iv[U];
    uint64_t nv[U], visv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      iv[u] = nth_set_bit(m, g + u * G);
      nv[u] = 0;
      visv[u] = 0;
      if (iv[u] >= 0) {
        const size_t wi = (static_cast<size_t>(vb) * 64 + iv[u]) * TW + j;
        nv[u] = a.N[wi];
        if (SHORTEST) visv[u] = a.Vis[wi];
      }
    }
    m = drop_low_bits(m, U * G);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = iv[u];
      const uint32_t v = vb * 64 + (i >= 0 ? i : 0);
      uint64_t nb = 0;
      if (i >= 0) {
        const size_t wi = static_cast<size_t>(v) * TW + j;
        const uint64_t n = nv[u];
        if (n) a.N[wi] = 0;
        const uint64_t vis = visv[u];
        nb = n & mk.live & ~vis;
        if (nb) emit_word<SHORTEST>(a, wi, nb, vis, mk.racc, c);
      }
      const uint64_t gb = group_bits<TW>(__ballot(nb != 0), g);
      const uint64_t deg = shfl64(degl, i >= 0 ? i : 0);
      if (i >= 0 && gb) {
        if (j == 0) {
          actmask |= 1ull << i;
          c.act += 1;
          c.actdeg += deg;
          if (a.Fnm) a.Fnm[v] = gb;
        }
        c.te += deg * __popcll(nb & mk.live_next);
        if (nb) c.wdeg += deg;
      }
    }
  }
  actmask = wave_or(actmask);
  if (lane == 0) {
    // pull items of light vertices set bits of this word too
    if (actmask && a.actbm_n)
      atomicOr(reinterpret_cast<ull *>(&a.actbm_n[vb]), static_cast<ull>(actmask));
    a.touched[vb] = 0;
  }
}

// Zero the active rows of a frontier buffer, their word masks (Fm may be
// nullptr) and its bitmap word vb.
template <int TW>
__device__ __forceinline__ void clear_block(uint64_t *F, uint64_t *Fm, uint64_t *actbm,
                                            uint32_t vb) {
  constexpr int G = 64 / TW;
  const int lane = lane_id();
  const int g = lane / TW, j = lane % TW;
  uint64_t m = actbm[vb];
  if (!m) return;
  while (m) {
    const int i = nth_set_bit(m, g);
    m = drop_low_bits(m, G);
    if (i >= 0) {
      F[(static_cast<size_t>(vb) * 64 + i) * TW + j] = 0;
      if (Fm && j == 0) Fm[static_cast<size_t>(vb) * 64 + i] = 0;
    }
  }
  if (lane == 0) actbm[vb] = 0;
}

// As clear_block, but writes only the words each row's word mask names.
// Every writer of a frontier row writes its mask too (emits in the pull and
// the epilogue, the L0 initialiser), so the mask covers all nonzero words,
// and a row that holds bits in a few words costs a few sector writes.
template <int TW>
__device__ __forceinline__ void clear_block_masked(uint64_t *F, uint64_t *Fm, uint64_t *actbm,
                                                   uint32_t vb) {
  constexpr int G = 64 / TW;
  const int lane = lane_id();
  const int g = lane / TW, j = lane % TW;
  uint64_t m = actbm[vb];
  if (!m) return;
  const size_t v0 = static_cast<size_t>(vb) * 64;
  // lane i reads (and zeroes) the mask of vertex v0 + i if it is active
  const uint64_t mine = ((m >> lane) & 1) ? Fm[v0 + lane] : 0;
  if (mine) Fm[v0 + lane] = 0;
  while (m) {
    const int i = nth_set_bit(m, g);
    m = drop_low_bits(m, G);
    const uint64_t wm = shfl64(mine, i >= 0 ? i : 0);
    if (i >= 0 && ((wm >> j) & 1)) F[(v0 + i) * TW + j] = 0;
  }
  if (lane == 0) actbm[vb] = 0;
}

// Item builder for one 256-word slice of an active bitmap: per-thread counts
// of kPushChunk-neighbour chunks, block-level exclusive scan, one atomicAdd
// per block.  Must be called by the whole block (uses __syncthreads).

// With fm set, only the vertices whose word mask meets pw (word push).
__device__ __forceinline__ void items_slice(const uint64_t *actbm, uint32_t nvb,
                                            uint32_t base, const uint64_t *xoff,
                                            ull *items, ull *nitems, ull *s_wsum,
                                            ull *s_base, const uint64_t *fm = nullptr,
                                            uint64_t pw = 0) {
  const int lane = lane_id(), wid = wave_id();
  const uint32_t wb = base + threadIdx.x;
  uint64_t m = wb < nvb ? actbm[wb] : 0;
  if (fm) {
    // eight mask loads in flight per round (a dense bitmap word would
    // otherwise cost 64 dependent round trips)
    uint64_t keep = 0;
    for (uint64_t mm = m; mm;) {
      int ii[8];
      uint64_t wm[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        ii[t] = mm ? __ffsll(static_cast<long long>(mm)) - 1 : -1;
        mm &= mm - 1;
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) wm[t] = ii[t] >= 0 ? fm[wb * 64 + ii[t]] : 0;
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (wm[t] & pw) keep |= 1ull << ii[t];
    }
    m = keep;
  }
  ull cnt = 0;
  for (uint64_t mm = m; mm; mm &= mm - 1) {
    const uint32_t v = wb * 64 + (__ffsll(static_cast<long long>(mm)) - 1);
    const uint64_t d = xoff[v + 1] - xoff[v];
    cnt += (d + kPushChunk - 1) / kPushChunk;
  }
  ull x = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const ull y = shfl_up64(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    ull tot = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
      const ull t = s_wsum[w];
      s_wsum[w] = tot;
      tot += t;
    }
    *s_base = tot ? atomicAdd(nitems, tot) : 0;
  }
  __syncthreads();
  ull pos = *s_base + s_wsum[wid] + x - cnt;
  for (uint64_t mm = m; mm; mm &= mm - 1) {
    const uint32_t v = wb * 64 + (__ffsll(static_cast<long long>(mm)) - 1);
    const uint64_t d = xoff[v + 1] - xoff[v];
    const uint64_t nch = (d + kPushChunk - 1) / kPushChunk;
    for (uint64_t q = 0; q < nch; ++q) items[pos++] = v | (q << 32);
  }
  __syncthreads();
}

// --------------------------------------------------------- plain kernels
template <int TW, bool SHORTEST, bool SPARSE>
__global__ __launch_bounds__(kBlock) void k_pull(HopArgs a) {
  __shared__ uint32_t s_nbr[kWavesPerBlock][64];
  __shared__ uint64_t s_nm[kWavesPerBlock][SPARSE ? 64 : 1];
  __shared__ uint16_t s_pairs[kWavesPerBlock][SPARSE ? kPairCap : 1];
  __shared__ uint64_t s_acc[kWavesPerBlock][64];
  __shared__ uint8_t s_wl[kWavesPerBlock][64];
  const int wid = wave_id();
  const LaneMasks mk = lane_masks(a, lane_id() % TW);
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  LocalCtr c;
  const int j = lane_id() % TW;
  const uint64_t n = a.npitems;
  uint64_t it = blockIdx.x * kWavesPerBlock + wid;
  PullHead h{};
  if (it < n) h = pull_head<TW, SHORTEST>(a, a.pitems[it], j);
  while (it < n) {
    const uint64_t nx = it + nwaves;
    PullHead hn{};
    if (nx < n) hn = pull_head<TW, SHORTEST>(a, a.pitems[nx], j);
    pull_body<TW, SHORTEST, SPARSE>(a, h, s_nbr[wid], s_nm[wid], s_pairs[wid], s_acc[wid],
                                    s_wl[wid], mk, c);
    h = hn;
    it = nx;
  }
  flush_ctr(c, a.ctr, a.front_next, a.wdeg, j);
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_push(HopArgs a) {
  const int wid = wave_id();
  const uint64_t nitems = *a.nitems;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves)
    push_item<TW>(a, a.items[it]);
}

template <int TW, bool SHORTEST>
__global__ __launch_bounds__(kBlock) void k_push_epi(HopArgs a) {
  const int wid = wave_id();
  const LaneMasks mk = lane_masks(a, lane_id() % TW);
  const uint32_t n = a.ewords ? a.newords : (a.V + 63) / 64;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  const uint32_t w = blockIdx.x * kWavesPerBlock + wid;
  LocalCtr c;
  for (uint64_t k = 0; k * nwaves < n; ++k) {
    const uint64_t i = skewed_unit(k, w, nwaves);
    if (i < n) epi_block<TW, SHORTEST>(a, a.ewords ? a.ewords[i] : static_cast<uint32_t>(i), mk, c);
  }
  flush_ctr(c, a.ctr, a.front_next, a.wdeg, lane_id() % TW);
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_wpush(HopArgs a) {
  const int wid = wave_id();
  const uint64_t nitems = *a.nitems;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves)
    wpush_item<TW>(a, a.items[it]);
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_clear(uint64_t *F, uint64_t *Fm, uint64_t *actbm,
                                                 uint32_t V) {
  const int wid = wave_id();
  const uint32_t nvb = (V + 63) / 64;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  const uint32_t w = blockIdx.x * kWavesPerBlock + wid;
  for (uint64_t k = 0; k * nwaves < nvb; ++k) {
    const uint64_t vb = skewed_unit(k, w, nwaves);
    if (vb < nvb) clear_block<TW>(F, Fm, actbm, static_cast<uint32_t>(vb));
  }
}

__global__ __launch_bounds__(kBlock) void k_items(const uint64_t *actbm, uint32_t V,
                                                 const uint64_t *xoff, ull *items,
                                                 ull *nitems, const uint64_t *fm,
                                                 uint64_t pw);

// Launchers (vs_vexpand.hip).
void launch_pull(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                 const HopArgs &a);
void launch_push(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                 const HopArgs &a);
void launch_epi(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                const HopArgs &a);
void launch_clear(uint32_t tw, uint32_t grid, hipStream_t s, uint64_t *F, uint64_t *Fm,
                  uint64_t *bm, uint32_t V);
void launch_items(uint32_t grid, hipStream_t s, const uint64_t *bm, uint32_t V,
                  const uint64_t *xoff, ull *items, ull *nitems,
                  const uint64_t *fm = nullptr, uint64_t pw = 0);
void launch_wpush(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                  const HopArgs &a);
uint32_t choose_tile_words(uint32_t requested, uint32_t W);

} // namespace vs
