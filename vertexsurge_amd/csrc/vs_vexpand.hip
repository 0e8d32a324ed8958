// Single-query VExpand engine: host-driven hop loop over the pull / push
// kernels of vs_kernels.hh, with per-hop push/pull choice, Shortest early
// stop, FindShortest, per-layer export and the final union.
//
// Reference control flow being replaced (src/objects/operator/vexpand.cpp):
//   before_execution  :166-220  sources -> L0 (and visit), k0 = 1024 stacks
//   children/after_child :222-284 one Expand per hop; Shortest stops when
//                        visit_count does not change (that empty hop still
//                        counts in k_actual_max); FindShortest assigns
//                        distances and stops when all targets are found
//   merge_to_MT       :298-354  union of L_k_min .. L_k_actual_max, or visit
//                        itself for Shortest with k_min == 0
//   copy_to_multi_MT  :356-378  with_distance: one matrix per hop
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>

#include "vs_kernels.hh"
#include "vs_sched.hh"

namespace vs {

__global__ __launch_bounds__(kBlock) void k_items(const uint64_t *actbm, uint32_t V,
                                                 const uint64_t *xoff, ull *items,
                                                 ull *nitems, const uint64_t *fm,
                                                 uint64_t pw) {
  __shared__ ull s_wsum[kWavesPerBlock];
  __shared__ ull s_base;
  const uint32_t nvb = (V + 63) / 64;
  for (uint32_t base = blockIdx.x * kBlock; base < nvb; base += gridDim.x * kBlock)
    items_slice(actbm, nvb, base, xoff, items, nitems, s_wsum, &s_base, fm, pw);
}

// L0: row r has only its source vertex (ExpandStatus::init_first_m_layer,
// src/include/operator/expand_impl.hh:250-277).
__global__ void k_init(const uint32_t *src, uint32_t S, uint32_t V, uint32_t TW,
                       uint64_t *F, uint64_t *Fm, uint64_t *actbm, uint64_t *Vis,
                       uint64_t *R) {
  const uint32_t nvb = (V + 63) / 64;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < S;
       r += gridDim.x * blockDim.x) {
    const uint32_t s = src[r];
    const uint32_t w = r / 64, t = w / TW, j = w % TW;
    const ull bit = 1ull << (r % 64);
    const size_t wi = (static_cast<size_t>(t) * V + s) * TW + j;
    atomicOr(reinterpret_cast<ull *>(&F[wi]), bit);
    atomicOr(reinterpret_cast<ull *>(&Fm[static_cast<size_t>(t) * V + s]), 1ull << j);
    if (Vis) atomicOr(reinterpret_cast<ull *>(&Vis[wi]), bit);
    if (R) atomicOr(reinterpret_cast<ull *>(&R[wi]), bit);
    atomicOr(reinterpret_cast<ull *>(&actbm[static_cast<size_t>(t) * nvb + s / 64]),
             1ull << (s % 64));
  }
}

// Result of Shortest with k_min == 1 is visit minus the L0 bits.  Counts
// words that become zero so the final nonzero-word count stays exact.
__global__ void k_clear_sources(const uint32_t *src, uint32_t S, uint32_t V,
                                uint32_t TW, uint64_t *R, ull *zeroed) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < S;
       r += gridDim.x * blockDim.x) {
    const uint32_t s = src[r];
    const uint32_t w = r / 64, t = w / TW, j = w % TW;
    const ull bit = 1ull << (r % 64);
    const size_t wi = (static_cast<size_t>(t) * V + s) * TW + j;
    const ull old = atomicAnd(reinterpret_cast<ull *>(&R[wi]), ~bit);
    if (old == bit) atomicAdd(zeroed, 1ull);
  }
}

// FindShortest target test for layer L_hop (vexpand.cpp:246-258).
__global__ void k_fs_check(const uint64_t *L, const uint32_t *tgt, uint32_t S,
                           uint32_t V, uint32_t TW, int32_t *dist, int32_t hop,
                           ull *found) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < S;
       r += gridDim.x * blockDim.x) {
    if (dist[r] >= 0) continue;
    const uint32_t w = r / 64, t = w / TW, j = w % TW;
    const uint64_t word = L[(static_cast<size_t>(t) * V + tgt[r]) * TW + j];
    if ((word >> (r % 64)) & 1) {
      dist[r] = hop;
      atomicAdd(found, 1ull);
    }
  }
}

template <int TW>
static void launch_pull_tw(bool shortest, uint32_t grid, hipStream_t s,
                           const HopArgs &a) {
  const bool sparse = a.Fm != nullptr;
  if (shortest && sparse)
    hipLaunchKernelGGL((k_pull<TW, true, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else if (shortest)
    hipLaunchKernelGGL((k_pull<TW, true, false>), dim3(grid), dim3(kBlock), 0, s, a);
  else if (sparse)
    hipLaunchKernelGGL((k_pull<TW, false, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_pull<TW, false, false>), dim3(grid), dim3(kBlock), 0, s, a);
}

template <int TW>
static void launch_push_tw(bool shortest, uint32_t grid, hipStream_t s,
                           const HopArgs &a) {
  hipLaunchKernelGGL((k_push<TW>), dim3(grid), dim3(kBlock), 0, s, a);
  if (shortest) hipLaunchKernelGGL((k_push_epi<TW, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else hipLaunchKernelGGL((k_push_epi<TW, false>), dim3(grid), dim3(kBlock), 0, s, a);
}

#define VS_TW_CALL(tw, FN, ...)            \
  switch (tw) {                            \
    case 1: FN<1>(__VA_ARGS__); break;     \
    case 2: FN<2>(__VA_ARGS__); break;     \
    case 4: FN<4>(__VA_ARGS__); break;     \
    case 8: FN<8>(__VA_ARGS__); break;     \
    case 16: FN<16>(__VA_ARGS__); break;   \
    case 32: FN<32>(__VA_ARGS__); break;   \
    default: FN<64>(__VA_ARGS__); break;   \
  }

void launch_pull(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                 const HopArgs &a) {
  VS_TW_CALL(tw, launch_pull_tw, shortest, grid, s, a);
}

void launch_push(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                 const HopArgs &a) {
  VS_TW_CALL(tw, launch_push_tw, shortest, grid, s, a);
}

template <int TW>
static void launch_epi_tw(bool shortest, uint32_t grid, hipStream_t s, const HopArgs &a) {
  if (shortest) hipLaunchKernelGGL((k_push_epi<TW, true>), dim3(grid), dim3(kBlock), 0, s, a);
  else hipLaunchKernelGGL((k_push_epi<TW, false>), dim3(grid), dim3(kBlock), 0, s, a);
}

void launch_epi(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                const HopArgs &a) {
  VS_TW_CALL(tw, launch_epi_tw, shortest, grid, s, a);
}

template <int TW>
static void launch_clear_tw(uint32_t grid, hipStream_t s, uint64_t *F, uint64_t *Fm,
                            uint64_t *bm, uint32_t V) {
  hipLaunchKernelGGL((k_clear<TW>), dim3(grid), dim3(kBlock), 0, s, F, Fm, bm, V);
}

void launch_clear(uint32_t tw, uint32_t grid, hipStream_t s, uint64_t *F, uint64_t *Fm,
                  uint64_t *bm, uint32_t V) {
  VS_TW_CALL(tw, launch_clear_tw, grid, s, F, Fm, bm, V);
}


void launch_items(uint32_t grid, hipStream_t s, const uint64_t *bm, uint32_t V,
                  const uint64_t *xoff, ull *items, ull *nitems, const uint64_t *fm,
                  uint64_t pw) {
  hipLaunchKernelGGL(k_items, dim3(grid), dim3(kBlock), 0, s, bm, V, xoff, items,
                     nitems, fm, pw);
}

template <int TW>
static void launch_wpush_tw(bool shortest, uint32_t grid, hipStream_t s, const HopArgs &a) {
  (void)shortest;
  hipLaunchKernelGGL((k_wpush<TW>), dim3(grid), dim3(kBlock), 0, s, a);
}

void launch_wpush(uint32_t tw, bool shortest, uint32_t grid, hipStream_t s,
                  const HopArgs &a) {
  VS_TW_CALL(tw, launch_wpush_tw, shortest, grid, s, a);
}

uint32_t choose_tile_words(uint32_t requested, uint32_t W) {
  uint32_t tw = requested ? requested : 64;
  if (tw > 64) tw = 64;
  tw = pow2ceil(tw);
  const uint32_t need = pow2ceil(std::max<uint32_t>(W, 1));
  return std::min(tw, need);
}

// Compulsory traffic of hop i (bytes), from the counters of L_{i-1}, L_i.
uint64_t hop_bytes(const HopCounters &prev, const HopCounters &cur, bool shortest,
                   bool any_union, uint64_t vis_prev, uint64_t vis_cur,
                   uint64_t r_prev, uint64_t r_cur) {
  uint64_t b = 8 * prev.nz + 8 * cur.nz + 4 * prev.actdeg + 16 * prev.act;
  if (shortest) b += 8 * (vis_prev + vis_cur);
  if (any_union) b += 8 * (r_prev + r_cur);
  return b;
}

// A scoped set of pool buffers returned on every exit path.
struct Lease {
  vs_graph_s *g;
  struct Item { void *p; size_t bytes; bool clean; bool keep; };
  std::vector<Item> items;
  explicit Lease(vs_graph_s *gg) : g(gg) {}
  vs_status get(size_t bytes, bool zero, void **out) {
    VS_TRY(g->pool.get(bytes, zero, g->stream, out));
    items.push_back({*out, bytes, false, false});  // dirty unless marked clean
    return VS_OK;
  }
  void mark(void *p, bool clean) {
    for (auto &it : items) if (it.p == p) it.clean = clean;
  }
  void keep(void *p) {
    for (auto &it : items) if (it.p == p) it.keep = true;
  }
  ~Lease() {
    for (auto &it : items)
      if (!it.keep) g->pool.put(it.p, it.bytes, it.clean);
  }
};

struct Events {
  std::vector<hipEvent_t> ev;
  ~Events() { for (auto e : ev) (void)hipEventDestroy(e); }
  hipEvent_t make() {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    ev.push_back(e);
    return e;
  }
};

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (!a || !b || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0;
  }
  return ms;
}

// Host-side counters of L0 for tile t (sources are on the host anyway).
static void l0_counters(const uint32_t *src, uint32_t S, uint32_t TW, uint32_t t,
                        const std::vector<uint64_t> &xoff, bool vis, bool r,
                        HopCounters &c, uint64_t *wdeg) {
  std::memset(&c, 0, sizeof(c));
  std::fill(wdeg, wdeg + 64, 0ull);
  const uint64_t r0 = static_cast<uint64_t>(t) * TW * 64;
  const uint64_t r1 = std::min<uint64_t>(S, r0 + static_cast<uint64_t>(TW) * 64);
  std::unordered_set<uint64_t> words;
  std::unordered_set<uint32_t> verts;
  for (uint64_t row = r0; row < r1; ++row) {
    const uint32_t s = src[row];
    const uint64_t d = xoff[s + 1] - xoff[s];
    c.pop += 1;
    c.te += d;
    if (words.insert((row / 64) << 32 | s).second) {
      c.nz += 1;
      c.nzdeg += d;
      wdeg[(row / 64) % TW] += d;
    }
    if (verts.insert(s).second) {
      c.act += 1;
      c.actdeg += d;
    }
  }
  c.visnew = vis ? c.nz : 0;
  c.rnew = r ? c.nz : 0;
}

struct RunOut {
  vs_result_s *res = nullptr;
};

static vs_status vexpand_run(vs_graph_s *g, const uint32_t *sources, uint32_t S,
                             const vs_vexpand_params &p, const uint32_t *targets,
                             int32_t *dist, vs_result_s **out,
                             vs_vexpand_stats *st) {
  const bool find_shortest = targets != nullptr;
  const int path = find_shortest ? VS_PATH_FIND_SHORTEST : p.path_type;
  if (path != VS_PATH_ANY && path != VS_PATH_SHORTEST && path != VS_PATH_FIND_SHORTEST) {
    set_error("vs_vexpand: unknown path_type %d", p.path_type);
    return VS_ERR_INVALID;
  }
  if (!find_shortest && p.path_type == VS_PATH_FIND_SHORTEST) {
    set_error("vs_vexpand: FindShortest needs targets; use vs_find_shortest");
    return VS_ERR_INVALID;
  }
  if (p.k_max >= VS_MAX_HOPS) {
    set_error("vs_vexpand: k_max %u exceeds %d", p.k_max, VS_MAX_HOPS - 1);
    return VS_ERR_INVALID;
  }
  if (p.sched < VS_SCHED_HEURISTIC || p.sched > VS_SCHED_LEARNED) {
    set_error("vs_vexpand: unknown sched %d", p.sched);
    return VS_ERR_INVALID;
  }
  if (p.tile_words > 64 || (p.tile_words & (p.tile_words - 1))) {
    set_error("vs_vexpand: tile_words %u is not a power of two <= 64", p.tile_words);
    return VS_ERR_INVALID;
  }
  if (S > 0 && !sources) { set_error("vs_vexpand: null sources"); return VS_ERR_INVALID; }
  for (uint32_t r = 0; r < S; ++r) {
    if (sources[r] >= g->V || (targets && targets[r] >= g->V)) {
      set_error("vs_vexpand: row %u vertex out of range (V=%u)", r, g->V);
      return VS_ERR_INVALID;
    }
  }
  VS_HIP(hipSetDevice(g->device));
  const DevCSR *push = nullptr, *pull = nullptr;
  VS_TRY(graph_csrs(g, p.direction, &push, &pull));

  const bool shortest = path != VS_PATH_ANY;
  const bool with_distance = !find_shortest && p.with_distance;
  const uint32_t V = g->V, nvb = (V + 63) / 64;
  const uint32_t W = (S + 63) / 64;
  const uint32_t k_min = p.k_min, k_max = p.k_max;
  const bool timing = (p.flags & VS_VX_TIMING) != 0;
  // Row packing (below) and the learned tile-width choice read the sources'
  // 2-hop volumes sum_{u in N(s)} deg(u), from one small kernel.
  const bool pack = !(p.flags & VS_VX_KEEP_ORDER) && W > 1 &&
                    k_max >= (find_shortest ? 2u : 3u);
  const bool learn_tw = !find_shortest && p.sched == VS_SCHED_LEARNED && g->sched &&
                        p.tile_words == 0 && g->tuning.tile_words == 0;
  std::vector<uint64_t> key;
  QueryFeatures qf;
  if (S && (pack || learn_tw || timing)) {
    key.resize(S);
    VS_TRY(source_volumes(g, push, sources, S, key.data()));
    qf.S = S;
    qf.k_max = k_max;
    qf.V = g->V;
    qf.E = push->E;
    for (uint32_t r = 0; r < S; ++r) {
      qf.vol2 += key[r];
      qf.degsum += push->hoff[sources[r] + 1] - push->hoff[sources[r]];
    }
  }
  uint32_t TW = choose_tile_words(p.tile_words ? p.tile_words : g->tuning.tile_words, W);
  bool tw_learned = false;
  if (learn_tw && S) {
    const uint32_t pick = sched_tile_words(g->sched, qf, choose_tile_words(64, W), TW);
    tw_learned = pick != TW;
    TW = pick;
  }
  const uint32_t T = S ? (W + TW - 1) / TW : 0;
  const size_t tile_words = static_cast<size_t>(V) * TW;
  const size_t mat_bytes = static_cast<size_t>(T) * tile_words * 8;
  const size_t bm_bytes = static_cast<size_t>(T) * nvb * 8;
  hipStream_t s = g->stream;

  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->tile_words = TW;
    st->num_tiles = T;
    st->src_degsum = qf.degsum;
    st->src_vol2 = qf.vol2;
    st->tile_words_learned = tw_learned ? 1u : 0u;
  }
  // Row packing: the rows run in packed order, sources and targets permuted
  // to match; pos[r] is caller row r's packed row (same tile).  The result
  // keeps the packed order and pos as its stored-row map, which the
  // row-indexed outputs apply when they are read.  Packing pays from hop 2
  // on (fewer needed words per pull): used from k_max = 3, or 2 for
  // FindShortest, whose per-row distances are permuted on the host.
  const uint32_t *const caller_sources = sources;
  const uint32_t RT = 64 * TW;
  std::vector<uint32_t> pk_src, pk_tgt, pos;
  if (pack) {
    std::vector<uint32_t> order(S);
    for (uint32_t r = 0; r < S; ++r) order[r] = r;
    for (uint32_t t0 = 0; t0 < S; t0 += RT) {
      const uint32_t t1 = std::min<uint32_t>(S, t0 + RT);
      std::stable_sort(order.begin() + t0, order.begin() + t1,
                       [&](uint32_t x, uint32_t y) { return key[x] > key[y]; });
    }
    pk_src.resize(S);
    pos.resize(S);
    for (uint32_t i = 0; i < S; ++i) {
      pk_src[i] = sources[order[i]];
      pos[order[i]] = i;
    }
    if (targets) {
      pk_tgt.resize(S);
      for (uint32_t i = 0; i < S; ++i) pk_tgt[i] = targets[order[i]];
      targets = pk_tgt.data();
    }
    sources = pk_src.data();
  }
  std::unique_ptr<vs_result_s> res(new vs_result_s);
  res->g = g;
  res->S = S;
  res->V = V;
  res->tw = TW;
  res->ntiles = T;
  if (!find_shortest && S) res->row_map.assign(caller_sources, caller_sources + S);

  if (S == 0) {  // the reference warns and skips (vexpand.cpp:170-173)
    res->k_actual_max = 0;
    if (!with_distance) res->mats.push_back(Mat{});
    g->refs++;
    *out = res.release();
    return VS_OK;
  }

  Lease lease(g);
  Events evs;
  hipEvent_t ev0 = evs.make(), ev_hops0 = evs.make(), ev_hops1 = evs.make(),
             ev1 = evs.make();
  VS_HIP(hipEventRecord(ev0, s));

  void *d_src = nullptr, *d_tgt = nullptr, *d_dist = nullptr, *d_found = nullptr;
  VS_TRY(lease.get(static_cast<size_t>(S) * 4, false, &d_src));
  VS_HIP(hipMemcpyAsync(d_src, sources, static_cast<size_t>(S) * 4, hipMemcpyHostToDevice, s));
  if (find_shortest) {
    VS_TRY(lease.get(static_cast<size_t>(S) * 4, false, &d_tgt));
    VS_TRY(lease.get(static_cast<size_t>(S) * 4, false, &d_dist));
    VS_TRY(lease.get(8, true, &d_found));
    VS_HIP(hipMemcpyAsync(d_tgt, targets, static_cast<size_t>(S) * 4, hipMemcpyHostToDevice, s));
  }

  // frontier buffers: ping-pong, or one per layer when layers are exported
  std::vector<uint64_t *> layer(k_max + 1, nullptr);
  uint64_t *Fbuf[2] = {nullptr, nullptr}, *BM[2] = {nullptr, nullptr};
  uint64_t *FM[2] = {nullptr, nullptr};  // word masks of the two frontiers in use
  void *tmp = nullptr;
  VS_TRY(lease.get(mat_bytes, true, &tmp));
  Fbuf[0] = static_cast<uint64_t *>(tmp);
  const size_t wm_bytes = static_cast<size_t>(T) * V * 8;
  VS_TRY(lease.get(wm_bytes, true, &tmp));
  FM[0] = static_cast<uint64_t *>(tmp);
  VS_TRY(lease.get(wm_bytes, true, &tmp));
  FM[1] = static_cast<uint64_t *>(tmp);
  VS_TRY(lease.get(bm_bytes, true, &tmp));
  BM[0] = static_cast<uint64_t *>(tmp);
  VS_TRY(lease.get(bm_bytes, true, &tmp));
  BM[1] = static_cast<uint64_t *>(tmp);
  if (!with_distance) {
    VS_TRY(lease.get(mat_bytes, true, &tmp));
    Fbuf[1] = static_cast<uint64_t *>(tmp);
  }
  layer[0] = Fbuf[0];
  uint64_t *Vis = nullptr, *R = nullptr;
  if (shortest) {
    VS_TRY(lease.get(mat_bytes, true, &tmp));
    Vis = static_cast<uint64_t *>(tmp);
  }
  const bool r_accum = !with_distance && !find_shortest &&
                       (!shortest || k_min >= 2);
  if (r_accum) {
    VS_TRY(lease.get(mat_bytes, true, &tmp));
    R = static_cast<uint64_t *>(tmp);
  }
  uint64_t *N = nullptr, *touched = nullptr;
  ull *items = nullptr;
  const size_t items_cap = static_cast<size_t>(V) + push->E / kPushChunk + 64;
  auto ensure_push = [&]() -> vs_status {
    if (N) return VS_OK;
    VS_TRY(lease.get(mat_bytes, true, &tmp));
    N = static_cast<uint64_t *>(tmp);
    VS_TRY(lease.get(bm_bytes, true, &tmp));
    touched = static_cast<uint64_t *>(tmp);
    VS_TRY(lease.get(static_cast<size_t>(T) * items_cap * 8, false, &tmp));
    items = static_cast<ull *>(tmp);
    return VS_OK;
  };
  const size_t nctr = static_cast<size_t>(k_max + 1) * T;
  void *d_ctr_v = nullptr;
  VS_TRY(lease.get(nctr * sizeof(HopCounters), true, &d_ctr_v));
  HopCounters *d_ctr = static_cast<HopCounters *>(d_ctr_v);
  // per (hop, tile) per-word OR of the layer's rows; L0 holds every row
  void *d_front_v = nullptr;
  VS_TRY(lease.get(nctr * 64 * 8, true, &d_front_v));
  uint64_t *d_front = static_cast<uint64_t *>(d_front_v);
  {
    std::vector<uint64_t> f0(static_cast<size_t>(T) * 64, 0);
    for (uint32_t r = 0; r < S; ++r) {
      const uint32_t w = r / 64;
      f0[static_cast<size_t>(w / TW) * 64 + (w % TW)] |= 1ull << (r % 64);
    }
    VS_HIP(hipMemcpyAsync(d_front, f0.data(), f0.size() * 8, hipMemcpyHostToDevice, s));
    VS_HIP(hipStreamSynchronize(s));  // f0 leaves scope
  }
  std::vector<HopCounters> hctr(nctr);
  auto C = [&](uint32_t hop, uint32_t t) -> HopCounters & { return hctr[static_cast<size_t>(hop) * T + t]; };
  // per (hop, tile) per-word push degrees of L_hop: kWParts partial copies
  // on the device, folded here after each hop
  const size_t wparts = static_cast<size_t>(kWParts) * 64;
  void *d_wdeg_v = nullptr;
  VS_TRY(lease.get(nctr * wparts * 8, true, &d_wdeg_v));
  ull *d_wdeg = static_cast<ull *>(d_wdeg_v);
  std::vector<uint64_t> hwdeg(nctr * 64, 0), wpart(static_cast<size_t>(T) * wparts);
  auto Wd = [&](uint32_t hop, uint32_t t) -> uint64_t * {
    return &hwdeg[(static_cast<size_t>(hop) * T + t) * 64];
  };
  const uint64_t wpush_budget = word_push_budget(g->tuning, push->E);

  // ---- L0
  {
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((S + 255) / 256, 1024));
    hipLaunchKernelGGL(k_init, dim3(grid), dim3(256), 0, s,
                       static_cast<const uint32_t *>(d_src), S, V, TW, Fbuf[0], FM[0],
                       BM[0], Vis, (R && k_min == 0) ? R : nullptr);
    VS_HIP(hipGetLastError());
    for (uint32_t t = 0; t < T; ++t)
      l0_counters(sources, S, TW, t, push->hoff, shortest, R && k_min == 0, C(0, t), Wd(0, t));
  }
  uint32_t found = 0;
  if (find_shortest) {
    std::vector<int32_t> d0(S);
    for (uint32_t r = 0; r < S; ++r) {
      d0[r] = sources[r] == targets[r] ? 0 : -1;
      found += d0[r] == 0;
    }
    VS_HIP(hipMemcpyAsync(d_dist, d0.data(), static_cast<size_t>(S) * 4, hipMemcpyHostToDevice, s));
    VS_HIP(hipStreamSynchronize(s));  // d0 leaves scope
  }

  const uint32_t pull_grid = static_cast<uint32_t>(g->num_cus) * 8;
  const uint32_t clear_grid = grid_for(nvb, g->num_cus);
  const uint32_t item_grid = std::max<uint32_t>(1, std::min<uint32_t>((nvb + kBlock - 1) / kBlock, g->num_cus * 4));
  const uint32_t push_grid = static_cast<uint32_t>(g->num_cus) * 8;

  VS_HIP(hipEventRecord(ev_hops0, s));
  uint32_t k_actual = 0;
  int cur = 0;
  uint64_t vis_nz = 0, r_nz = 0;
  for (uint32_t t = 0; t < T; ++t) {
    vis_nz += C(0, t).visnew;
    r_nz += C(0, t).rnew;
  }
  std::vector<int> tile_mode(T, 0);
  for (uint32_t hop = 1; hop <= k_max; ++hop) {
    uint64_t *F = with_distance ? layer[hop - 1] : Fbuf[cur];
    uint64_t *Fn;
    if (with_distance) {
      VS_TRY(lease.get(mat_bytes, true, &tmp));
      layer[hop] = static_cast<uint64_t *>(tmp);
      Fn = layer[hop];
    } else {
      Fn = Fbuf[cur ^ 1];
      layer[hop] = Fn;
    }
    uint64_t *bmF = BM[cur], *bmN = BM[cur ^ 1];
    hipEvent_t eh0 = timing ? evs.make() : nullptr, eh1 = timing ? evs.make() : nullptr;
    if (timing) VS_HIP(hipEventRecord(eh0, s));
    const bool acc_r = R && hop >= k_min;
    const uint64_t sparse_pct = g->tuning.sparse_pct ? g->tuning.sparse_pct : kSparsePct;
    for (uint32_t t = 0; t < T; ++t) {
      const HopCounters &pc = C(hop - 1, t);
      tile_mode[t] = 0;
      if (pc.pop == 0) continue;  // empty frontier: L_hop of this tile is empty
      int mode;
      if (p.sched == VS_SCHED_PUSH) mode = 1;
      else if (p.sched == VS_SCHED_PULL) mode = 2;
      else {
        SchedFeatures f;
        f.hop = hop;
        f.rows = std::min<uint64_t>(S - static_cast<uint64_t>(t) * TW * 64,
                                    static_cast<uint64_t>(TW) * 64);
        f.tw = TW;
        f.V = V;
        f.E = push->E;
        f.pop = pc.pop;
        f.act = pc.act;
        f.actdeg = pc.actdeg;
        f.vis_frac = shortest ? static_cast<double>(vis_nz) /
                                    (static_cast<double>(T) * tile_words)
                              : 0.0;
        mode = (p.sched == VS_SCHED_LEARNED && g->sched)
                   ? sched_decide(g->sched, f, g->tuning)
                   : heuristic_decide(f, g->tuning);
      }
      tile_mode[t] = mode;
      HopArgs a{};
      a.V = V;
      a.S = S;
      a.word0 = t * TW;
      // most in-edges come from active vertices: skip the bitmap lookup
      a.dense = pc.actdeg * 4 > 3 * pull->E ? 1u : 0u;
      a.poff = pull->off;
      a.padj = pull->adj;
      a.xoff = push->off;
      a.xadj = push->adj;
      a.F = F + t * tile_words;
      // read the word masks only when the rows a pull gathers (weighted by
      // how often: degree) are mostly empty words
      a.Fm = pc.nzdeg * 100 < pc.actdeg * TW * sparse_pct ? FM[cur] + static_cast<size_t>(t) * V
                                                         : nullptr;
      a.actbm = bmF + static_cast<size_t>(t) * nvb;
      a.Fn = Fn + t * tile_words;
      a.Fnm = FM[cur ^ 1] + static_cast<size_t>(t) * V;
      a.actbm_n = bmN + static_cast<size_t>(t) * nvb;
      a.Vis = Vis ? Vis + t * tile_words : nullptr;
      a.R = acc_r ? R + t * tile_words : nullptr;
      a.ctr = d_ctr + static_cast<size_t>(hop) * T + t;
      a.live = nullptr;
      a.live_next = nullptr;
      a.racc = nullptr;
      a.front = d_front + (static_cast<size_t>(hop - 1) * T + t) * 64;
      a.front_next = d_front + (static_cast<size_t>(hop) * T + t) * 64;
      a.wdeg = d_wdeg + (static_cast<size_t>(hop) * T + t) * wparts;
      a.pw = 0;
      a.Fwm = nullptr;
      hipEvent_t k0 = timing ? evs.make() : nullptr, k1 = timing ? evs.make() : nullptr;
      if (mode == 1) {
        VS_TRY(ensure_push());
        a.N = N + t * tile_words;
        a.touched = touched + static_cast<size_t>(t) * nvb;
        a.items = items + static_cast<size_t>(t) * items_cap;
        a.nitems = &d_ctr[static_cast<size_t>(hop - 1) * T + t].items;
        launch_items(item_grid, s, a.actbm, V, push->off,
                     const_cast<ull *>(a.items), const_cast<ull *>(a.nitems));
        if (timing) VS_HIP(hipEventRecord(k0, s));
        launch_push(TW, shortest, push_grid, s, a);
        if (timing) VS_HIP(hipEventRecord(k1, s));
      } else {
        VS_TRY(ensure_pull_items(*const_cast<DevCSR *>(pull), V, g->tuning.pull_hubs));
        a.pitems = pull->pitems;
        a.padj = pull->hadj;  // hubs-first lists
        a.hubcnt = pull->hubcnt;
        a.npitems = pull->npitems;
        a.items = nullptr;
        a.nitems = nullptr;
        a.N = nullptr;
        a.touched = nullptr;
        // word push: only when whole rows are gathered (the word-mask pull
        // reads only the words that hold bits anyway)
        const uint64_t pw = a.Fm ? 0 : word_push_pick(Wd(hop - 1, t), TW, wpush_budget);
        if (pull->has_heavy || pw) {  // hub partials / pushed words meet in N
          VS_TRY(ensure_push());
          a.N = N + t * tile_words;
          a.touched = touched + static_cast<size_t>(t) * nvb;
        }
        if (timing) VS_HIP(hipEventRecord(k0, s));
        if (pw) {
          a.pw = pw;
          a.Fwm = FM[cur] + static_cast<size_t>(t) * V;
          a.items = items + static_cast<size_t>(t) * items_cap;
          a.nitems = &d_ctr[static_cast<size_t>(hop - 1) * T + t].items;
          launch_items(item_grid, s, a.actbm, V, push->off, const_cast<ull *>(a.items),
                       const_cast<ull *>(a.nitems), a.Fwm, pw);
          launch_wpush(TW, shortest, push_grid, s, a);
          if (st) st->wpush_words[hop - 1] += __builtin_popcountll(pw);
        }
        launch_pull(TW, shortest, pull_grid, s, a);
        if (timing) VS_HIP(hipEventRecord(k1, s));
        if (pull->has_heavy) {  // only the words that hold chunked vertices
          a.ewords = pull->hwords;
          a.newords = pull->nhwords;
          launch_epi(TW, shortest, push_grid, s, a);
        }
      }
      VS_HIP(hipGetLastError());
      if (timing && st) {
        VS_HIP(hipEventSynchronize(k1));
        const double ms = ev_ms(k0, k1);
        if (mode == 1) { st->ms_push += ms; st->n_push++; }
        else { st->ms_pull += ms; st->n_pull++; }
      }
    }
    if (timing) VS_HIP(hipEventRecord(eh1, s));
    VS_HIP(hipMemcpyAsync(&C(hop, 0), d_ctr + static_cast<size_t>(hop) * T,
                          sizeof(HopCounters) * T, hipMemcpyDeviceToHost, s));
    VS_HIP(hipMemcpyAsync(wpart.data(), d_wdeg + static_cast<size_t>(hop) * T * wparts,
                          wpart.size() * 8, hipMemcpyDeviceToHost, s));
    if (find_shortest) {
      const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((S + 255) / 256, 1024));
      hipLaunchKernelGGL(k_fs_check, dim3(grid), dim3(256), 0, s, Fn,
                         static_cast<const uint32_t *>(d_tgt), S, V, TW,
                         static_cast<int32_t *>(d_dist), static_cast<int32_t>(hop),
                         static_cast<ull *>(d_found));
      VS_HIP(hipGetLastError());
    }
    ull found_now = 0;
    if (find_shortest)
      VS_HIP(hipMemcpyAsync(&found_now, d_found, 8, hipMemcpyDeviceToHost, s));
    VS_HIP(hipStreamSynchronize(s));
    k_actual = hop;
    for (uint32_t t = 0; t < T; ++t)
      for (size_t q = 0; q < wparts; ++q) Wd(hop, t)[q % 64] += wpart[t * wparts + q];
    uint64_t newbits = 0, prev_vis = vis_nz, prev_r = r_nz, te = 0, bytes = 0;
    uint64_t frontier_bits = 0, frontier_act = 0, frontier_deg = 0;
    for (uint32_t t = 0; t < T; ++t) {
      newbits += C(hop, t).pop;
      vis_nz += C(hop, t).visnew;
      r_nz += C(hop, t).rnew;
      te += C(hop - 1, t).te;
      frontier_bits += C(hop - 1, t).pop;
      frontier_act += C(hop - 1, t).act;
      frontier_deg += C(hop - 1, t).actdeg;
    }
    const bool any_union = !shortest && acc_r;
    for (uint32_t t = 0; t < T; ++t) {
      // per-tile share of vis / r nonzero words is not tracked separately;
      // attribute the whole-query terms once (tile 0)
      const uint64_t b = hop_bytes(C(hop - 1, t), C(hop, t), false, false, 0, 0, 0, 0) +
                         (t == 0 ? (shortest ? 8 * (prev_vis + vis_nz) : 0) +
                                       (any_union ? 8 * (prev_r + r_nz) : 0)
                                 : 0);
      bytes += b;
      if (st) {
        if (tile_mode[t] == 1) st->bytes_push += b;
        else if (tile_mode[t] == 2) st->bytes_pull += b;
      }
    }
    if (st) {
      st->te += te;
      st->te_hop[hop - 1] = te;
      st->bytes_hop[hop - 1] = bytes;
      st->bytes_alg += bytes;
      st->frontier_bits[hop - 1] = frontier_bits;
      st->frontier_act[hop - 1] = frontier_act;
      st->frontier_deg[hop - 1] = frontier_deg;
      st->vis_frac[hop - 1] = shortest ? static_cast<double>(prev_vis) /
                                             (static_cast<double>(T) * tile_words)
                                       : 0.0;
      st->new_bits[hop - 1] = newbits;
      int m = 0;
      for (uint32_t t = 0; t < T; ++t) if (tile_mode[t]) m = (m == 0 || m == tile_mode[t]) ? tile_mode[t] : 3;
      st->mode_hop[hop - 1] = m;
      if (timing) st->ms_hop[hop - 1] = ev_ms(eh0, eh1);
    }
    // Shortest / FindShortest stop: visit_count unchanged (vexpand.cpp:229-233,
    // 267-272); FindShortest also stops when every target is found (:262-266).
    bool stop = shortest && newbits == 0;
    if (find_shortest && found + found_now >= S) stop = true;
    // frontier L_{hop-1} is consumed: zero its rows (ping-pong) or its bitmap
    if (!with_distance) {
      for (uint32_t t = 0; t < T; ++t)
        launch_clear(TW, clear_grid, s, F + t * tile_words, FM[cur] + static_cast<size_t>(t) * V,
                     bmF + static_cast<size_t>(t) * nvb, V);
      VS_HIP(hipGetLastError());
    } else {
      VS_HIP(hipMemsetAsync(bmF, 0, bm_bytes, s));
      VS_HIP(hipMemsetAsync(FM[cur], 0, wm_bytes, s));
    }
    cur ^= 1;
    if (stop) break;
  }
  VS_HIP(hipEventRecord(ev_hops1, s));

  // ---- finalize
  uint64_t r_final_nz = 0;
  std::vector<int32_t> pk_dist;  // distances in packed row order
  if (find_shortest) {
    int32_t *dd = dist;
    if (pack) {
      pk_dist.resize(S);
      dd = pk_dist.data();
    }
    VS_HIP(hipMemcpyAsync(dd, d_dist, static_cast<size_t>(S) * 4, hipMemcpyDeviceToHost, s));
  } else if (with_distance) {
    for (uint32_t i = k_min; i <= k_actual; ++i) {
      Mat m;
      m.d = layer[i];
      m.bytes = mat_bytes;
      m.distance = i;
      res->mats.push_back(m);
      lease.keep(layer[i]);
      for (uint32_t t = 0; t < T; ++t) r_final_nz += C(i, t).nz;
    }
  } else {
    Mat m;
    m.bytes = mat_bytes;
    if (shortest && k_min <= 1) {
      m.d = Vis;
      r_final_nz = vis_nz;
      if (k_min == 1) {
        void *z = nullptr;
        VS_TRY(lease.get(8, true, &z));
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((S + 255) / 256, 1024));
        hipLaunchKernelGGL(k_clear_sources, dim3(grid), dim3(256), 0, s,
                           static_cast<const uint32_t *>(d_src), S, V, TW, Vis,
                           static_cast<ull *>(z));
        VS_HIP(hipGetLastError());
        ull zh = 0;
        VS_HIP(hipMemcpyAsync(&zh, z, 8, hipMemcpyDeviceToHost, s));
        VS_HIP(hipStreamSynchronize(s));
        r_final_nz -= zh;
      }
    } else {
      m.d = R;
      r_final_nz = r_nz;
    }
    lease.keep(m.d);
    res->mats.push_back(m);
  }
  // the last frontier (ping-pong mode) still holds L_k_actual: clear it so
  // the buffer returns to the pool clean
  if (!with_distance) {
    uint64_t *F = Fbuf[cur];
    for (uint32_t t = 0; t < T; ++t)
      launch_clear(TW, clear_grid, s, F + t * tile_words, FM[cur] + static_cast<size_t>(t) * V,
                   BM[cur] + static_cast<size_t>(t) * nvb, V);
    VS_HIP(hipGetLastError());
    lease.mark(Fbuf[0], true);
    lease.mark(Fbuf[1], true);
  } else {
    VS_HIP(hipMemsetAsync(BM[cur], 0, bm_bytes, s));
    VS_HIP(hipMemsetAsync(FM[cur], 0, wm_bytes, s));
    // layers that are not exported go back dirty (default)
  }
  lease.mark(FM[0], true);
  lease.mark(FM[1], true);
  lease.mark(BM[0], true);
  lease.mark(BM[1], true);
  if (N) {
    lease.mark(N, true);
    lease.mark(touched, true);
  }
  if (Vis && res->mats.empty()) lease.mark(Vis, false);
  if (pack) res->stored = pos;  // the matrices stay in packed row order
  VS_HIP(hipEventRecord(ev1, s));
  VS_HIP(hipStreamSynchronize(s));
  if (find_shortest && pack)
    for (uint32_t r = 0; r < S; ++r) dist[r] = pk_dist[pos[r]];
  res->k_actual_max = k_actual;
  if (st) {
    st->k_actual_max = k_actual;
    st->bytes_alg += 8 * r_final_nz;
    st->visit_count = 0;
    if (shortest) {
      uint64_t vc = 0;
      for (uint32_t hop = 0; hop <= k_actual; ++hop)
        for (uint32_t t = 0; t < T; ++t) vc += C(hop, t).pop;
      st->visit_count = vc;
    }
    st->ms_total = ev_ms(ev0, ev1);
    st->ms_hops = ev_ms(ev_hops0, ev_hops1);
  }
  if (find_shortest) {
    res.reset();
    *out = nullptr;
    return VS_OK;
  }
  g->refs++;
  *out = res.release();
  return VS_OK;
}

} // namespace vs

using namespace vs;

extern "C" {

vs_status vs_vexpand(vs_graph_t g, const uint32_t *sources, uint32_t S,
                     const vs_vexpand_params *p, vs_result_t *out,
                     vs_vexpand_stats *stats) try {
  clear_error();
  if (!g || !p || !out) { set_error("vs_vexpand: null argument"); return VS_ERR_INVALID; }
  *out = nullptr;
  vs_result_s *r = nullptr;
  vs_status s = vexpand_run(g, sources, S, *p, nullptr, nullptr, &r, stats);
  if (s == VS_OK) *out = r;
  return s;
} VS_CATCH("vs_vexpand")

vs_status vs_find_shortest(vs_graph_t g, const uint32_t *sources,
                           const uint32_t *targets, uint32_t S,
                           int32_t direction, uint32_t k_max, int32_t *dist,
                           vs_vexpand_stats *stats) try {
  clear_error();
  if (!g || (S > 0 && (!sources || !targets || !dist))) {
    set_error("vs_find_shortest: null argument");
    return VS_ERR_INVALID;
  }
  if (S == 0) return VS_OK;
  vs_vexpand_params p{};
  p.k_min = 1;
  p.k_max = k_max;
  p.path_type = VS_PATH_FIND_SHORTEST;
  p.direction = direction;
  p.sched = VS_SCHED_HEURISTIC;
  vs_result_s *r = nullptr;
  return vexpand_run(g, sources, S, p, targets, dist, &r, stats);
} VS_CATCH("vs_find_shortest")

} // extern "C"
