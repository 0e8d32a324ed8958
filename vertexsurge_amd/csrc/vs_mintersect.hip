// Multi-way intersection over sorted adjacency lists.
//
// The reference (Mintersect::run_internal, src/objects/operator/
// mintersect.cpp:48-345) builds, per base tuple, an all-ones bitmap of V bits,
// ANDs one DenseColMainMatrix column (or a scanned flat dependency) per
// dependency, then pops the set bits in ascending row order
// (BitmapIterator, src/include/storage/bitmap.hh:91-126).  Cost per tuple is
// O(V/64) words per dependency whatever the degrees.
//
// Here each dependency is a CSR in HBM with sorted lists, so the same set is
// the intersection of arity sorted lists.  One wave owns one tuple: the
// shortest list drives, 64 elements per step (one per lane); each lane looks
// its element up in the other lists by binary search, shortest list first and
// each further list only for the lanes still in every list so far (few
// survive the first: a random element is in a list about as often as the
// list is long relative to V), each search starting where the previous
// step's largest searched element landed; ballot + mbcnt give the ascending
// output positions.  Each search first bisects the list's fence -- every 64th entry
// of the adjacency array, a copy 64 times smaller that stays cached for the
// long (hub) lists -- and then the at most 64 entries between two fence
// posts, one or two cache lines: a lookup in a list of a million entries
// misses the cache once or twice instead of at every one of its last eight
// or so probes.
//
// One intersection pass when the lists are wanted: the split also writes each
// tuple's bound (its driver length: the intersection is a subset of the
// driver), whose exclusive scan places every tuple's survivors in a staging
// buffer; the pass writes counts and staged lists at once, then the counts'
// scan gives the final offsets and a copy kernel moves each tuple's run to
// its place.  When the staging buffer (4 bytes per driver entry) would pass
// a size limit, two passes instead: count -> exclusive scan -> fill.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include <rocprim/device/device_scan.hpp>

#include "vs_device.hh"

namespace vs {

constexpr int kMaxArity = 8;

struct MiArgs {
  uint32_t arity;
  uint32_t pad;
  uint64_t T;
  const uint32_t *tuples;
  const uint64_t *off[kMaxArity];
  const uint32_t *adj[kMaxArity];
  const uint32_t *fence[kMaxArity];  // fence[d][f] = adj[d][64 f]
  uint32_t V[kMaxArity];
  uint32_t *counts;
  const uint64_t *offsets;
  uint32_t *lists;
  // tuples split by the length of their shortest list: at most 8, 16, 32
  // entries (eight, four, two tuples per wave), and the rest (one per wave)
  const uint32_t *order[4];
  const ull *norder;  // device counts: the four lists, then the giants (kCounters)
};

constexpr int kBuckets = 4;
__host__ __device__ constexpr uint32_t bucket_len(int k) { return 8u << k; }  // 8, 16, 32
// Tuples whose shortest list exceeds kGiantLen entries ("giants": a few
// hundred of the 4 M random 3-paths of R-MAT 24, up to 40 K entries) are
// counted separately and stored from the back of the long list; each gets
// a wave of its own.  Strided with the others, the wave holding the largest
// took 2.6x the mean work and the kernel waited for it.  The next class
// ("large", over kLargeLen entries, stored from the back of the 32-entry
// bucket's list) is dealt out first, one per wave, so that no wave draws
// several of them on top of its share of the rest.
constexpr int kLong = kBuckets - 1, kGiant = kBuckets, kLarge = kBuckets + 1;
constexpr int kCounters = kBuckets + 2;
constexpr uint64_t kGiantLen = 8192, kLargeLen = 2048;

// A tuple's lists: start and length of each, and the shortest (the driver).
// Returns false if an id is outside its dependency (empty intersection).
template <int N>
__device__ __forceinline__ bool tuple_lists(const MiArgs &a, uint64_t t, uint64_t (&b)[N],
                                            uint32_t (&n)[N], int &drv, uint64_t &best) {
  bool ok = true;
  drv = 0;
  best = ~0ull;
#pragma unroll
  for (int d = 0; d < N; ++d) {
    b[d] = 0;
    n[d] = 0;
    const uint32_t x = a.tuples[t * N + d];
    if (x >= a.V[d]) {
      ok = false;
      continue;
    }
    b[d] = a.off[d][x];
    const uint64_t len = a.off[d][x + 1] - b[d];
    n[d] = static_cast<uint32_t>(len);
    if (len < best) {
      best = len;
      drv = d;
    }
  }
  return ok;
}

// lower_bound of x in list d (positions relative to the list's start, the
// search starting at `from`: everything before it is < x) for the lanes with
// `on`.  Two phases: the fence (every 64th adjacency entry), then at most 64
// entries between two fence posts.
__device__ __forceinline__ uint32_t search_list(const MiArgs &a, int d, uint64_t b, uint32_t n,
                                                uint32_t from, uint32_t x, bool on) {
  uint32_t lo = from, hi = on ? n : 0;
  // fence posts of the list sit at the multiples of 64 in [b + lo, b + hi),
  // fence indices fa + [0, fhi)
  const uint64_t fa = (b + lo + 63) >> 6;
  const uint32_t fnr = on ? static_cast<uint32_t>(((b + hi + 63) >> 6) - fa) : 0;
  uint32_t flo = 0, fhi = fnr;
  const uint32_t *fence = a.fence[d];
  while (flo < fhi) {  // first fence post >= x
    const uint32_t mid = (flo + fhi) >> 1;
    if (fence[fa + mid] < x) flo = mid + 1;
    else fhi = mid;
  }
  if (on && fnr > 0) {
    // post flo-1 is < x: the answer lies after it
    if (flo > 0) lo = max(lo, static_cast<uint32_t>(64 * (fa + flo - 1) + 1 - b));
    // post flo is >= x: the answer is at or before it
    if (flo < fnr) hi = static_cast<uint32_t>(64 * (fa + flo) - b);
  }
  const uint32_t *adj = a.adj[d] + b;
  while (lo < hi) {  // lower_bound within at most 64 entries
    const uint32_t mid = (lo + hi) >> 1;
    if (adj[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Per lane: is x in every list but the driver's?  The other lists are
// searched one after the other, shortest first, each only for the lanes
// whose x is in every list searched so far: a random x is in a list about
// as often as the list is long relative to V, so the shortest list keeps
// few lanes (on the 3-path workload 5 % of the driver elements), and the
// longer lists are searched for those only.  lo[d] returns x's lower bound
// in list d (from[d] for a lane that did not search it); searched[d] is the
// ballot of the lanes that searched list d.
template <int N>
__device__ __forceinline__ bool member_all(const MiArgs &a, const uint64_t (&b)[N],
                                           const uint32_t (&n)[N], int drv,
                                           const uint32_t (&from)[N], uint32_t x, bool have,
                                           uint32_t (&lo)[N], uint64_t (&searched)[N]) {
  // rank of each list by length (ties by index); the driver is skipped
  int rank[N];
#pragma unroll
  for (int d = 0; d < N; ++d) {
    rank[d] = 0;
#pragma unroll
    for (int e = 0; e < N; ++e) rank[d] += (n[e] < n[d]) || (n[e] == n[d] && e < d);
  }
  bool keep = have;
#pragma unroll
  for (int d = 0; d < N; ++d) {
    lo[d] = from[d];
    searched[d] = 0;
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
#pragma unroll
    for (int d = 0; d < N; ++d) {
      if (rank[d] != r || d == drv) continue;
      const uint64_t on = __ballot(keep);
      searched[d] = on;
      if (on == 0) continue;
      lo[d] = search_list(a, d, b[d], n[d], from[d], x, keep);
      keep = keep && lo[d] < n[d] && a.adj[d][b[d] + lo[d]] == x;
    }
  }
  return keep;
}

// What a pass writes: counts only, lists at a.offsets only (their counts are
// known), or both (lists staged at a.offsets = the bounds' scan).
constexpr int kMiCount = 0, kMiFill = 1, kMiStage = 2;

// Tuples with a long shortest list: one wave per tuple, 64 driver elements
// per step; each step's searches start where lane 63's landed.
template <int MODE, int N>
__device__ __forceinline__ void mi_long_tuple(const MiArgs &a, uint64_t t, int lane) {
  constexpr bool WRITE = MODE != kMiCount, COUNT = MODE != kMiFill;
  uint64_t b[N];
  uint32_t n[N];
  int drv;
  uint64_t best;
  const bool ok = tuple_lists<N>(a, t, b, n, drv, best);
  uint32_t total = 0;
  uint64_t out = WRITE ? a.offsets[t] : 0;
  if (ok && best > 0) {
    const uint32_t *dl = a.adj[drv];
    uint32_t from[N];
#pragma unroll
    for (int d = 0; d < N; ++d) from[d] = 0;
    for (uint64_t base = b[drv]; base < b[drv] + best; base += 64) {
      const uint64_t idx = base + lane;
      const bool have = idx < b[drv] + best;
      const uint32_t x = have ? dl[idx] : 0xFFFFFFFFu;
      uint32_t lo[N];
      uint64_t searched[N];
      const bool keep = member_all<N>(a, b, n, drv, from, x, have, lo, searched);
      // the next step's elements are all larger than this step's: a list
      // resumes where its highest searching lane landed
#pragma unroll
      for (int d = 0; d < N; ++d)
        if (searched[d]) from[d] = __shfl(lo[d], 63 - __clzll(searched[d]));
      const uint64_t m = __ballot(keep);
      if (WRITE && keep) a.lists[out + lane_rank(m)] = x;
      out += __popcll(m);
      total += __popcll(m);
    }
  }
  if (COUNT && lane == 0) a.counts[t] = total;
}

// Tuples with a long shortest list: one wave per tuple, 64 driver elements
// per step; each step's searches start where lane 63's landed.  The giants
// (stored backwards from the end of the long list) go first, one wave each
// while there are fewer of them than waves; the other waves share the rest.
template <int MODE, int N>
__global__ __launch_bounds__(kBlock) void k_mintersect(MiArgs a) {
  const int lane = lane_id(), wid = wave_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + wid;
  const uint64_t cnt = a.norder[kLong], ng = a.norder[kGiant], nl = a.norder[kLarge];
  const uint32_t *lst = a.order[kLong], *lrg = a.order[kLong - 1];
  if (ng < nwaves) {
    if (w < ng) {
      mi_long_tuple<MODE, N>(a, lst[a.T - 1 - w], lane);
    } else {
      const uint64_t rw = nwaves - ng;
      for (uint64_t i = w - ng; i < nl; i += rw) mi_long_tuple<MODE, N>(a, lrg[a.T - 1 - i], lane);
      for (uint64_t i = w - ng; i < cnt; i += rw) mi_long_tuple<MODE, N>(a, lst[i], lane);
    }
  } else {
    for (uint64_t i = w; i < ng; i += nwaves) mi_long_tuple<MODE, N>(a, lst[a.T - 1 - i], lane);
    for (uint64_t i = w; i < nl; i += nwaves) mi_long_tuple<MODE, N>(a, lrg[a.T - 1 - i], lane);
    for (uint64_t i = w; i < cnt; i += nwaves) mi_long_tuple<MODE, N>(a, lst[i], lane);
  }
}

// Tuples of bucket K, whose shortest list has at most GS = bucket_len(K)
// entries: 64 / GS tuples per wave, GS lanes each, one step.
template <int MODE, int N, int K>
__global__ __launch_bounds__(kBlock) void k_mintersect_short(MiArgs a) {
  constexpr bool WRITE = MODE != kMiCount, COUNT = MODE != kMiFill;
  constexpr int GS = bucket_len(K), NG = 64 / GS;
  const int lane = lane_id(), wid = wave_id();
  const int g = lane / GS, gl = lane % GS;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  const uint64_t cnt = a.norder[K];
  for (uint64_t w = blockIdx.x * kWavesPerBlock + wid; w * NG < cnt; w += nwaves) {
    const uint64_t i = w * NG + g;
    const bool mine = i < cnt;
    const uint64_t t = mine ? a.order[K][i] : 0;
    uint64_t b[N] = {};
    uint32_t n[N] = {};
    int drv = 0;
    uint64_t best = 0;
    const bool ok = mine && tuple_lists<N>(a, t, b, n, drv, best);
    const bool have = ok && static_cast<uint64_t>(gl) < best;
    const uint32_t x = have ? a.adj[drv][b[drv] + gl] : 0xFFFFFFFFu;
    uint32_t from[N], lo[N];
    uint64_t searched[N];
#pragma unroll
    for (int d = 0; d < N; ++d) from[d] = 0;
    const bool keep = member_all<N>(a, b, n, drv, from, x, have, lo, searched);
    const uint64_t m = __ballot(keep);
    const uint64_t gm = (m >> (g * GS)) & ((1ull << GS) - 1);
    if (WRITE && keep)
      a.lists[a.offsets[t] + __popcll(gm & ((1ull << gl) - 1))] = x;
    if (COUNT && mine && gl == 0) a.counts[t] = static_cast<uint32_t>(__popcll(gm));
  }
}

// Splits the tuples by the length of their shortest list into the buckets'
// lists; lists[k] has room for every tuple.  bound, if set, receives each
// tuple's driver length (0 for a tuple with an id outside its list set).
// A block takes kSplitRounds x kBlock consecutive tuples at a time and
// reserves each bucket's positions for all of them with one atomic (wave
// scans of the per-thread counts, then the block's waves in LDS): the four
// bucket counters are the only shared words, and one atomic per wave and
// bucket kept them busy for most of the kernel.
constexpr int kSplitRounds = 8;

template <int N>
__global__ __launch_bounds__(kBlock) void k_mi_split(MiArgs a, uint32_t *lists, ull *norder,
                                                    uint32_t *bound) {
  __shared__ uint32_t s_wave[kWavesPerBlock][kCounters];
  __shared__ ull s_base[kCounters];
  const int lane = lane_id(), wid = wave_id();
  constexpr uint64_t kChunk = static_cast<uint64_t>(kBlock) * kSplitRounds;
  // block-uniform loop: every thread reaches the barriers
  for (uint64_t c0 = blockIdx.x * kChunk; c0 < a.T; c0 += gridDim.x * kChunk) {
    int kk[kSplitRounds];
    uint32_t cnt[kCounters] = {};
#pragma unroll
    for (int r = 0; r < kSplitRounds; ++r) {
      const uint64_t t = c0 + static_cast<uint64_t>(r) * kBlock + threadIdx.x;
      kk[r] = -1;
      if (t < a.T) {
        uint64_t b[N];
        uint32_t n[N];
        int drv;
        uint64_t best;
        if (!tuple_lists<N>(a, t, b, n, drv, best)) best = 0;
        int k = best > kGiantLen ? kGiant : best > kLargeLen ? kLarge : kLong;
        for (int q = kBuckets - 2; q >= 0; --q)
          if (best <= bucket_len(q)) k = q;
        if (bound) bound[t] = static_cast<uint32_t>(best);  // a list holds < 2^32 ids
        kk[r] = k;
#pragma unroll
        for (int q = 0; q < kCounters; ++q) cnt[q] += k == q;
      }
    }
    uint32_t inc[kCounters];
#pragma unroll
    for (int q = 0; q < kCounters; ++q) {
      inc[q] = cnt[q];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc[q], d);
        if (lane >= d) inc[q] += y;
      }
      if (lane == 63) s_wave[wid][q] = inc[q];
    }
    __syncthreads();
    if (threadIdx.x < kCounters) {
      const int q = threadIdx.x;
      uint32_t tot = 0;
      for (int w = 0; w < kWavesPerBlock; ++w) {
        const uint32_t x = s_wave[w][q];
        s_wave[w][q] = tot;
        tot += x;
      }
      s_base[q] = tot ? atomicAdd(&norder[q], static_cast<ull>(tot)) : 0;
    }
    __syncthreads();
    uint64_t pos[kCounters];
#pragma unroll
    for (int q = 0; q < kCounters; ++q) pos[q] = s_base[q] + s_wave[wid][q] + inc[q] - cnt[q];
#pragma unroll
    for (int r = 0; r < kSplitRounds; ++r) {
      const uint32_t t = static_cast<uint32_t>(c0 + static_cast<uint64_t>(r) * kBlock +
                                               threadIdx.x);  // T < 2^32 - 1
      if (kk[r] < 0) continue;  // past T
#pragma unroll
      for (int q = 0; q < kBuckets; ++q)
        if (kk[r] == q) lists[q * a.T + pos[q]++] = t;
      // giants from the back of the long list (long + giants <= T), large
      // tuples from the back of the 32-entry bucket's (that bucket + large <= T)
      if (kk[r] == kGiant) lists[kLong * a.T + (a.T - 1 - pos[kGiant]++)] = t;
      if (kk[r] == kLarge) lists[(kLong - 1) * a.T + (a.T - 1 - pos[kLarge]++)] = t;
    }
    __syncthreads();  // s_wave / s_base are rewritten by the next chunk
  }
}

template <int MODE, int N>
static void launch_fn(uint32_t grid, hipStream_t s, const MiArgs &a) {
  hipLaunchKernelGGL((k_mintersect<MODE, N>), dim3(grid), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL((k_mintersect_short<MODE, N, 0>), dim3(grid), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL((k_mintersect_short<MODE, N, 1>), dim3(grid), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL((k_mintersect_short<MODE, N, 2>), dim3(grid), dim3(kBlock), 0, s, a);
}

template <int N>
static void launch_n(int mode, uint32_t grid, hipStream_t s, const MiArgs &a) {
  if (mode == kMiFill) launch_fn<kMiFill, N>(grid, s, a);
  else if (mode == kMiStage) launch_fn<kMiStage, N>(grid, s, a);
  else launch_fn<kMiCount, N>(grid, s, a);
}

template <int N>
static void launch_split(uint32_t grid, hipStream_t s, const MiArgs &a, uint32_t *lists,
                         ull *norder, uint32_t *bound) {
  hipLaunchKernelGGL((k_mi_split<N>), dim3(grid), dim3(kBlock), 0, s, a, lists, norder, bound);
}

// arity as a template parameter keeps each variant's per-list state in
// registers
static void launch_mintersect(int mode, uint32_t arity, uint32_t grid, hipStream_t s,
                              const MiArgs &a) {
  switch (arity) {
    case 1: launch_n<1>(mode, grid, s, a); break;
    case 2: launch_n<2>(mode, grid, s, a); break;
    case 3: launch_n<3>(mode, grid, s, a); break;
    case 4: launch_n<4>(mode, grid, s, a); break;
    case 5: launch_n<5>(mode, grid, s, a); break;
    case 6: launch_n<6>(mode, grid, s, a); break;
    case 7: launch_n<7>(mode, grid, s, a); break;
    default: launch_n<8>(mode, grid, s, a); break;
  }
}

static void launch_mi_split(uint32_t arity, uint32_t grid, hipStream_t s, const MiArgs &a,
                            uint32_t *lists, ull *norder, uint32_t *bound) {
  switch (arity) {
    case 1: launch_split<1>(grid, s, a, lists, norder, bound); break;
    case 2: launch_split<2>(grid, s, a, lists, norder, bound); break;
    case 3: launch_split<3>(grid, s, a, lists, norder, bound); break;
    case 4: launch_split<4>(grid, s, a, lists, norder, bound); break;
    case 5: launch_split<5>(grid, s, a, lists, norder, bound); break;
    case 6: launch_split<6>(grid, s, a, lists, norder, bound); break;
    case 7: launch_split<7>(grid, s, a, lists, norder, bound); break;
    default: launch_split<8>(grid, s, a, lists, norder, bound); break;
  }
}

// Moves each tuple's staged run stg[soff[t] .. + counts[t]) to
// lists[off[t] ..].  One thread per tuple copies a short run itself; the
// wave copies the long runs of its lanes together, 64 entries per step.
constexpr uint32_t kCompactSerial = 16;

__global__ __launch_bounds__(kBlock) void k_mi_compact(const uint32_t *__restrict__ counts,
                                                      const uint64_t *__restrict__ soff,
                                                      const uint64_t *__restrict__ off,
                                                      const uint32_t *__restrict__ stg,
                                                      uint32_t *__restrict__ lists, uint64_t T) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t lim = (T + 63) / 64 * 64;  // whole waves stay in the loop
  for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < lim;
       t += stride) {
    const bool valid = t < T;
    const uint32_t n = valid ? counts[t] : 0;
    const uint64_t src = n ? soff[t] : 0, dst = n ? off[t] : 0;
    const bool big = n > kCompactSerial;
    if (!big)
      for (uint32_t i = 0; i < n; ++i) lists[dst + i] = stg[src + i];
    for (uint64_t m = __ballot(big); m; m &= m - 1) {
      const int l = __ffsll(static_cast<long long>(m)) - 1;
      const uint32_t nl = __shfl(n, l);
      const uint64_t sl = shfl64(src, l), dl = shfl64(dst, l);
      for (uint32_t i = lane_id(); i < nl; i += 64) lists[dl + i] = stg[sl + i];
    }
  }
}

// fence[f] = adj[64 f]
__global__ void k_fence(const uint32_t *__restrict__ adj, uint64_t E, uint32_t *__restrict__ fence) {
  const uint64_t nf = (E + 63) / 64;
  for (uint64_t f = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; f < nf;
       f += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    fence[f] = adj[f * 64];
}

// starts: bit i set if entry i begins a list (of a vertex u >= 1; entry 0
// needs no bit), one thread per vertex
__global__ __launch_bounds__(kBlock) void k_list_starts(const uint64_t *__restrict__ off,
                                                        uint32_t V, uint64_t E,
                                                        uint64_t *starts) {
  for (uint64_t u = 1 + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u < V;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = off[u];
    if (i < E) atomicOr(reinterpret_cast<ull *>(&starts[i >> 6]), 1ull << (i & 63));
  }
}

// *bad = 1 if two neighbouring entries of one list are not strictly
// ascending; one thread per entry (hub lists spread over the whole grid)
__global__ __launch_bounds__(kBlock) void k_check_sets(const uint32_t *__restrict__ adj,
                                                       uint64_t E,
                                                       const uint64_t *__restrict__ starts,
                                                       int *bad) {
  bool ok = true;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i + 1 < E;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (adj[i] >= adj[i + 1] && !((starts[(i + 1) >> 6] >> ((i + 1) & 63)) & 1)) ok = false;
  }
  if (!ok) *bad = 1;
}

// Builds the fence of a CSR and checks, once, that its lists are sets: the
// searches need strictly ascending lists (a repeated id in the driving list
// would be emitted twice, where the reference's bitmap AND yields it once).
// A build that fails part-way leaves c.sets at -1 and is redone next time.
vs_status build_fence(vs_graph_s *g, DevCSR &c) {
  if (!c.fence || c.sets < 0) {
    const uint64_t nf = (c.E + 63) / 64;
    void *p = nullptr, *flag = nullptr;
    VS_HIP(hipMalloc(&flag, sizeof(int)));
    struct Free { void *p; ~Free() { (void)hipFree(p); } } fr{flag};
    if (!c.fence) {
      VS_HIP(hipMalloc(&p, std::max<uint64_t>(nf, 1) * 4));
      c.fence = static_cast<uint32_t *>(p);
    }
    VS_HIP(hipMemsetAsync(flag, 0, sizeof(int), g->stream));
    void *starts = nullptr;
    const size_t sbytes = ((c.E + 63) / 64) * 8;
    if (nf) {
      hipLaunchKernelGGL(k_fence, dim3(grid_for((nf + 63) / 64, g->num_cus)), dim3(kBlock), 0,
                         g->stream, c.adj, c.E, c.fence);
      VS_HIP(hipGetLastError());
      VS_TRY(g->pool.get(sbytes, true, g->stream, &starts));
      hipLaunchKernelGGL(k_list_starts, dim3(grid_for((g->V + 63) / 64, g->num_cus)), dim3(kBlock),
                         0, g->stream, c.off, g->V, c.E, static_cast<uint64_t *>(starts));
      hipLaunchKernelGGL(k_check_sets, dim3(grid_for((c.E + 63) / 64, g->num_cus)), dim3(kBlock),
                         0, g->stream, c.adj, c.E, static_cast<const uint64_t *>(starts),
                         static_cast<int *>(flag));
      VS_HIP(hipGetLastError());
    }
    int bad = 0;
    VS_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    // the intersection may run on another graph's stream
    VS_HIP(hipStreamSynchronize(g->stream));
    if (starts) g->pool.put(starts, sbytes, false);
    c.sets = bad ? 0 : 1;
  }
  return VS_OK;
}

static vs_status ensure_fence(vs_graph_s *g, DevCSR &c) {
  VS_TRY(build_fence(g, c));
  if (c.sets != 1) {
    set_error("vs_mintersect: an adjacency list of a dependency repeats or misorders ids; "
              "intersections need ascending sets (vs_store_open builds them from a store)");
    return VS_ERR_INVALID;
  }
  return VS_OK;
}

__global__ void k_last_offset(const uint32_t *counts, uint64_t *offsets, uint64_t T) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    offsets[T] = T ? offsets[T - 1] + counts[T - 1] : 0;
}

// bytes of the distinct adjacency lists touched (4 * deg + 16 each)
__global__ void k_mark(const uint32_t *tuples, uint64_t n, uint32_t arity,
                       uint32_t V0, uint64_t *bm) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t x = tuples[i];
    if (x < V0) atomicOr(reinterpret_cast<ull *>(&bm[x >> 6]), 1ull << (x & 63));
  }
}

__global__ void k_marked_bytes(const uint64_t *bm, uint32_t V, const uint64_t *off,
                               ull *bytes) {
  uint64_t acc = 0;
  const uint32_t nvb = (V + 63) / 64;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nvb;
       w += gridDim.x * blockDim.x) {
    for (uint64_t m = bm[w]; m; m &= m - 1) {
      const uint32_t v = w * 64 + (__ffsll(static_cast<long long>(m)) - 1);
      acc += 4 * (off[v + 1] - off[v]) + 16;
    }
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(bytes, static_cast<ull>(acc));
}

// offsets[0..T) = exclusive scan of counts, offsets[T] = their sum, also
// returned in *sum (the stream is synchronised).
static vs_status scan_counts(vs_graph_s *g, hipStream_t s, const uint32_t *counts,
                             uint64_t *offsets, uint64_t T, uint64_t *sum) {
  size_t tmp_bytes = 0;
  VS_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, counts, offsets, static_cast<uint64_t>(0),
                                 static_cast<size_t>(T), rocprim::plus<uint64_t>(), s));
  void *tmp = nullptr;
  VS_TRY(g->pool.get(tmp_bytes, false, s, &tmp));
  hipError_t e = rocprim::exclusive_scan(tmp, tmp_bytes, counts, offsets,
                                         static_cast<uint64_t>(0), static_cast<size_t>(T),
                                         rocprim::plus<uint64_t>(), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_last_offset, dim3(1), dim3(64), 0, s, counts, offsets, T);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(sum, offsets + T, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  g->pool.put(tmp, tmp_bytes, false);
  if (e != hipSuccess) {
    set_error("vs_mintersect scan: %s", hipGetErrorString(e));
    return VS_ERR_HIP;
  }
  return VS_OK;
}

// Largest staging buffer (bytes) one call may take for its single pass:
// VS_MI_STAGE_MAX_MB from the environment, 4096 MB by default; 0 always
// takes two passes.
static uint64_t stage_limit_bytes() {
  const char *e = getenv("VS_MI_STAGE_MAX_MB");
  if (!e || !*e) return 4096ull << 20;
  char *end = nullptr;
  const unsigned long long mb = strtoull(e, &end, 10);
  if (end == e) return 4096ull << 20;
  return static_cast<uint64_t>(mb) << 20;
}

vs_status mintersect_run(const vs_graph_t *deps, uint32_t arity, const uint32_t *d_tuples,
                         uint64_t T, uint32_t *d_counts, uint64_t *d_offsets,
                         const ListsFor *lists_for, uint64_t *total, vs_mintersect_stats *st) {
  if (!deps || arity == 0 || arity > kMaxArity) {
    set_error("vs_mintersect: arity %u not in 1..%d", arity, kMaxArity);
    return VS_ERR_INVALID;
  }
  for (uint32_t d = 0; d < arity; ++d) {
    if (!deps[d]) { set_error("vs_mintersect: null dependency %u", d); return VS_ERR_INVALID; }
    if (deps[d]->device != deps[0]->device) {
      set_error("vs_mintersect: dependencies on different devices");
      return VS_ERR_INVALID;
    }
  }
  if (st) std::memset(st, 0, sizeof(*st));
  if (total) *total = 0;
  if (T == 0) {
    if (d_offsets) {
      const uint64_t z = 0;
      VS_HIP(hipMemcpy(d_offsets, &z, 8, hipMemcpyHostToDevice));
    }
    return VS_OK;
  }
  if (!d_tuples || !d_counts) { set_error("vs_mintersect: null tuples/counts"); return VS_ERR_INVALID; }
  vs_graph_s *g = deps[0];
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  MiArgs a{};
  a.arity = arity;
  a.T = T;
  a.tuples = d_tuples;
  for (uint32_t d = 0; d < arity; ++d) {
    VS_TRY(ensure_fence(deps[d], deps[d]->fwd));
    a.off[d] = deps[d]->fwd.off;
    a.adj[d] = deps[d]->fwd.adj;
    a.fence[d] = deps[d]->fwd.fence;
    a.V[d] = deps[d]->V;
  }
  a.counts = d_counts;
  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
  if (T >= 0xFFFFFFFFull) {
    set_error("vs_mintersect: %llu tuples in one call (at most 2^32 - 2)",
              static_cast<unsigned long long>(T));
    return VS_ERR_INVALID;
  }
  const uint32_t grid = grid_for(T, g->num_cus);
  // pool buffers of this call, returned on every exit
  struct Back {
    Pool &p;
    void *ptr[4] = {};
    size_t bytes[4] = {};
    int n = 0;
    void add(void *q, size_t b) { ptr[n] = q; bytes[n++] = b; }
    ~Back() { for (int i = 0; i < n; ++i) p.put(ptr[i], bytes[i], false); }
  } back{g->pool};
  // tuple lists by driver length: room for T ids in each bucket, counters
  void *lp = nullptr, *np = nullptr;
  const size_t lbytes = kBuckets * T * 4, nbytes = kCounters * 8;
  VS_TRY(g->pool.get(lbytes, false, s, &lp));
  back.add(lp, lbytes);
  VS_TRY(g->pool.get(nbytes, false, s, &np));
  back.add(np, nbytes);
  VS_HIP(hipMemsetAsync(np, 0, nbytes, s));
  for (int k = 0; k < kBuckets; ++k) a.order[k] = static_cast<const uint32_t *>(lp) + k * T;
  a.norder = static_cast<const ull *>(np);
  const uint64_t stage_max = stage_limit_bytes();
  const bool try_stage = d_offsets && lists_for && stage_max > 0;
  VS_HIP(hipEventRecord(e0, s));
  // with staging in view the split also writes each tuple's bound into counts
  const uint64_t split_chunks = (T + kBlock * kSplitRounds - 1) / (kBlock * kSplitRounds);
  launch_mi_split(arity, grid_for(split_chunks * kWavesPerBlock, g->num_cus), s, a,
                  static_cast<uint32_t *>(lp), static_cast<ull *>(np),
                  try_stage ? d_counts : nullptr);
  VS_HIP(hipGetLastError());
  uint64_t tot = 0;
  bool staged = false;
  if (try_stage) {
    void *so = nullptr;
    const size_t sobytes = (T + 1) * 8;
    VS_TRY(g->pool.get(sobytes, false, s, &so));
    back.add(so, sobytes);
    uint64_t *soff = static_cast<uint64_t *>(so);
    uint64_t bound = 0;
    VS_TRY(scan_counts(g, s, d_counts, soff, T, &bound));
    // staging sizes rounded up to powers of two, so calls of similar sizes
    // reuse one pooled buffer
    const bool fits = bound <= stage_max / 4;
    size_t sbytes = size_t{1} << 20;
    while (fits && sbytes < bound * 4) sbytes <<= 1;
    void *stg = nullptr;
    if (fits && g->pool.get(sbytes, false, s, &stg) == VS_OK) {
      back.add(stg, sbytes);
      a.offsets = soff;
      a.lists = static_cast<uint32_t *>(stg);
      launch_mintersect(kMiStage, arity, grid, s, a);
      VS_HIP(hipGetLastError());
      VS_TRY(scan_counts(g, s, d_counts, d_offsets, T, &tot));
      if (total) *total = tot;
      uint32_t *d_lists = nullptr;
      VS_TRY((*lists_for)(tot, &d_lists));
      if (tot && d_lists) {  // a provider may decline the lists (nullptr)
        hipLaunchKernelGGL(k_mi_compact, dim3(grid_for((T + 63) / 64, g->num_cus)),
                           dim3(kBlock), 0, s, d_counts, soff, d_offsets,
                           static_cast<const uint32_t *>(stg), d_lists, T);
        VS_HIP(hipGetLastError());
      }
      staged = true;
    } else {
      clear_error();  // no room to stage: two passes
    }
  }
  if (!staged) {
    launch_mintersect(kMiCount, arity, grid, s, a);
    VS_HIP(hipGetLastError());
    if (d_offsets) {
      VS_TRY(scan_counts(g, s, d_counts, d_offsets, T, &tot));
      if (total) *total = tot;
      if (lists_for) {
        uint32_t *d_lists = nullptr;
        VS_TRY((*lists_for)(tot, &d_lists));
        if (tot && d_lists) {
          a.offsets = d_offsets;
          a.lists = d_lists;
          launch_mintersect(kMiFill, arity, grid, s, a);
          VS_HIP(hipGetLastError());
        }
      }
    }
  }
  VS_HIP(hipEventRecord(e1, s));
  VS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  VS_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (!d_offsets) {
    // count-only: total from a host reduction of the counts
    std::vector<uint32_t> c(T);
    VS_HIP(hipMemcpy(c.data(), d_counts, T * 4, hipMemcpyDeviceToHost));
    for (uint32_t x : c) tot += x;
    if (total) *total = tot;
  }
  if (st) {
    st->ms = ms;
    st->total = tot;
    // compulsory bytes: distinct lists touched + tuples + counts + offsets + lists
    uint64_t b = 4ull * arity * T + 4 * T + (d_offsets ? 8 * (T + 1) : 0) +
                 (lists_for ? 4 * tot : 0);
    for (uint32_t d = 0; d < arity; ++d) {
      bool seen = false;
      for (uint32_t q = 0; q < d; ++q) seen |= deps[q] == deps[d];
      if (seen) continue;
      // one bitmap per distinct dependency graph over all columns bound to it
      void *bm = nullptr, *cnt = nullptr;
      const size_t bm_bytes = static_cast<size_t>((deps[d]->V + 63) / 64) * 8;
      VS_TRY(g->pool.get(bm_bytes, true, s, &bm));
      VS_TRY(g->pool.get(8, true, s, &cnt));
      // mark all columns bound to this dependency
      std::vector<uint32_t> cols;
      for (uint32_t q = 0; q < arity; ++q) if (deps[q] == deps[d]) cols.push_back(q);
      if (cols.size() == arity) {
        hipLaunchKernelGGL(k_mark, dim3(1024), dim3(256), 0, s, d_tuples, T * arity, arity,
                           deps[d]->V, static_cast<uint64_t *>(bm));
      } else {
        std::vector<uint32_t> h(T * arity), sel;
        VS_HIP(hipMemcpy(h.data(), d_tuples, T * arity * 4, hipMemcpyDeviceToHost));
        sel.reserve(T * cols.size());
        for (uint64_t t = 0; t < T; ++t) for (uint32_t q : cols) sel.push_back(h[t * arity + q]);
        void *dsel = nullptr;
        VS_TRY(g->pool.get(sel.size() * 4, false, s, &dsel));
        VS_HIP(hipMemcpyAsync(dsel, sel.data(), sel.size() * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_mark, dim3(1024), dim3(256), 0, s,
                           static_cast<const uint32_t *>(dsel), static_cast<uint64_t>(sel.size()),
                           1u, deps[d]->V, static_cast<uint64_t *>(bm));
        VS_HIP(hipStreamSynchronize(s));
        g->pool.put(dsel, sel.size() * 4, false);
      }
      hipLaunchKernelGGL(k_marked_bytes, dim3(1024), dim3(256), 0, s,
                         static_cast<const uint64_t *>(bm), deps[d]->V, deps[d]->fwd.off,
                         static_cast<ull *>(cnt));
      VS_HIP(hipGetLastError());
      uint64_t hb = 0;
      VS_HIP(hipMemcpyAsync(&hb, cnt, 8, hipMemcpyDeviceToHost, s));
      VS_HIP(hipStreamSynchronize(s));
      g->pool.put(bm, bm_bytes, false);
      g->pool.put(cnt, 8, false);
      b += hb;
    }
    st->bytes_alg = b;
  }
  return VS_OK;
}

// ------------------------------------------------- allocation-free variant
// vs_mintersect_device_noalloc runs the two-pass intersection (split ->
// count -> scan -> fill) in a fixed device array instead of pool buffers,
// one copy per GPU, held until the process ends:
//   [0, 256)                       the bucket, giant and large counters (ull)
//   [256, kNaHeader)               per-block sums of the counts scan and
//                                  their total (ull, at most 1153 entries)
//   [kNaHeader, kNaScratchBytes)   the four bucket lists, T ids each
// Host code never passes the array's name to a kernel: it only stands for
// the device address, which hipGetSymbolAddress gives per device (with that
// device current) on the first call there.
constexpr size_t kNaHeader = 64 << 10;
constexpr size_t kNaScratchBytes =
    kNaHeader + static_cast<size_t>(kBuckets) * VS_NOALLOC_MAX_TUPLES * 4;
constexpr uint32_t kNaScanPer = 16;                         // counts per thread
constexpr uint32_t kNaScanChunk = kBlock * kNaScanPer;      // counts per block
constexpr int kNaMaxDevices = 64;
static_assert(256 + 8 * ((VS_NOALLOC_MAX_TUPLES + kNaScanChunk - 1) / kNaScanChunk + 1) <=
                  kNaHeader,
              "scan partials outgrow the scratch header");

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

static std::mutex g_na_mu;
static unsigned char *g_na_ptr[kNaMaxDevices];

// Device address of this GPU's scratch (the calling thread has it current).
static vs_status na_scratch(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kNaMaxDevices) {
    set_error("vs_mintersect_device_noalloc: device %d beyond %d", dev, kNaMaxDevices);
    return VS_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error("vs_mintersect_device_noalloc: scratch address on device %d: %s", dev,
                hipGetErrorString(e));
      return VS_ERR_HIP;
    }
    g_na_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_na_ptr[dev];
  return VS_OK;
}

// Zeroes n words (a fill kernel, so a captured launch sequence holds no
// memset node).
__global__ void k_na_zero(ull *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0;
}

// part[b] = sum of counts[b * kNaScanChunk .. + kNaScanChunk)
__global__ __launch_bounds__(kBlock) void k_na_block_sums(const uint32_t *__restrict__ counts,
                                                         uint64_t T, ull *__restrict__ part) {
  __shared__ ull s_w[kWavesPerBlock];
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kNaScanChunk;
  ull acc = 0;
#pragma unroll
  for (uint32_t i = 0; i < kNaScanPer; ++i) {
    const uint64_t k = b0 + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
    if (k < T) acc += counts[k];
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) s_w[wave_id()] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    ull t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += s_w[w];
    part[blockIdx.x] = t;
  }
}

// Exclusive scan of part[0 .. nb) in place and part[nb] = the total; one wave.
__global__ __launch_bounds__(64) void k_na_scan_part(ull *part, uint32_t nb) {
  const int lane = lane_id();
  ull base = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
    const uint32_t i = b0 + lane;
    const ull x = i < nb ? part[i] : 0;
    ull inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const ull y = shfl_up64(inc, d);
      if (lane >= d) inc += y;
    }
    if (i < nb) part[i] = base + inc - x;
    base += shfl64(inc, 63);
  }
  if (lane == 0) part[nb] = base;
}

// offsets[k] = exclusive prefix sum of counts; offsets[T] = the total.  A
// thread scans kNaScanPer consecutive counts of its block's chunk.
__global__ __launch_bounds__(kBlock) void k_na_offsets(const uint32_t *__restrict__ counts,
                                                      uint64_t T, const ull *__restrict__ part,
                                                      uint32_t nb, uint64_t *__restrict__ offsets) {
  __shared__ ull s_w[kWavesPerBlock];
  const int lane = lane_id(), wid = wave_id();
  const uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * kNaScanChunk +
                      static_cast<uint64_t>(threadIdx.x) * kNaScanPer;
  uint32_t c[kNaScanPer];
  ull sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < kNaScanPer; ++i) {
    c[i] = k0 + i < T ? counts[k0 + i] : 0u;
    sum += c[i];
  }
  ull inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const ull y = shfl_up64(inc, d);
    if (lane >= d) inc += y;
  }
  if (lane == 63) s_w[wid] = inc;
  __syncthreads();
  ull run = part[blockIdx.x] + inc - sum;
  for (int w = 0; w < wid; ++w) run += s_w[w];
#pragma unroll
  for (uint32_t i = 0; i < kNaScanPer; ++i) {
    if (k0 + i < T) offsets[k0 + i] = run;
    run += c[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offsets[T] = part[nb];
}

static vs_status mintersect_noalloc(const vs_graph_t *deps, uint32_t arity,
                                    const uint32_t *d_tuples, uint64_t T, uint32_t *d_counts,
                                    uint64_t *d_offsets, uint32_t *d_lists, uint64_t list_cap,
                                    uint64_t *total) {
  const char *fn = "vs_mintersect_device_noalloc";
  if (!deps || arity == 0 || arity > kMaxArity) {
    set_error("%s: arity %u not in 1..%d", fn, arity, kMaxArity);
    return VS_ERR_INVALID;
  }
  for (uint32_t d = 0; d < arity; ++d) {
    if (!deps[d]) { set_error("%s: null dependency %u", fn, d); return VS_ERR_INVALID; }
    if (deps[d]->device != deps[0]->device) {
      set_error("%s: dependencies on different devices", fn);
      return VS_ERR_INVALID;
    }
  }
  if (!d_offsets) { set_error("%s: needs d_offsets (T + 1 entries)", fn); return VS_ERR_INVALID; }
  if (T > VS_NOALLOC_MAX_TUPLES) {
    set_error("%s: %llu tuples exceed the scratch limit of %u per call", fn,
              static_cast<unsigned long long>(T), VS_NOALLOC_MAX_TUPLES);
    return VS_ERR_INVALID;
  }
  if (T > 0 && (!d_tuples || !d_counts)) {
    set_error("%s: null tuples/counts", fn);
    return VS_ERR_INVALID;
  }
  if (total) *total = 0;
  vs_graph_s *g = deps[0];
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  unsigned char *base = nullptr;
  VS_TRY(na_scratch(g->device, &base));
  ull *norder = reinterpret_cast<ull *>(base);
  ull *part = reinterpret_cast<ull *>(base + 256);
  uint32_t *buckets = reinterpret_cast<uint32_t *>(base + kNaHeader);
  MiArgs a{};
  a.arity = arity;
  a.T = T;
  a.tuples = d_tuples;
  for (uint32_t d = 0; d < arity; ++d) {
    const DevCSR &c = deps[d]->fwd;
    if (!c.fence || c.sets < 0) {
      set_error("%s: dependency %u has no intersection fence (building it when the graph "
                "was created failed; a vs_mintersect call builds it)", fn, d);
      return VS_ERR_INVALID;
    }
    if (c.sets != 1) {
      set_error("%s: an adjacency list of dependency %u repeats or misorders ids; "
                "intersections need ascending sets", fn, d);
      return VS_ERR_INVALID;
    }
    a.off[d] = c.off;
    a.adj[d] = c.adj;
    a.fence[d] = c.fence;
    a.V[d] = deps[d]->V;
  }
  const uint32_t nb = static_cast<uint32_t>((T + kNaScanChunk - 1) / kNaScanChunk);
  if (T == 0) {  // offsets[0] = 0
    hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<ull *>(d_offsets), 1u);
    VS_HIP(hipGetLastError());
    VS_HIP(hipStreamSynchronize(s));
    return VS_OK;
  }
  a.counts = d_counts;
  for (int k = 0; k < kBuckets; ++k) a.order[k] = buckets + static_cast<size_t>(k) * T;
  a.norder = norder;
  hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, s, norder, static_cast<uint32_t>(kCounters));
  const uint32_t grid = grid_for(T, g->num_cus);
  const uint64_t split_chunks = (T + kBlock * kSplitRounds - 1) / (kBlock * kSplitRounds);
  launch_mi_split(arity, grid_for(split_chunks * kWavesPerBlock, g->num_cus), s, a, buckets,
                  norder, nullptr);
  launch_mintersect(kMiCount, arity, grid, s, a);
  hipLaunchKernelGGL(k_na_block_sums, dim3(nb), dim3(kBlock), 0, s,
                     static_cast<const uint32_t *>(d_counts), T, part);
  hipLaunchKernelGGL(k_na_scan_part, dim3(1), dim3(64), 0, s, part, nb);
  hipLaunchKernelGGL(k_na_offsets, dim3(nb), dim3(kBlock), 0, s,
                     static_cast<const uint32_t *>(d_counts), T,
                     static_cast<const ull *>(part), nb, d_offsets);
  VS_HIP(hipGetLastError());
  uint64_t tot = 0;
  VS_HIP(hipMemcpyAsync(&tot, d_offsets + T, 8, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  if (total) *total = tot;
  if (!d_lists) return VS_OK;
  if (tot > list_cap) {
    set_error("%s: list capacity %llu < total %llu", fn, static_cast<unsigned long long>(list_cap),
              static_cast<unsigned long long>(tot));
    return VS_ERR_CAPACITY;
  }
  a.offsets = d_offsets;
  a.lists = d_lists;
  launch_mintersect(kMiFill, arity, grid, s, a);
  VS_HIP(hipGetLastError());
  VS_HIP(hipStreamSynchronize(s));
  return VS_OK;
}

vs_status mintersect_impl(const vs_graph_t *deps, uint32_t arity, const uint32_t *d_tuples,
                          uint64_t T, uint32_t *d_counts, uint64_t *d_offsets,
                          uint32_t *d_lists, uint64_t list_cap, uint64_t *total,
                          vs_mintersect_stats *st) {
  const ListsFor into = [&](uint64_t tot, uint32_t **out) -> vs_status {
    *out = d_lists;
    if (tot > list_cap) {
      set_error("vs_mintersect: list capacity %llu < total %llu",
                static_cast<unsigned long long>(list_cap), static_cast<unsigned long long>(tot));
      return VS_ERR_CAPACITY;
    }
    return VS_OK;
  };
  return mintersect_run(deps, arity, d_tuples, T, d_counts, d_offsets, d_lists ? &into : nullptr,
                        total, st);
}

} // namespace vs

using namespace vs;

extern "C" {

vs_status vs_mintersect_device_noalloc(const vs_graph_t *deps, uint32_t arity,
                                       const uint32_t *d_tuples, uint64_t T,
                                       uint32_t *d_counts, uint64_t *d_offsets,
                                       uint32_t *d_lists, uint64_t list_cap,
                                       uint64_t *total) try {
  clear_error();
  return mintersect_noalloc(deps, arity, d_tuples, T, d_counts, d_offsets, d_lists, list_cap,
                            total);
} VS_CATCH("vs_mintersect_device_noalloc")

vs_status vs_mintersect_device(const vs_graph_t *deps, uint32_t arity,
                               const uint32_t *d_tuples, uint64_t T,
                               uint32_t *d_counts, uint64_t *d_offsets,
                               uint32_t *d_lists, uint64_t list_cap,
                               uint64_t *total, vs_mintersect_stats *stats) try {
  clear_error();
  return mintersect_impl(deps, arity, d_tuples, T, d_counts, d_offsets, d_lists,
                         list_cap, total, stats);
} VS_CATCH("vs_mintersect_device")

vs_status vs_mintersect(const vs_graph_t *deps, uint32_t arity,
                        const uint32_t *tuples, uint64_t T, uint32_t *counts,
                        uint64_t *offsets, uint32_t *lists, uint64_t list_cap,
                        uint64_t *total, vs_mintersect_stats *stats) try {
  clear_error();
  if (!deps || arity == 0 || arity > kMaxArity || !deps[0]) {
    set_error("vs_mintersect: bad dependencies / arity %u", arity);
    return VS_ERR_INVALID;
  }
  if (T > 0 && (!tuples || !counts)) {
    set_error("vs_mintersect: null tuples/counts");
    return VS_ERR_INVALID;
  }
  if (lists && !offsets) {
    set_error("vs_mintersect: lists need offsets");
    return VS_ERR_INVALID;
  }
  vs_graph_s *g = deps[0];
  VS_HIP(hipSetDevice(g->device));
  struct Bufs {
    vs_graph_s *g;
    void *t = nullptr, *c = nullptr, *o = nullptr, *l = nullptr;
    size_t tb = 0, cb = 0, ob = 0, lb = 0;
    ~Bufs() {
      g->pool.put(t, tb, false);
      g->pool.put(c, cb, false);
      g->pool.put(o, ob, false);
      g->pool.put(l, lb, false);
    }
  } b{g};
  if (T == 0) {
    if (total) *total = 0;
    if (offsets) offsets[0] = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return VS_OK;
  }
  b.tb = T * arity * 4;
  b.cb = T * 4;
  b.ob = offsets ? (T + 1) * 8 : 0;
  VS_TRY(g->pool.get(b.tb, false, g->stream, &b.t));
  VS_TRY(g->pool.get(b.cb, false, g->stream, &b.c));
  if (b.ob) VS_TRY(g->pool.get(b.ob, false, g->stream, &b.o));
  VS_HIP(hipMemcpyAsync(b.t, tuples, b.tb, hipMemcpyHostToDevice, g->stream));
  uint64_t tot = 0;
  // one intersection pass; the device list buffer is taken once the total
  // is known (counts and offsets are returned on a capacity error too)
  const ListsFor into = [&](uint64_t n, uint32_t **out) -> vs_status {
    *out = nullptr;
    if (n > list_cap) {
      set_error("vs_mintersect: list capacity %llu < total %llu",
                static_cast<unsigned long long>(list_cap), static_cast<unsigned long long>(n));
      return VS_ERR_CAPACITY;
    }
    b.lb = std::max<uint64_t>(n, 1) * 4;
    VS_TRY(g->pool.get(b.lb, false, g->stream, &b.l));
    *out = static_cast<uint32_t *>(b.l);
    return VS_OK;
  };
  const vs_status rs = mintersect_run(deps, arity, static_cast<uint32_t *>(b.t), T,
                                      static_cast<uint32_t *>(b.c),
                                      static_cast<uint64_t *>(b.o), lists ? &into : nullptr,
                                      &tot, stats);
  if (rs != VS_OK && rs != VS_ERR_CAPACITY) return rs;
  if (total) *total = tot;
  VS_HIP(hipMemcpy(counts, b.c, b.cb, hipMemcpyDeviceToHost));
  if (offsets) VS_HIP(hipMemcpy(offsets, b.o, b.ob, hipMemcpyDeviceToHost));
  if (rs != VS_OK) return rs;
  if (lists && tot) VS_HIP(hipMemcpy(lists, b.l, tot * 4, hipMemcpyDeviceToHost));
  return VS_OK;
} VS_CATCH("vs_mintersect")

} // extern "C"
