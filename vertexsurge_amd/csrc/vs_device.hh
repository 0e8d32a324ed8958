// Device-side internal declarations (hipcc only): graph handle, buffer pool,
// result handle, and wave-level helpers for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <vector>

#include "vs_internal.hh"

#define VS_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      vs::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                \
                    hipGetErrorString(e_));                                   \
      return VS_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define VS_TRY(call)                                                          \
  do {                                                                        \
    vs_status s_ = (call);                                                    \
    if (s_ != VS_OK) return s_;                                               \
  } while (0)

typedef unsigned long long ull;

namespace vs {

// Work item of the push kernel: vertex id in the low 32 bits, chunk of
// kPushChunk neighbours in the high 32 bits.  Hubs become many items so one
// wave never owns a whole hub list.
constexpr uint32_t kPushChunk = 256;
// Pull work item: destination vertex | chunk of kPullChunk in-neighbours.
// A vertex with more in-edges (a hub) spans several items on different waves;
// their partial ORs meet in the push accumulator and the epilogue finishes it.
constexpr uint32_t kPullChunk = 512;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = 64 * kWavesPerBlock;

// Device buffers of exactly-matching size are recycled per graph.  A buffer
// comes back "clean" (all zero) or "dirty"; get(zero=true) memsets only dirty
// ones.
class Pool {
 public:
  vs_status get(size_t bytes, bool zero, hipStream_t s, void **out);
  void put(void *p, size_t bytes, bool clean);
  void trim();
  // bytes of the cached (free) buffers
  size_t cached() const {
    size_t n = 0;
    for (const auto &kv : free_) n += kv.first;
    return n;
  }
  ~Pool() { trim(); }

 private:
  std::multimap<size_t, std::pair<void *, bool>> free_;
};

// One CSR resident in HBM.
struct DevCSR {
  uint64_t *off = nullptr;
  uint32_t *adj = nullptr;
  uint64_t E = 0;
  std::vector<uint64_t> hoff; // host copy of offsets (degrees for host stats)
  bool owned = false;
  // static pull work items over this CSR's lists (built on first pull)
  unsigned long long *pitems = nullptr;
  uint64_t npitems = 0;
  bool has_heavy = false;     // some list longer than kPullChunk
  // bitmap words (v / 64) holding a vertex with several pull items: the only
  // words a pull hop's epilogue has to visit
  uint32_t *hwords = nullptr;
  uint32_t nhwords = 0;
  // every 64th list entry (fence[f] = adj[64 f]), built on the first
  // intersection over this CSR
  uint32_t *fence = nullptr;
  int sets = -1;  // lists strictly ascending: 1 yes, 0 no, -1 not checked
  // Pull copy of the lists with the hub neighbours first (hubs = the
  // kPullHubs vertices that appear in the most lists, i.e. whose frontier
  // rows are gathered most often; most-listed first), the rest ascending,
  // and per-vertex hub counts.  Pull gathers hub rows with cached loads and the rest with
  // streaming loads, so the rows reused across destinations stay in L2.
  uint32_t *hadj = nullptr;
  uint32_t *hubcnt = nullptr;
  uint32_t nhubs = 0;  // hub count the lists were built for
};

// default hub count: 131072 rows of 512 B (tile width 64) = 64 MB, a hot set
// the 256 MB Infinity Cache can hold next to the streaming gathers
constexpr uint32_t kPullHubs = 131072;
// default word-mask threshold: pull gathers sparse rows through their word
// masks when the degree-weighted share of nonzero words is below this (%).
// On R-MAT 24 / 4096 sources the hop-3 frontier sits between 35 and 50: its
// dense pull took 41.4 ms and fetched 195 GB, the word-mask pull 47.1 ms and
// 244 GB; hop 2 (far below 35) stays on the word masks.
constexpr uint32_t kSparsePct = 35;
// default word-push budget, per mille of E: on R-MAT 24 / 4096 sources the
// hop-3 frontier's sparsest word reaches 2.6 % of E by push degree (and its
// lane keeps every destination's pull going through all of its in-edges),
// the next one 14 %
constexpr uint32_t kWPushPml = 40;

// Builds the pull items of a CSR once (vertex | chunk << 32 per kPullChunk
// in-edges, vertices without edges omitted).
// builds the pull items once, and the hub-first lists for `hubs` hubs
// (0 = kPullHubs), rebuilding them when the hub count changes
vs_status ensure_pull_items(DevCSR &c, uint32_t V, uint32_t hubs);

} // namespace vs

struct vs_sched_s;
namespace vs {
// drops one reference to a scheduler, freeing it with the last (vs_mlp.hip)
void sched_release(vs_sched_s *s);
}

struct vs_graph_s {
  int device = 0;
  int num_cus = 256;
  uint32_t V = 0;
  uint64_t E = 0;
  bool symmetric = false;
  vs::DevCSR fwd, rev, both;      // rev/both built on demand for directed graphs
  std::vector<uint32_t> hadj;     // host copy of adj (directed graphs only)
  hipStream_t stream = nullptr;
  vs::Pool pool;
  vs_tuning tuning{0, 16, 0, 0};
  vs_sched_s *sched = nullptr;
  int refs = 1;                   // graph + live results
  bool destroyed = false;
};

namespace vs {

// key[r] = sum over u in N(src[r]) of deg(u) in the push CSR (the source's
// 2-hop volume), computed on the device; S host ids in, S keys out.
vs_status source_volumes(vs_graph_s *g, const DevCSR *push, const uint32_t *src, size_t S,
                         uint64_t *key);
// (push CSR, pull CSR) for a direction: push follows edges forward from a
// frontier vertex, pull gathers over the edges that arrive at a vertex.
vs_status graph_csrs(vs_graph_s *g, int direction, const DevCSR **push,
                     const DevCSR **pull);
void graph_release(vs_graph_s *g);
// Builds a CSR's intersection fence (every 64th adjacency entry) and checks
// once whether its lists are strictly ascending sets (DevCSR::sets); done at
// graph creation for the forward CSR (vs_mintersect.hip).
vs_status build_fence(vs_graph_s *g, DevCSR &c);

struct Mat {
  uint64_t *d = nullptr;
  size_t bytes = 0;
  uint32_t distance = UINT32_MAX;
};

} // namespace vs

struct vs_result_s {
  vs_graph_s *g = nullptr;
  // S rows (sources); V columns: the graph's vertices, or the selected
  // columns of a column select (then col_map holds their vertex ids)
  uint32_t S = 0, V = 0, tw = 1, ntiles = 0, k_actual_max = 0;
  std::vector<vs::Mat> mats;
  std::vector<uint32_t> row_map;  // source vertex of each row
  std::vector<uint32_t> col_map;  // empty: column c is vertex c
  // Row order of the stored matrices.  Empty: stored row r is caller row r.
  // Otherwise the engine expanded the rows in an order of its own -- sorted
  // within each tile of 64 * tw rows by the source's 2-hop volume -- and
  // stored[r] is the stored row of caller row r (same tile).  Every output
  // indexed by row (row counts, digests, exports, get, the flat table, plain
  // TableConvert) is given in caller order; row-order-free outputs (counts,
  // column counts, row-OR, column select) read the stored matrices as they are.
  std::vector<uint32_t> stored;
};

namespace vs {

// ----------------------------------------------------------- device helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
// Wave index within the block.  readfirstlane tells the compiler the value is
// wave-uniform, so the work-item indices derived from it, and what is loaded
// through them, live in scalar registers.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  // number of set bits of m below this lane
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Orders LDS traffic between the lanes of one wave (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t x, int m) {
  const uint32_t lo = __shfl_xor(static_cast<uint32_t>(x), m);
  const uint32_t hi = __shfl_xor(static_cast<uint32_t>(x >> 32), m);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  const uint32_t lo = __shfl(static_cast<uint32_t>(x), src);
  const uint32_t hi = __shfl(static_cast<uint32_t>(x >> 32), src);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// OR over the 64/TW lane groups that share word index lane % TW.
template <int TW>
__device__ __forceinline__ uint64_t group_or(uint64_t x) {
#pragma unroll
  for (int m = TW; m < 64; m <<= 1) x |= shfl_xor64(x, m);
  return x;
}

__device__ __forceinline__ ull shfl_up64(ull x, int d) {
  const uint32_t lo = __shfl_up(static_cast<uint32_t>(x), d);
  const uint32_t hi = __shfl_up(static_cast<uint32_t>(x >> 32), d);
  return (static_cast<ull>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) x += shfl_xor64(x, m);
  return x;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) x |= shfl_xor64(x, m);
  return x;
}

// Bits of global word w that belong to real source rows (< S).
__device__ __forceinline__ uint64_t valid_mask(uint32_t w, uint32_t S) {
  const uint64_t r0 = static_cast<uint64_t>(w) * 64;
  if (r0 >= S) return 0;
  const uint64_t n = S - r0;
  return n >= 64 ? ~0ull : ((1ull << n) - 1);
}

template <int TW>
__device__ __forceinline__ uint64_t group_bits(uint64_t bal, int g) {
  if constexpr (TW == 64) return bal;
  else return (bal >> (g * TW)) & ((1ull << TW) - 1);
}

// index of the g-th (0-based) lowest set bit of m, or -1
__device__ __forceinline__ int nth_set_bit(uint64_t m, int g) {
  for (int q = 0; q < g && m; ++q) m &= m - 1;
  return m ? __ffsll(static_cast<long long>(m)) - 1 : -1;
}

__device__ __forceinline__ uint64_t drop_low_bits(uint64_t m, int n) {
  for (int q = 0; q < n && m; ++q) m &= m - 1;
  return m;
}

__host__ __device__ __forceinline__ uint64_t digest_hash(uint32_t c) {
  uint64_t z = 0x5EEDC0DEull + (static_cast<uint64_t>(c) + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint32_t pow2ceil(uint32_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

inline uint32_t grid_for(uint64_t waves, int num_cus, int blocks_per_cu = 8) {
  uint64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  const uint64_t cap = static_cast<uint64_t>(num_cus) * blocks_per_cu;
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  return static_cast<uint32_t>(blocks);
}

// Result-side kernels (vs_graph.hip).
vs_status mat_row_counts(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                         uint32_t tw, uint32_t ntiles, uint64_t *d_counts);
vs_status mat_row_digests(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                          uint32_t tw, uint32_t ntiles, uint64_t *d_digests);
// native [tile][col][tw] -> one bitmap of V bits per row at out + row*pitch
// (words nvb .. pitch-1 of each row are left as they are); with d_dest, stored
// row r goes to row d_dest[r] instead
vs_status mat_to_rowmajor(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                          uint32_t tw, uint64_t *out, uint64_t pitch,
                          const uint32_t *d_dest = nullptr);
// In place, within each tile of a [T][V][tw] matrix: word bit i of tile t
// (row t * 64 tw + i) takes the bit of row d_unpack[t * 64 tw + i] of the
// same tile.
vs_status mat_unpack_rows(vs_graph_s *g, uint64_t *M, uint32_t V, uint32_t T, uint32_t tw,
                          const uint16_t *d_unpack);

// Matrix i of a result in caller row order: the stored matrix itself when
// its rows are stored in that order, else a pooled copy with the rows put
// back (returned to the pool by the destructor).
class CallerOrder {
 public:
  CallerOrder() = default;
  CallerOrder(const CallerOrder &) = delete;
  CallerOrder &operator=(const CallerOrder &) = delete;
  ~CallerOrder();
  vs_status open(vs_result_s *r, uint32_t i);
  const uint64_t *d = nullptr;

 private:
  vs_graph_s *g_ = nullptr;
  void *tmp_ = nullptr;
  size_t bytes_ = 0;
};

// The flat pairs of matrix `which` of r (the union of all when UINT32_MAX)
// in (row, target) order into device buffers from the graph's pool (the
// caller returns them, `bytes` each); with want false only *n (vs_flat.hip).
vs_status flat_pairs_device(vs_result_s *r, uint32_t which, bool want, uint32_t **d_start,
                            uint32_t **d_target, size_t *bytes, uint64_t *n);
// Multi-way intersection over edge relations with every array in device
// memory (vs_mintersect_device, vs_mintersect.hip): lists into d_lists
// (capacity list_cap), or, in mintersect_run, into the buffer lists_for
// returns for the total once it is known (nullptr: counts only).
using ListsFor = std::function<vs_status(uint64_t, uint32_t **)>;
vs_status mintersect_run(vs_graph_s *const *deps, uint32_t arity, const uint32_t *d_tuples,
                         uint64_t T, uint32_t *d_counts, uint64_t *d_offsets,
                         const ListsFor *lists_for, uint64_t *total, vs_mintersect_stats *st);
vs_status mintersect_impl(vs_graph_s *const *deps, uint32_t arity, const uint32_t *d_tuples,
                          uint64_t T, uint32_t *d_counts, uint64_t *d_offsets,
                          uint32_t *d_lists, uint64_t list_cap, uint64_t *total,
                          vs_mintersect_stats *st);

// Device table of the stored -> caller row map of r (S uint32, pooled; the
// destructor returns it), or nullptr when the rows are in caller order.
class DestRows {
 public:
  DestRows() = default;
  DestRows(const DestRows &) = delete;
  DestRows &operator=(const DestRows &) = delete;
  ~DestRows();
  vs_status open(vs_result_s *r);
  const uint32_t *d = nullptr;

 private:
  vs_graph_s *g_ = nullptr;
  void *p_ = nullptr;
  size_t bytes_ = 0;
};

} // namespace vs
