// Bit-matrix tables on the device: column-major matrices and the operators
// that sit between VExpand and Mintersect in a pattern plan.
//
// Reference: DenseColMainMatrix (src/include/storage/bitmap.hh:262-326,
// src/objects/storage/bitmap.cpp:580-908), the table maps
// (src/include/execution/matrix_table.hh:4-125,
// src/objects/execution/bit_matrix_table.cpp:12-62), TableConvert and
// MatrixTableSelect (src/objects/operator/big_operators2.cpp:16-213),
// MatrixColumnCount (src/objects/operator/matrix_column_count.cpp:19-87),
// MatrixRowOrSum (src/objects/operator/matrix_row_or_sum.cpp:14-55) and the
// ColMain branch of Mintersect (src/objects/operator/mintersect.cpp:139-243).
//
// The reference transposes bit by bit (big_operators2.cpp:49-54) and ANDs
// whole columns per tuple on the CPU.  Here a transpose is a 64x64 ballot
// block per wave, a plain conversion is a word permutation, and the
// intersection streams columns through waves in fixed chunks (count pass ->
// scan -> fill pass, outputs in ascending row order), every pass HBM-bound.
#include <algorithm>
#include <functional>
#include <cstring>
#include <memory>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "vs_device.hh"

struct vs_colmat_s {
  vs_graph_s *g = nullptr;
  uint32_t rows = 0, cols = 0;
  uint64_t pitch = 0;  // words per column, align8(ceil(rows / 64))
  uint64_t *d = nullptr;
  size_t bytes = 0;
  std::vector<uint32_t> row_map, col_map;
  uint32_t *d_row_map = nullptr;
  size_t row_map_bytes = 0;
  // column lookup by id: (id << 32 | column) ascending, first column per id
  uint64_t *d_col_keys = nullptr;
  size_t col_keys_bytes = 0;
  uint32_t ncol_keys = 0;
};

namespace vs {

constexpr int kMaxColmatDeps = 8;
constexpr uint64_t kMcChunkWords = 4096;  // words of one column per wave item
constexpr uint64_t kMcMaxItems = 1ull << 22;
constexpr uint32_t kNoCol = 0xFFFFFFFFu;

inline uint64_t colmat_pitch(uint32_t rows) {
  return ((static_cast<uint64_t>(rows) + 63) / 64 + 7) / 8 * 8;
}

// ---------------------------------------------------------------- kernels

// Clears the bits of rows >= rows in every column (partial last word and
// the align8 padding words).
__global__ void k_cm_clear_pad(uint64_t *__restrict__ d, uint32_t cols, uint64_t pitch,
                               uint32_t rows) {
  const uint64_t used = (static_cast<uint64_t>(rows) + 63) / 64;
  const uint64_t tail = pitch - used + ((rows % 64) ? 1 : 0);
  const uint64_t n = static_cast<uint64_t>(cols) * tail;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t c = i / tail, k = i % tail;
    const uint64_t w = pitch - tail + k;
    uint64_t &x = d[c * pitch + w];
    if (w < used) x &= (1ull << (rows % 64)) - 1;
    else x = 0;
  }
}

// Native [tile][vertex][TW] -> ColMain with rows = source rows: column v is
// the W words of vertex v (across tiles), then zero padding to the pitch.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_cm_from_native(const uint64_t *__restrict__ M,
                                                           uint32_t V, uint32_t W,
                                                           uint64_t pitch,
                                                           uint64_t *__restrict__ out) {
  const uint64_t n = static_cast<uint64_t>(V) * pitch;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = i / pitch, w = i % pitch;
    out[i] = w < W ? M[((w / TW) * V + v) * TW + (w % TW)] : 0;
  }
}

// Native -> native with columns sel[0..n): out[t][k][j] = M[t][sel[k]][j].
template <int TW>
__global__ __launch_bounds__(kBlock) void k_select_native(const uint64_t *__restrict__ M,
                                                          uint32_t V, uint32_t ntiles,
                                                          const uint32_t *__restrict__ sel,
                                                          uint32_t n,
                                                          uint64_t *__restrict__ out) {
  const uint64_t total = static_cast<uint64_t>(ntiles) * n * TW;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t j = i % TW, tk = i / TW;
    const uint64_t t = tk / n, k = tk % n;
    out[i] = M[(t * V + sel[k]) * TW + j];
  }
}

// Column popcounts of a native matrix: lane j of a TW-lane group reads word j
// of one vertex in every tile; the group sums its lanes.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_colcount_native(const uint64_t *__restrict__ M,
                                                            uint32_t V, uint32_t ntiles,
                                                            uint32_t *__restrict__ counts) {
  constexpr int G = 64 / TW;
  const int lane = lane_id(), wid = wave_id();
  const int g = lane / TW, j = lane % TW;
  const uint64_t nitems = (static_cast<uint64_t>(V) + G - 1) / G;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves) {
    const uint64_t v = it * G + g;
    uint32_t c = 0;
    if (v < V)
      for (uint32_t t = 0; t < ntiles; ++t)
        c += __popcll(M[(static_cast<uint64_t>(t) * V + v) * TW + j]);
#pragma unroll
    for (int m = 1; m < TW; m <<= 1) c += __shfl_xor(c, m);
    if (j == 0 && v < V) counts[v] = c;
  }
}

// Column popcounts of a ColMain matrix: one wave per column.
__global__ __launch_bounds__(kBlock) void k_colcount_cm(const uint64_t *__restrict__ d,
                                                        uint32_t cols, uint64_t pitch,
                                                        uint64_t used,
                                                        uint32_t *__restrict__ counts) {
  const int lane = lane_id(), wid = wave_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t c = blockIdx.x * kWavesPerBlock + wid; c < cols; c += nwaves) {
    uint64_t s = 0;
    for (uint64_t w = lane; w < used; w += 64) s += __popcll(d[c * pitch + w]);
    s = wave_sum(s);
    if (lane == 0) counts[c] = static_cast<uint32_t>(s);
  }
}

// (count << 32 | id) sort keys; ids through the column map when given.
__global__ void k_count_keys(const uint32_t *__restrict__ counts,
                             const uint32_t *__restrict__ map, uint32_t n,
                             uint64_t *__restrict__ keys) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t id = map ? map[i] : static_cast<uint32_t>(i);
    keys[i] = (static_cast<uint64_t>(counts[i]) << 32) | id;
  }
}

struct McArgs {
  uint32_t nd;
  uint32_t rows;
  uint64_t pitch;
  uint64_t used;     // words holding rows
  uint64_t nchunk;   // chunks per column
  uint64_t T;
  const uint32_t *tuples;
  uint32_t *colidx;  // T * nd resolved columns (kNoCol = absent)
  const uint64_t *mat[kMaxColmatDeps];
  uint32_t cols[kMaxColmatDeps];
  const uint64_t *keys[kMaxColmatDeps];
  uint32_t nkeys[kMaxColmatDeps];
  const uint32_t *row_map;
};

// MatrixTable::find_col_idx (bit_matrix_table.cpp:48-57): through the
// column map's lookup (first column of the id), else the id itself when it
// is a column index.
__device__ __forceinline__ uint32_t find_col(const McArgs &a, uint32_t d, uint32_t id) {
  const uint64_t *k = a.keys[d];
  if (!k) return id < a.cols[d] ? id : kNoCol;
  uint32_t lo = 0, hi = a.nkeys[d];
  const uint64_t x = static_cast<uint64_t>(id) << 32;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (k[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < a.nkeys[d] && (k[lo] >> 32) == id) ? static_cast<uint32_t>(k[lo]) : kNoCol;
}

__global__ void k_mc_resolve(McArgs a, ull *found) {
  for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < a.T;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    bool ok = true;
    for (uint32_t d = 0; d < a.nd; ++d) {
      const uint32_t c = find_col(a, d, a.tuples[t * a.nd + d]);
      a.colidx[t * a.nd + d] = c;
      ok &= c != kNoCol;
    }
    if (ok) atomicAdd(found, 1ull);
  }
}

// word w of the AND of tuple t's columns (cols already resolved, all present)
__device__ __forceinline__ uint64_t and_word(const McArgs &a, const uint32_t *ci, uint64_t w) {
  uint64_t x = a.mat[0][static_cast<uint64_t>(ci[0]) * a.pitch + w];
  for (uint32_t d = 1; d < a.nd; ++d) x &= a.mat[d][static_cast<uint64_t>(ci[d]) * a.pitch + w];
  return x;
}

// Count pass: one wave per (tuple, chunk) item of the batch starting at t0.
__global__ __launch_bounds__(kBlock) void k_mc_count(McArgs a, uint64_t t0, uint64_t nitems,
                                                     uint32_t *__restrict__ part) {
  const int lane = lane_id(), wid = wave_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves) {
    const uint64_t t = t0 + it / a.nchunk, ch = it % a.nchunk;
    const uint32_t *ci = a.colidx + t * a.nd;
    bool ok = true;
    for (uint32_t d = 0; d < a.nd; ++d) ok &= ci[d] != kNoCol;
    uint64_t s = 0;
    if (ok) {
      const uint64_t w1 = min(a.used, (ch + 1) * kMcChunkWords);
      for (uint64_t w = ch * kMcChunkWords + lane; w < w1; w += 64) s += __popcll(and_word(a, ci, w));
    }
    s = wave_sum(s);
    if (lane == 0) part[it] = static_cast<uint32_t>(s);
  }
}

// counts[t] = sum of the tuple's chunk counts
__global__ void k_mc_sum(const uint32_t *__restrict__ part, uint64_t nchunk, uint64_t nt,
                         uint32_t *__restrict__ counts) {
  for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < nt;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t s = 0;
    for (uint64_t c = 0; c < nchunk; ++c) s += part[t * nchunk + c];
    counts[t] = s;
  }
}

// Fill pass: the item's set rows in ascending order, from its position
// base + pscan[item]; each 64-word step gives lanes their output slots with
// a wave prefix sum of the words' popcounts.
__global__ __launch_bounds__(kBlock) void k_mc_fill(McArgs a, uint64_t t0, uint64_t nitems,
                                                    const uint64_t *__restrict__ pscan,
                                                    uint64_t base, uint32_t *__restrict__ lists) {
  const int lane = lane_id(), wid = wave_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves) {
    const uint64_t t = t0 + it / a.nchunk, ch = it % a.nchunk;
    const uint32_t *ci = a.colidx + t * a.nd;
    bool ok = true;
    for (uint32_t d = 0; d < a.nd; ++d) ok &= ci[d] != kNoCol;
    if (!ok) continue;
    uint64_t pos = base + pscan[it];
    const uint64_t w0 = ch * kMcChunkWords, w1 = min(a.used, w0 + kMcChunkWords);
    for (uint64_t wb = w0; wb < w1; wb += 64) {
      const uint64_t w = wb + lane;
      uint64_t x = w < w1 ? and_word(a, ci, w) : 0;
      const uint32_t c = __popcll(x);
      uint32_t inc = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
      }
      uint64_t o = pos + inc - c;
      while (x) {
        const uint32_t b = __ffsll(static_cast<long long>(x)) - 1;
        x &= x - 1;
        const uint32_t r = static_cast<uint32_t>(w * 64 + b);
        lists[o++] = a.row_map ? a.row_map[r] : r;
      }
      pos += __shfl(inc, 63);
    }
  }
}

#define VS_TW_DISPATCH_T(tw, KERNEL, ...)                                    \
  switch (tw) {                                                              \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;               \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;               \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;               \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;               \
    case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;             \
    case 32: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break;             \
    default: hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__); break;             \
  }

// --------------------------------------------------------------- host side

// Device scratch from the graph's pool, returned on scope exit.
struct Scratch {
  vs_graph_s *g;
  void *p = nullptr;
  size_t bytes = 0;
  explicit Scratch(vs_graph_s *gg) : g(gg) {}
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
  vs_status get(size_t b, bool zero) {
    bytes = std::max<size_t>(b, 8);
    return g->pool.get(bytes, zero, g->stream, &p);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
  ~Scratch() {
    if (p) g->pool.put(p, bytes, false);
  }
};

static vs_status colmat_new(vs_graph_s *g, uint32_t rows, uint32_t cols, bool zero,
                            std::unique_ptr<vs_colmat_s> &out) {
  std::unique_ptr<vs_colmat_s> m(new vs_colmat_s);
  m->g = g;
  m->rows = rows;
  m->cols = cols;
  m->pitch = colmat_pitch(rows);
  m->bytes = std::max<size_t>(static_cast<size_t>(cols) * m->pitch * 8, 8);
  void *p = nullptr;
  VS_TRY(g->pool.get(m->bytes, zero, g->stream, &p));
  m->d = static_cast<uint64_t *>(p);
  g->refs++;
  out = std::move(m);
  return VS_OK;
}

static void colmat_release(vs_colmat_s *m) {
  if (!m) return;
  vs_graph_s *g = m->g;
  (void)hipSetDevice(g->device);
  (void)hipStreamSynchronize(g->stream);
  if (m->d) g->pool.put(m->d, m->bytes, false);
  if (m->d_row_map) g->pool.put(m->d_row_map, m->row_map_bytes, false);
  if (m->d_col_keys) g->pool.put(m->d_col_keys, m->col_keys_bytes, false);
  graph_release(g);
  delete m;
}

static vs_status colmat_upload_maps(vs_colmat_s *m) {
  vs_graph_s *g = m->g;
  if (m->d_row_map) { g->pool.put(m->d_row_map, m->row_map_bytes, false); m->d_row_map = nullptr; }
  if (m->d_col_keys) { g->pool.put(m->d_col_keys, m->col_keys_bytes, false); m->d_col_keys = nullptr; }
  m->ncol_keys = 0;
  if (!m->row_map.empty()) {
    m->row_map_bytes = m->row_map.size() * 4;
    void *p = nullptr;
    VS_TRY(g->pool.get(m->row_map_bytes, false, g->stream, &p));
    m->d_row_map = static_cast<uint32_t *>(p);
    VS_HIP(hipMemcpyAsync(p, m->row_map.data(), m->row_map_bytes, hipMemcpyHostToDevice,
                          g->stream));
  }
  if (!m->col_map.empty()) {
    std::vector<uint64_t> keys(m->col_map.size());
    for (size_t c = 0; c < keys.size(); ++c)
      keys[c] = (static_cast<uint64_t>(m->col_map[c]) << 32) | c;
    std::sort(keys.begin(), keys.end());
    // keep the first column of every id (set_col_hash_map's insert keeps the
    // first, bit_matrix_table.cpp:33-36)
    keys.erase(std::unique(keys.begin(), keys.end(),
                           [](uint64_t x, uint64_t y) { return (x >> 32) == (y >> 32); }),
               keys.end());
    m->ncol_keys = static_cast<uint32_t>(keys.size());
    m->col_keys_bytes = keys.size() * 8;
    void *p = nullptr;
    VS_TRY(g->pool.get(m->col_keys_bytes, false, g->stream, &p));
    m->d_col_keys = static_cast<uint64_t *>(p);
    VS_HIP(hipMemcpyAsync(p, keys.data(), m->col_keys_bytes, hipMemcpyHostToDevice,
                          g->stream));
  }
  VS_HIP(hipStreamSynchronize(g->stream));
  return VS_OK;
}

// id of column c: through the column map; columns past a short map keep
// their index
inline uint32_t col_id(const std::vector<uint32_t> &col_map, uint32_t c) {
  return c < col_map.size() ? col_map[c] : c;
}

// column counts on the device (uint32 per column) -> host ids/counts in the
// requested order
static vs_status finish_col_counts(vs_graph_s *g, const uint32_t *d_counts, uint32_t n,
                                   const std::vector<uint32_t> &col_map, int32_t order,
                                   uint32_t *ids, int32_t *counts) {
  if (n == 0) return VS_OK;
  if (!ids || !counts) { set_error("column counts: null output"); return VS_ERR_INVALID; }
  hipStream_t s = g->stream;
  if (order == VS_ORDER_NONE) {
    std::vector<uint32_t> h(n);
    VS_HIP(hipMemcpyAsync(h.data(), d_counts, n * 4ull, hipMemcpyDeviceToHost, s));
    VS_HIP(hipStreamSynchronize(s));
    for (uint32_t c = 0; c < n; ++c) {
      ids[c] = col_id(col_map, c);
      counts[c] = static_cast<int32_t>(h[c]);
    }
    return VS_OK;
  }
  if (order != VS_ORDER_ASC && order != VS_ORDER_DESC) {
    set_error("column counts: unknown order %d", order);
    return VS_ERR_INVALID;
  }
  Scratch map(g), keys(g), sorted(g), tmp(g);
  std::vector<uint32_t> full;
  if (!col_map.empty()) {
    full.resize(n);
    for (uint32_t c = 0; c < n; ++c) full[c] = col_id(col_map, c);
    VS_TRY(map.get(n * 4ull, false));
    VS_HIP(hipMemcpyAsync(map.p, full.data(), n * 4ull, hipMemcpyHostToDevice, s));
  }
  VS_TRY(keys.get(n * 8ull, false));
  VS_TRY(sorted.get(n * 8ull, false));
  hipLaunchKernelGGL(k_count_keys, dim3(grid_for((n + 255) / 256, g->num_cus)), dim3(256), 0, s,
                     d_counts, col_map.empty() ? nullptr : map.as<uint32_t>(), n,
                     keys.as<uint64_t>());
  VS_HIP(hipGetLastError());
  size_t tb = 0;
  if (order == VS_ORDER_ASC) {
    VS_HIP(rocprim::radix_sort_keys(nullptr, tb, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    static_cast<size_t>(n), 0, 64, s));
    VS_TRY(tmp.get(tb, false));
    VS_HIP(rocprim::radix_sort_keys(tmp.p, tb, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    static_cast<size_t>(n), 0, 64, s));
  } else {
    VS_HIP(rocprim::radix_sort_keys_desc(nullptr, tb, keys.as<uint64_t>(),
                                         sorted.as<uint64_t>(), static_cast<size_t>(n), 0,
                                         64, s));
    VS_TRY(tmp.get(tb, false));
    VS_HIP(rocprim::radix_sort_keys_desc(tmp.p, tb, keys.as<uint64_t>(),
                                         sorted.as<uint64_t>(), static_cast<size_t>(n), 0,
                                         64, s));
  }
  std::vector<uint64_t> h(n);
  VS_HIP(hipMemcpyAsync(h.data(), sorted.p, n * 8ull, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  for (uint32_t c = 0; c < n; ++c) {
    ids[c] = static_cast<uint32_t>(h[c]);
    counts[c] = static_cast<int32_t>(h[c] >> 32);
  }
  return VS_OK;
}

static vs_status finish_row_or(vs_graph_s *g, const uint32_t *d_counts, uint32_t n,
                               const std::vector<uint32_t> &col_map, uint32_t *ids,
                               uint32_t cap, uint32_t *nout) {
  std::vector<uint32_t> h(n);
  if (n) {
    VS_HIP(hipMemcpyAsync(h.data(), d_counts, n * 4ull, hipMemcpyDeviceToHost, g->stream));
    VS_HIP(hipStreamSynchronize(g->stream));
  }
  uint32_t k = 0;
  for (uint32_t c = 0; c < n; ++c) k += h[c] != 0;
  if (nout) *nout = k;
  if (!ids) return VS_OK;
  if (cap < k) {
    set_error("row-or: capacity %u < %u columns", cap, k);
    return VS_ERR_CAPACITY;
  }
  k = 0;
  for (uint32_t c = 0; c < n; ++c)
    if (h[c]) ids[k++] = col_id(col_map, c);
  return VS_OK;
}

static vs_status result_mat(vs_result_t r, uint32_t i) {
  if (!r) { set_error("null result"); return VS_ERR_INVALID; }
  if (i >= r->mats.size()) {
    set_error("matrix index %u out of range (%zu matrices)", i, r->mats.size());
    return VS_ERR_INVALID;
  }
  return VS_OK;
}

static vs_status result_col_counts_dev(vs_result_t r, uint32_t i, Scratch &cnt) {
  vs_graph_s *g = r->g;
  VS_TRY(cnt.get(r->V * 4ull, true));
  if (r->V == 0 || r->S == 0 || !r->mats[i].d) return VS_OK;
  const uint32_t G = 64 / r->tw;
  const uint32_t grid = grid_for((r->V + G - 1) / G, g->num_cus);
  VS_TW_DISPATCH_T(r->tw, k_colcount_native, dim3(grid), dim3(kBlock), 0, g->stream,
                   r->mats[i].d, r->V, r->ntiles, cnt.as<uint32_t>());
  VS_HIP(hipGetLastError());
  return VS_OK;
}

static vs_status colmat_col_counts_dev(vs_colmat_s *m, Scratch &cnt) {
  vs_graph_s *g = m->g;
  VS_TRY(cnt.get(m->cols * 4ull, true));
  if (m->cols == 0) return VS_OK;
  const uint32_t grid = grid_for(m->cols, g->num_cus);
  hipLaunchKernelGGL(k_colcount_cm, dim3(grid), dim3(kBlock), 0, g->stream, m->d, m->cols,
                     m->pitch, (static_cast<uint64_t>(m->rows) + 63) / 64, cnt.as<uint32_t>());
  VS_HIP(hipGetLastError());
  return VS_OK;
}

// Count pass over tuples [t0, t0 + nt) into part (nt * nchunk entries).
static vs_status mc_count(vs_graph_s *g, const McArgs &a, uint64_t t0, uint64_t nt,
                          uint32_t *part) {
  const uint64_t nitems = nt * a.nchunk;
  hipLaunchKernelGGL(k_mc_count, dim3(grid_for(nitems, g->num_cus)), dim3(kBlock), 0,
                     g->stream, a, t0, nitems, part);
  VS_HIP(hipGetLastError());
  return VS_OK;
}

__global__ void k_mc_last(const uint32_t *counts, uint64_t *offsets, uint64_t T) {
  if (threadIdx.x == 0 && blockIdx.x == 0) offsets[T] = offsets[T - 1] + counts[T - 1];
}

// Device core of the ColMain intersection: tuples, counts and offsets (T + 1,
// the exclusive scan of counts and the total) in device memory.  Once the
// total is known, lists_for(total) gives the device list buffer (nullptr:
// counts and offsets only; it may also fail, which ends the call with its
// status).  *nfound counts the tuples whose columns all resolved.
static vs_status mc_device(const vs_colmat_t *deps, uint32_t nd, const uint32_t *d_tuples,
                           uint64_t T, uint32_t *d_counts, uint64_t *d_offsets,
                           const std::function<vs_status(uint64_t, uint32_t **)> &lists_for,
                           uint64_t *total, uint64_t *nfound, float *ms) {
  *total = 0;
  *nfound = 0;
  *ms = 0.f;
  if (!deps || nd == 0 || nd > kMaxColmatDeps) {
    set_error("vs_mintersect_colmat: %u dependencies (1..%d)", nd, kMaxColmatDeps);
    return VS_ERR_INVALID;
  }
  for (uint32_t d = 0; d < nd; ++d) {
    if (!deps[d]) { set_error("vs_mintersect_colmat: null dependency %u", d); return VS_ERR_INVALID; }
    if (deps[d]->g->device != deps[0]->g->device) {
      set_error("vs_mintersect_colmat: dependencies on different devices");
      return VS_ERR_INVALID;
    }
    // the reference asserts one row count for all ColMain dependencies
    // (mintersect.cpp:70-77)
    if (deps[d]->rows != deps[0]->rows) {
      set_error("vs_mintersect_colmat: row counts differ (%u vs %u)", deps[d]->rows,
                deps[0]->rows);
      return VS_ERR_INVALID;
    }
  }
  vs_graph_s *g = deps[0]->g;
  VS_HIP(hipSetDevice(g->device));
  for (uint32_t d = 0; d < nd; ++d)
    if (deps[d]->g != g) VS_HIP(hipStreamSynchronize(deps[d]->g->stream));
  hipStream_t s = g->stream;
  if (T == 0) {
    const uint64_t z = 0;
    VS_HIP(hipMemcpyAsync(d_offsets, &z, 8, hipMemcpyHostToDevice, s));
    VS_HIP(hipStreamSynchronize(s));
    uint32_t *dl = nullptr;
    return lists_for(0, &dl);
  }
  McArgs a{};
  a.nd = nd;
  a.rows = deps[0]->rows;
  a.pitch = deps[0]->pitch;
  a.used = (static_cast<uint64_t>(a.rows) + 63) / 64;
  a.nchunk = std::max<uint64_t>(1, (a.used + kMcChunkWords - 1) / kMcChunkWords);
  a.T = T;
  for (uint32_t d = 0; d < nd; ++d) {
    a.mat[d] = deps[d]->d;
    a.cols[d] = deps[d]->cols;
    a.keys[d] = deps[d]->d_col_keys;
    a.nkeys[d] = deps[d]->ncol_keys;
  }
  // target row map: the first ColMain dependency's (mintersect.cpp:115-122)
  a.row_map = deps[0]->d_row_map;
  a.tuples = d_tuples;
  Scratch dc(g), found(g), part(g), tmp(g), pscan(g);
  VS_TRY(dc.get(T * nd * 4, false));
  VS_TRY(found.get(8, true));
  a.colidx = dc.as<uint32_t>();
  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } eg{e0, e1};
  VS_HIP(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_mc_resolve, dim3(grid_for((T + 255) / 256, g->num_cus)), dim3(256), 0, s,
                     a, found.as<ull>());
  VS_HIP(hipGetLastError());
  const uint64_t TB = std::max<uint64_t>(1, std::min<uint64_t>(T, kMcMaxItems / a.nchunk));
  VS_TRY(part.get(TB * a.nchunk * 4, false));
  for (uint64_t t0 = 0; t0 < T; t0 += TB) {
    const uint64_t nt = std::min(TB, T - t0);
    VS_TRY(mc_count(g, a, t0, nt, part.as<uint32_t>()));
    hipLaunchKernelGGL(k_mc_sum, dim3(grid_for((nt + 255) / 256, g->num_cus)), dim3(256), 0, s,
                       part.as<uint32_t>(), a.nchunk, nt, d_counts + t0);
    VS_HIP(hipGetLastError());
  }
  // offsets: exclusive scan of the counts, the total at T
  size_t tb = 0;
  VS_HIP(rocprim::exclusive_scan(nullptr, tb, d_counts, d_offsets, static_cast<uint64_t>(0),
                                 static_cast<size_t>(T), rocprim::plus<uint64_t>(), s));
  const size_t pitems = TB * a.nchunk;
  size_t tb2 = 0;
  VS_HIP(rocprim::exclusive_scan(nullptr, tb2, part.as<uint32_t>(), static_cast<uint64_t *>(nullptr),
                                 static_cast<uint64_t>(0), pitems, rocprim::plus<uint64_t>(), s));
  VS_TRY(tmp.get(std::max(tb, tb2), false));
  VS_HIP(rocprim::exclusive_scan(tmp.p, tb, d_counts, d_offsets, static_cast<uint64_t>(0),
                                 static_cast<size_t>(T), rocprim::plus<uint64_t>(), s));
  hipLaunchKernelGGL(k_mc_last, dim3(1), dim3(64), 0, s, d_counts, d_offsets, T);
  VS_HIP(hipGetLastError());
  uint64_t tot = 0, nf = 0;
  VS_HIP(hipMemcpyAsync(&tot, d_offsets + T, 8, hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(&nf, found.p, 8, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  *total = tot;
  *nfound = nf;
  uint32_t *d_lists = nullptr;
  VS_TRY(lists_for(tot, &d_lists));
  if (d_lists && tot) {
    VS_TRY(pscan.get(pitems * 8, false));
    for (uint64_t t0 = 0; t0 < T; t0 += TB) {
      const uint64_t nt = std::min(TB, T - t0);
      const uint64_t nitems = nt * a.nchunk;
      uint64_t base = 0;
      VS_HIP(hipMemcpyAsync(&base, d_offsets + t0, 8, hipMemcpyDeviceToHost, s));
      VS_TRY(mc_count(g, a, t0, nt, part.as<uint32_t>()));
      VS_HIP(rocprim::exclusive_scan(tmp.p, tb2, part.as<uint32_t>(), pscan.as<uint64_t>(),
                                     static_cast<uint64_t>(0), static_cast<size_t>(nitems),
                                     rocprim::plus<uint64_t>(), s));
      VS_HIP(hipStreamSynchronize(s));  // base is read below
      hipLaunchKernelGGL(k_mc_fill, dim3(grid_for(nitems, g->num_cus)), dim3(kBlock), 0, s, a,
                         t0, nitems, pscan.as<uint64_t>(), base, d_lists);
      VS_HIP(hipGetLastError());
    }
  }
  VS_HIP(hipEventRecord(e1, s));
  VS_HIP(hipEventSynchronize(e1));
  VS_HIP(hipEventElapsedTime(ms, e0, e1));
  return VS_OK;
}

static vs_status mintersect_colmat_impl(const vs_colmat_t *deps, uint32_t nd,
                                        const uint32_t *tuples, uint64_t T, uint32_t *counts,
                                        uint64_t *offsets, uint32_t *lists, uint64_t list_cap,
                                        uint64_t *total, vs_mintersect_stats *st) {
  if (st) std::memset(st, 0, sizeof(*st));
  if (total) *total = 0;
  if (!deps || nd == 0 || nd > kMaxColmatDeps || !deps[0]) {
    set_error("vs_mintersect_colmat: %u dependencies (1..%d)", nd, kMaxColmatDeps);
    return VS_ERR_INVALID;
  }
  if (lists && !offsets) { set_error("vs_mintersect_colmat: lists need offsets"); return VS_ERR_INVALID; }
  if (T == 0) {
    if (offsets) offsets[0] = 0;
    return VS_OK;
  }
  if (!tuples || !counts) { set_error("vs_mintersect_colmat: null tuples/counts"); return VS_ERR_INVALID; }
  vs_graph_s *g = deps[0]->g;
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  Scratch dt(g), dcnt(g), doff(g), dl(g);
  VS_TRY(dt.get(T * nd * 4, false));
  VS_TRY(dcnt.get(T * 4, false));
  VS_TRY(doff.get((T + 1) * 8, false));
  VS_HIP(hipMemcpyAsync(dt.p, tuples, T * nd * 4, hipMemcpyHostToDevice, s));
  bool over = false;
  auto lists_for = [&](uint64_t tot, uint32_t **out) -> vs_status {
    *out = nullptr;
    if (!lists || tot == 0) return VS_OK;
    if (tot > list_cap) {
      over = true;
      return VS_OK;
    }
    VS_TRY(dl.get(tot * 4, false));
    *out = dl.as<uint32_t>();
    return VS_OK;
  };
  uint64_t tot = 0, nfound = 0;
  float ms = 0.f;
  VS_TRY(mc_device(deps, nd, dt.as<uint32_t>(), T, dcnt.as<uint32_t>(), doff.as<uint64_t>(),
                   lists_for, &tot, &nfound, &ms));
  VS_HIP(hipMemcpyAsync(counts, dcnt.p, T * 4, hipMemcpyDeviceToHost, s));
  if (offsets) VS_HIP(hipMemcpyAsync(offsets, doff.p, (T + 1) * 8, hipMemcpyDeviceToHost, s));
  if (lists && !over && tot) VS_HIP(hipMemcpyAsync(lists, dl.p, tot * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  if (total) *total = tot;
  if (over) {
    set_error("vs_mintersect_colmat: list capacity %llu < total %llu",
              (unsigned long long)list_cap, (unsigned long long)tot);
    return VS_ERR_CAPACITY;
  }
  if (st) {
    st->ms = ms;
    st->total = tot;
    // compulsory bytes: the columns of every fully resolved tuple, the
    // tuples, counts, offsets and lists
    const uint64_t used = (static_cast<uint64_t>(deps[0]->rows) + 63) / 64;
    st->bytes_alg = nfound * nd * used * 8 + 4ull * nd * T + 4 * T +
                    (offsets ? 8 * (T + 1) : 0) + (lists ? 4 * tot : 0);
  }
  return VS_OK;
}

// ------------------------------------------- Mintersect, VarRow base table
// Base tuple t of the VarRow-base branch: column d is the pair's start
// vertex, or its target when bit d of `sel` is set.
__global__ void k_vr_tuples(const uint32_t *__restrict__ start, const uint32_t *__restrict__ target,
                            uint64_t P, uint32_t nd, uint32_t sel, uint32_t *__restrict__ tuples) {
  for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < P;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    for (uint32_t d = 0; d < nd; ++d) tuples[t * nd + d] = ((sel >> d) & 1) ? target[t] : start[t];
}

// (start, target) of base tuple t repeated over its output run
// offsets[t] .. + counts[t]: a thread writes a short run itself, the wave
// writes the long runs of its lanes together.
__global__ __launch_bounds__(kBlock) void k_vr_spread(const uint32_t *__restrict__ start,
                                                     const uint32_t *__restrict__ target,
                                                     const uint32_t *__restrict__ counts,
                                                     const uint64_t *__restrict__ offsets,
                                                     uint64_t P, uint32_t *__restrict__ out_s,
                                                     uint32_t *__restrict__ out_t) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t lim = (P + 63) / 64 * 64;  // whole waves stay in the loop
  for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < lim;
       t += stride) {
    const uint32_t n = t < P ? counts[t] : 0;
    const uint64_t o = n ? offsets[t] : 0;
    const uint32_t a = n ? start[t] : 0, b = n ? target[t] : 0;
    const bool big = n > 16;
    if (!big)
      for (uint32_t k = 0; k < n; ++k) {
        out_s[o + k] = a;
        out_t[o + k] = b;
      }
    for (uint64_t m = __ballot(big); m; m &= m - 1) {
      const int l = __ffsll(static_cast<long long>(m)) - 1;
      const uint32_t nl = __shfl(n, l), al = __shfl(a, l), bl = __shfl(b, l);
      const uint64_t ol = shfl64(o, l);
      for (uint32_t k = lane_id(); k < nl; k += 64) {
        out_s[ol + k] = al;
        out_t[ol + k] = bl;
      }
    }
  }
}

static vs_status mintersect_varrow(vs_result_t base, uint32_t i, const vs_colmat_t *mat_deps,
                                   const vs_graph_t *graph_deps, uint32_t nd,
                                   const uint32_t *var_idx, uint32_t *start, uint32_t *target,
                                   uint32_t *nv, uint64_t cap, uint64_t *n,
                                   vs_mintersect_stats *st) {
  const char *fn = "vs_mintersect_varrow";
  if (st) std::memset(st, 0, sizeof(*st));
  if (!n) { set_error("%s: null count", fn); return VS_ERR_INVALID; }
  *n = 0;
  VS_TRY(result_mat(base, i));
  if ((mat_deps == nullptr) == (graph_deps == nullptr)) {
    set_error("%s: give either ColMain or edge-relation dependencies", fn);
    return VS_ERR_INVALID;
  }
  if (nd == 0 || nd > kMaxColmatDeps || !var_idx) {
    set_error("%s: %u dependencies (1..%d) and their base variables", fn, nd, kMaxColmatDeps);
    return VS_ERR_INVALID;
  }
  if (start && (!target || !nv)) { set_error("%s: null target / new-vertex output", fn); return VS_ERR_INVALID; }
  uint32_t sel = 0;
  for (uint32_t d = 0; d < nd; ++d) {
    if (var_idx[d] > 1) {
      set_error("%s: base variable %u of dependency %u (0 = start, 1 = target)", fn, var_idx[d], d);
      return VS_ERR_INVALID;
    }
    const int dev = mat_deps ? (mat_deps[d] ? mat_deps[d]->g->device : -1)
                             : (graph_deps[d] ? graph_deps[d]->device : -1);
    if (dev != base->g->device) {
      set_error("%s: dependency %u missing or on another device", fn, d);
      return VS_ERR_INVALID;
    }
    sel |= var_idx[d] << d;
  }
  vs_graph_s *g = base->g;
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  // base tuples: the set bits of the base matrix, in (row, target) order
  // (VarRowMatrixTable iterators, bit_matrix_table.cpp:165-188)
  uint32_t *ps = nullptr, *pt = nullptr;
  size_t pbytes = 0;
  uint64_t P = 0;
  VS_TRY(flat_pairs_device(base, i, true, &ps, &pt, &pbytes, &P));
  struct Pairs {
    Pool &pool; void *a, *b; size_t n;
    ~Pairs() { if (a) pool.put(a, n, false); if (b) pool.put(b, n, false); }
  } pairs{g->pool, ps, pt, pbytes};
  if (P == 0) return VS_OK;
  Scratch tup(g), cnt(g), off(g), lst(g), os(g), ot(g);
  VS_TRY(tup.get(P * nd * 4, false));
  VS_TRY(cnt.get(P * 4, false));
  VS_TRY(off.get((P + 1) * 8, false));
  hipLaunchKernelGGL(k_vr_tuples, dim3(grid_for((P + 255) / 256, g->num_cus)), dim3(256), 0, s,
                     ps, pt, P, nd, sel, tup.as<uint32_t>());
  VS_HIP(hipGetLastError());
  uint64_t tot = 0;
  float ms = 0.f;
  const bool want = start != nullptr;
  bool over = false;
  uint32_t *d_lists = nullptr;
  if (mat_deps) {
    uint64_t nfound = 0;
    auto lists_for = [&](uint64_t total, uint32_t **out) -> vs_status {
      *out = nullptr;
      if (!want || total == 0) return VS_OK;
      if (total > cap) { over = true; return VS_OK; }
      VS_TRY(lst.get(total * 4, false));
      *out = d_lists = lst.as<uint32_t>();
      return VS_OK;
    };
    VS_TRY(mc_device(mat_deps, nd, tup.as<uint32_t>(), P, cnt.as<uint32_t>(), off.as<uint64_t>(),
                     lists_for, &tot, &nfound, &ms));
  } else {
    // one intersection pass; the list buffer is taken once the total is known
    const ListsFor lists_for = [&](uint64_t total, uint32_t **out) -> vs_status {
      *out = nullptr;
      if (total == 0) return VS_OK;
      if (total > cap) { over = true; return VS_OK; }
      VS_TRY(lst.get(total * 4, false));
      *out = d_lists = lst.as<uint32_t>();
      return VS_OK;
    };
    vs_mintersect_stats ms1{};
    VS_TRY(mintersect_run(graph_deps, nd, tup.as<uint32_t>(), P, cnt.as<uint32_t>(),
                          off.as<uint64_t>(), want ? &lists_for : nullptr, &tot, &ms1));
    ms = static_cast<float>(ms1.ms);
  }
  *n = tot;
  if (st) {
    st->ms = ms;
    st->total = tot;
  }
  if (!want) return VS_OK;
  if (over) {
    set_error("%s: capacity %llu < %llu tuples", fn, static_cast<unsigned long long>(cap),
              static_cast<unsigned long long>(tot));
    return VS_ERR_CAPACITY;
  }
  if (tot == 0) return VS_OK;
  VS_TRY(os.get(tot * 4, false));
  VS_TRY(ot.get(tot * 4, false));
  hipLaunchKernelGGL(k_vr_spread, dim3(grid_for((P + 63) / 64, g->num_cus)), dim3(kBlock), 0, s,
                     ps, pt, cnt.as<uint32_t>(), off.as<uint64_t>(), P, os.as<uint32_t>(),
                     ot.as<uint32_t>());
  VS_HIP(hipGetLastError());
  VS_HIP(hipMemcpyAsync(start, os.p, tot * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(target, ot.p, tot * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(nv, d_lists, tot * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  return VS_OK;
}

} // namespace vs

using namespace vs;

extern "C" {

vs_status vs_colmat_create(vs_graph_t g, uint32_t rows, uint32_t cols, const uint64_t *words,
                           vs_colmat_t *out) try {
  clear_error();
  if (!g || !out) { set_error("vs_colmat_create: null argument"); return VS_ERR_INVALID; }
  *out = nullptr;
  VS_HIP(hipSetDevice(g->device));
  std::unique_ptr<vs_colmat_s> m;
  VS_TRY(colmat_new(g, rows, cols, words == nullptr, m));
  if (words && cols) {
    VS_HIP(hipMemcpyAsync(m->d, words, static_cast<size_t>(cols) * m->pitch * 8,
                          hipMemcpyHostToDevice, g->stream));
    const uint64_t used = (static_cast<uint64_t>(rows) + 63) / 64;
    const uint64_t tail = m->pitch - used + ((rows % 64) ? 1 : 0);
    if (tail) {
      hipLaunchKernelGGL(k_cm_clear_pad, dim3(grid_for((cols * tail + 255) / 256, g->num_cus)),
                         dim3(256), 0, g->stream, m->d, cols, m->pitch, rows);
      VS_HIP(hipGetLastError());
    }
    VS_HIP(hipStreamSynchronize(g->stream));
  }
  *out = m.release();
  return VS_OK;
} VS_CATCH("vs_colmat_create")

vs_status vs_colmat_set_maps(vs_colmat_t m, const uint32_t *row_map, const uint32_t *col_map,
                             uint32_t n_col_map) try {
  clear_error();
  if (!m) { set_error("vs_colmat_set_maps: null matrix"); return VS_ERR_INVALID; }
  if (col_map && n_col_map > m->cols) {
    set_error("vs_colmat_set_maps: %u column ids for %u columns", n_col_map, m->cols);
    return VS_ERR_INVALID;
  }
  VS_HIP(hipSetDevice(m->g->device));
  if (row_map) m->row_map.assign(row_map, row_map + m->rows);
  else m->row_map.clear();
  if (col_map) m->col_map.assign(col_map, col_map + n_col_map);
  else m->col_map.clear();
  return colmat_upload_maps(m);
} VS_CATCH("vs_colmat_set_maps")

vs_status vs_colmat_info(vs_colmat_t m, uint32_t *rows, uint32_t *cols, uint64_t *pitch_words,
                         int *has_row_map, uint32_t *n_col_map) try {
  clear_error();
  if (!m) { set_error("vs_colmat_info: null matrix"); return VS_ERR_INVALID; }
  if (rows) *rows = m->rows;
  if (cols) *cols = m->cols;
  if (pitch_words) *pitch_words = m->pitch;
  if (has_row_map) *has_row_map = !m->row_map.empty();
  if (n_col_map) *n_col_map = static_cast<uint32_t>(m->col_map.size());
  return VS_OK;
} VS_CATCH("vs_colmat_info")

vs_status vs_colmat_get_maps(vs_colmat_t m, uint32_t *row_map, uint32_t *col_map) try {
  clear_error();
  if (!m) { set_error("vs_colmat_get_maps: null matrix"); return VS_ERR_INVALID; }
  if (row_map) {
    if (m->row_map.empty()) { set_error("matrix has no row map"); return VS_ERR_INVALID; }
    std::memcpy(row_map, m->row_map.data(), m->row_map.size() * 4);
  }
  if (col_map) {
    if (m->col_map.empty()) { set_error("matrix has no column map"); return VS_ERR_INVALID; }
    std::memcpy(col_map, m->col_map.data(), m->col_map.size() * 4);
  }
  return VS_OK;
} VS_CATCH("vs_colmat_get_maps")

vs_status vs_colmat_export(vs_colmat_t m, uint64_t *words, size_t bytes) try {
  clear_error();
  if (!m) { set_error("vs_colmat_export: null matrix"); return VS_ERR_INVALID; }
  const size_t need = static_cast<size_t>(m->cols) * m->pitch * 8;
  if (bytes < need) {
    set_error("vs_colmat_export: buffer %zu bytes < %zu", bytes, need);
    return VS_ERR_CAPACITY;
  }
  if (need == 0) return VS_OK;
  if (!words) { set_error("vs_colmat_export: null buffer"); return VS_ERR_INVALID; }
  VS_HIP(hipSetDevice(m->g->device));
  VS_HIP(hipMemcpyAsync(words, m->d, need, hipMemcpyDeviceToHost, m->g->stream));
  VS_HIP(hipStreamSynchronize(m->g->stream));
  return VS_OK;
} VS_CATCH("vs_colmat_export")

vs_status vs_colmat_get(vs_colmat_t m, uint32_t row, uint32_t col, int *bit) try {
  clear_error();
  if (!m || !bit || row >= m->rows || col >= m->cols) {
    set_error("vs_colmat_get: (%u,%u) outside the matrix", row, col);
    return VS_ERR_INVALID;
  }
  uint64_t w = 0;
  VS_HIP(hipSetDevice(m->g->device));
  VS_HIP(hipStreamSynchronize(m->g->stream));
  VS_HIP(hipMemcpy(&w, m->d + static_cast<uint64_t>(col) * m->pitch + row / 64, 8,
                   hipMemcpyDeviceToHost));
  *bit = static_cast<int>((w >> (row % 64)) & 1);
  return VS_OK;
} VS_CATCH("vs_colmat_get")

vs_status vs_colmat_free(vs_colmat_t m) try {
  clear_error();
  colmat_release(m);
  return VS_OK;
} VS_CATCH("vs_colmat_free")

vs_status vs_table_convert(vs_result_t r, uint32_t i, int transpose, vs_colmat_t *out) try {
  clear_error();
  VS_TRY(result_mat(r, i));
  if (!out) { set_error("vs_table_convert: null output"); return VS_ERR_INVALID; }
  *out = nullptr;
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  const Mat &mat = r->mats[i];
  std::unique_ptr<vs_colmat_s> m;
  if (!transpose) {
    VS_TRY(colmat_new(g, r->S, r->V, false, m));
    const uint64_t n = static_cast<uint64_t>(r->V) * m->pitch;
    if (n && r->S && mat.d) {
      CallerOrder co;  // ColMain rows are the caller's rows, in order
      VS_TRY(co.open(r, i));
      const uint32_t W = (r->S + 63) / 64;
      VS_TW_DISPATCH_T(r->tw, k_cm_from_native,
                       dim3(grid_for((n + 255) / 256, g->num_cus)), dim3(kBlock), 0, g->stream,
                       co.d, r->V, W, m->pitch, m->d);
      VS_HIP(hipGetLastError());
    } else if (n) {
      VS_HIP(hipMemsetAsync(m->d, 0, n * 8, g->stream));
    }
    m->row_map = r->row_map;
    m->col_map = r->col_map;
  } else {
    VS_TRY(colmat_new(g, r->V, r->S, true, m));
    if (r->S && r->V && mat.d) {
      DestRows dr;  // column k is caller row k
      VS_TRY(dr.open(r));
      VS_TRY(mat_to_rowmajor(g, mat, r->S, r->V, r->tw, m->d, m->pitch, dr.d));
    }
    m->row_map = r->col_map;
    m->col_map = r->row_map;
  }
  VS_TRY(colmat_upload_maps(m.get()));
  *out = m.release();
  return VS_OK;
} VS_CATCH("vs_table_convert")

vs_status vs_result_select_cols(vs_result_t r, uint32_t i, const uint32_t *col_ids, uint32_t n,
                                vs_result_t *out) try {
  clear_error();
  VS_TRY(result_mat(r, i));
  if (!out || (n && !col_ids)) { set_error("vs_result_select_cols: null argument"); return VS_ERR_INVALID; }
  *out = nullptr;
  // the reference handles only a source without a column map
  // (big_operators2.cpp:165-168)
  if (!r->col_map.empty()) {
    set_error("vs_result_select_cols: the result already has a column map");
    return VS_ERR_INVALID;
  }
  for (uint32_t k = 0; k < n; ++k)
    if (col_ids[k] >= r->V) {
      set_error("vs_result_select_cols: column %u >= %u", col_ids[k], r->V);
      return VS_ERR_INVALID;
    }
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  Mat m;
  m.distance = r->mats[i].distance;
  m.bytes = std::max<size_t>(static_cast<size_t>(r->ntiles) * n * r->tw * 8, 8);
  {
    void *p = nullptr;
    VS_TRY(g->pool.get(m.bytes, true, g->stream, &p));
    m.d = static_cast<uint64_t *>(p);
  }
  const uint64_t total = static_cast<uint64_t>(r->ntiles) * n * r->tw;
  if (total && r->mats[i].d) {
    Scratch sel(g);
    vs_status s = sel.get(n * 4ull, false);
    hipError_t e = hipSuccess;
    if (s == VS_OK) e = hipMemcpyAsync(sel.p, col_ids, n * 4ull, hipMemcpyHostToDevice, g->stream);
    if (s == VS_OK && e == hipSuccess) {
      VS_TW_DISPATCH_T(r->tw, k_select_native,
                       dim3(grid_for((total + 255) / 256, g->num_cus)), dim3(kBlock), 0,
                       g->stream, r->mats[i].d, r->V, r->ntiles, sel.as<uint32_t>(), n, m.d);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (s == VS_OK && e != hipSuccess) {
      set_error("vs_result_select_cols: %s", hipGetErrorString(e));
      s = VS_ERR_HIP;
    }
    if (s != VS_OK) {
      g->pool.put(m.d, m.bytes, false);
      return s;
    }
  }
  std::unique_ptr<vs_result_s> res(new vs_result_s);
  res->g = g;
  res->S = r->S;
  res->V = n;
  res->tw = r->tw;
  res->ntiles = r->ntiles;
  res->k_actual_max = r->k_actual_max;
  res->row_map = r->row_map;
  res->stored = r->stored;  // columns are selected, rows stay where they are
  res->col_map.assign(col_ids, col_ids + n);
  res->mats.push_back(m);
  g->refs++;
  *out = res.release();
  return VS_OK;
} VS_CATCH("vs_result_select_cols")

vs_status vs_result_col_map(vs_result_t r, uint32_t *col_ids) try {
  clear_error();
  if (!r || !col_ids) { set_error("vs_result_col_map: null argument"); return VS_ERR_INVALID; }
  if (r->col_map.empty()) { set_error("result has no column map"); return VS_ERR_INVALID; }
  std::memcpy(col_ids, r->col_map.data(), r->col_map.size() * 4);
  return VS_OK;
} VS_CATCH("vs_result_col_map")

vs_status vs_result_col_counts(vs_result_t r, uint32_t i, int32_t order, uint32_t *ids,
                               int32_t *counts) try {
  clear_error();
  VS_TRY(result_mat(r, i));
  VS_HIP(hipSetDevice(r->g->device));
  Scratch cnt(r->g);
  VS_TRY(result_col_counts_dev(r, i, cnt));
  return finish_col_counts(r->g, cnt.as<uint32_t>(), r->V, r->col_map, order, ids, counts);
} VS_CATCH("vs_result_col_counts")

vs_status vs_colmat_col_counts(vs_colmat_t m, int32_t order, uint32_t *ids, int32_t *counts) try {
  clear_error();
  if (!m) { set_error("vs_colmat_col_counts: null matrix"); return VS_ERR_INVALID; }
  VS_HIP(hipSetDevice(m->g->device));
  Scratch cnt(m->g);
  VS_TRY(colmat_col_counts_dev(m, cnt));
  return finish_col_counts(m->g, cnt.as<uint32_t>(), m->cols, m->col_map, order, ids, counts);
} VS_CATCH("vs_colmat_col_counts")

vs_status vs_result_row_or(vs_result_t r, uint32_t i, uint32_t *ids, uint32_t cap, uint32_t *n) try {
  clear_error();
  VS_TRY(result_mat(r, i));
  VS_HIP(hipSetDevice(r->g->device));
  Scratch cnt(r->g);
  VS_TRY(result_col_counts_dev(r, i, cnt));
  return finish_row_or(r->g, cnt.as<uint32_t>(), r->V, r->col_map, ids, cap, n);
} VS_CATCH("vs_result_row_or")

vs_status vs_colmat_row_or(vs_colmat_t m, uint32_t *ids, uint32_t cap, uint32_t *n) try {
  clear_error();
  if (!m) { set_error("vs_colmat_row_or: null matrix"); return VS_ERR_INVALID; }
  VS_HIP(hipSetDevice(m->g->device));
  Scratch cnt(m->g);
  VS_TRY(colmat_col_counts_dev(m, cnt));
  return finish_row_or(m->g, cnt.as<uint32_t>(), m->cols, m->col_map, ids, cap, n);
} VS_CATCH("vs_colmat_row_or")

vs_status vs_mintersect_varrow(vs_result_t base, uint32_t i, const vs_colmat_t *mat_deps,
                               const vs_graph_t *graph_deps, uint32_t nd, const uint32_t *var_idx,
                               uint32_t *start, uint32_t *target, uint32_t *nv, uint64_t cap,
                               uint64_t *n, vs_mintersect_stats *stats) try {
  clear_error();
  return mintersect_varrow(base, i, mat_deps, graph_deps, nd, var_idx, start, target, nv, cap, n,
                           stats);
} VS_CATCH("vs_mintersect_varrow")

vs_status vs_mintersect_colmat(const vs_colmat_t *deps, uint32_t nd, const uint32_t *tuples,
                               uint64_t T, uint32_t *counts, uint64_t *offsets, uint32_t *lists,
                               uint64_t list_cap, uint64_t *total, vs_mintersect_stats *stats) try {
  clear_error();
  return mintersect_colmat_impl(deps, nd, tuples, T, counts, offsets, lists, list_cap, total,
                                stats);
} VS_CATCH("vs_mintersect_colmat")

} // extern "C"
