// Flat export of a VExpand result: one (start, target[, distance]) tuple per
// set bit, the table VExpand's CSR strategy produces (merge_to_Flat,
// vexpand.cpp:380-446).  The bit matrices stay the computation; this is the
// output format.
//
// Pairs come from the union of the result's matrices.  With per-hop layers
// (with_distance) a pair's distance is that of the first layer holding it:
// Shortest layers are disjoint, and for Any this is the reference's dedup,
// which keeps the first (smallest) distance.  Order: by row, then by target
// column ascending (the reference's order depends on its threads).
//
//   count: one wave per (tile, chunk of kFlatChunk vertices): bit-sliced
//          per-row counts of the union words  ->  cnt[row][chunk]
//   scan:  exclusive scan over cnt in (row, chunk) order  ->  each row and
//          chunk's first output slot
//   fill:  one wave per (tile, word, chunk): lanes = 64 vertices at a time;
//          64 ballots transpose the 64 x 64 block (lane b gets the vertices
//          of row b), lane b writes its row's tuples in vertex order; the
//          first-hit layer of every (vertex, row) bit goes through LDS.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "vs_device.hh"

namespace vs {

constexpr uint32_t kFlatChunk = 8192;
constexpr uint32_t kFlatMaxLayers = 72;

struct FlatArgs {
  const uint64_t *m[kFlatMaxLayers];
  int32_t dist[kFlatMaxLayers];
  uint32_t nm;
  uint32_t S, V, ntiles, nch;
  uint32_t *cnt;           // [S][nch]
  const uint64_t *off;     // [S * nch + 1]
  const uint32_t *row_map; // [S] or nullptr
  const uint32_t *col_map; // [V] or nullptr
  uint32_t *start, *target;
  int32_t *distout;        // or nullptr
};

template <int TW>
__global__ __launch_bounds__(kBlock) void k_flat_count(FlatArgs f) {
  constexpr int G = 64 / TW;
  constexpr int NB = 14;  // counts up to kFlatChunk = 2^13
  const int lane = lane_id(), wid = wave_id();
  const int g = lane / TW, j = lane % TW;
  const uint64_t total = static_cast<uint64_t>(f.ntiles) * f.nch;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < total; it += nwaves) {
    const uint32_t t = static_cast<uint32_t>(it / f.nch), c = static_cast<uint32_t>(it % f.nch);
    const uint32_t v0 = c * kFlatChunk, v1 = min(f.V, v0 + kFlatChunk);
    uint64_t cnt[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) cnt[k] = 0;
    for (uint32_t v = v0 + g; v < v1; v += G) {
      const size_t w = (static_cast<size_t>(t) * f.V + v) * TW + j;
      uint64_t x = 0;
      for (uint32_t i = 0; i < f.nm; ++i) x |= f.m[i][w];
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const uint64_t c2 = cnt[k] & x;
        cnt[k] ^= x;
        x = c2;
      }
    }
    for (int b = 0; b < 64; ++b) {
      uint32_t n = 0;
#pragma unroll
      for (int k = 0; k < NB; ++k) n |= static_cast<uint32_t>((cnt[k] >> b) & 1) << k;
#pragma unroll
      for (int m = TW; m < 64; m <<= 1) n += __shfl_xor(n, m);
      const uint32_t row = (t * TW + j) * 64 + b;
      if (g == 0 && row < f.S) f.cnt[static_cast<size_t>(row) * f.nch + c] = n;
    }
  }
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_flat_fill(FlatArgs f) {
  __shared__ uint8_t s_fh[kWavesPerBlock][64][64];  // [vertex lane][row bit]: first layer
  const int lane = lane_id(), wid = wave_id();
  const uint64_t total = static_cast<uint64_t>(f.ntiles) * TW * f.nch;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < total; it += nwaves) {
    const uint32_t c = static_cast<uint32_t>(it % f.nch);
    const uint32_t tj = static_cast<uint32_t>(it / f.nch);  // t * TW + j
    const uint32_t t = tj / TW, j = tj % TW;
    const uint32_t row = tj * 64 + lane;  // this lane's row after the transpose
    const bool live = row < f.S;
    uint64_t pos = live ? f.off[static_cast<size_t>(row) * f.nch + c] : 0;
    const uint32_t sv = live ? (f.row_map ? f.row_map[row] : row) : 0;
    const uint32_t v0 = c * kFlatChunk, v1 = min(f.V, v0 + kFlatChunk);
    for (uint32_t vb = v0; vb < v1; vb += 64) {
      const uint32_t v = vb + lane;
      uint64_t u = 0;
      if (v < v1) {
        const size_t w = (static_cast<size_t>(t) * f.V + v) * TW + j;
        for (uint32_t i = 0; i < f.nm; ++i) {
          const uint64_t x = f.m[i][w];
          if (f.distout) {
            for (uint64_t nw = x & ~u; nw; nw &= nw - 1)
              s_fh[wid][lane][__ffsll(static_cast<long long>(nw)) - 1] = static_cast<uint8_t>(i);
          }
          u |= x;
        }
      }
      wave_sync();
      uint64_t mine = 0;  // vertices (bits) of this lane's row in this block
      for (int b = 0; b < 64; ++b) {
        const uint64_t mb = __ballot((u >> b) & 1);
        if (lane == b) mine = mb;
      }
      if (live) {
        while (mine) {
          const int q = __ffsll(static_cast<long long>(mine)) - 1;
          mine &= mine - 1;
          const uint32_t tv = vb + q;
          f.start[pos] = sv;
          f.target[pos] = f.col_map ? f.col_map[tv] : tv;
          if (f.distout) f.distout[pos] = f.dist[s_fh[wid][q][lane]];
          ++pos;
        }
      }
      wave_sync();
    }
  }
}

template <int TW>
static void launch_flat(bool fill, uint32_t grid, hipStream_t s, const FlatArgs &f) {
  if (fill) hipLaunchKernelGGL(k_flat_fill<TW>, dim3(grid), dim3(kBlock), 0, s, f);
  else hipLaunchKernelGGL(k_flat_count<TW>, dim3(grid), dim3(kBlock), 0, s, f);
}

static void launch_flat_tw(uint32_t tw, bool fill, uint32_t grid, hipStream_t s,
                           const FlatArgs &f) {
  switch (tw) {
    case 1: launch_flat<1>(fill, grid, s, f); break;
    case 2: launch_flat<2>(fill, grid, s, f); break;
    case 4: launch_flat<4>(fill, grid, s, f); break;
    case 8: launch_flat<8>(fill, grid, s, f); break;
    case 16: launch_flat<16>(fill, grid, s, f); break;
    case 32: launch_flat<32>(fill, grid, s, f); break;
    default: launch_flat<64>(fill, grid, s, f); break;
  }
}


// The flat pairs of matrix `which` of r (every matrix's union when which is
// UINT32_MAX), in (row, target) order, into device buffers from the graph's
// pool that the caller returns: d_start / d_target hold *n entries (when
// want is false only *n is computed).  Distances are not produced.
vs_status flat_pairs_device(vs_result_s *r, uint32_t which, bool want, uint32_t **d_start,
                            uint32_t **d_target, size_t *bytes, uint64_t *n) {
  *n = 0;
  *d_start = *d_target = nullptr;
  *bytes = 0;
  const size_t nm = which == UINT32_MAX ? r->mats.size() : 1;
  if (nm == 0 || r->S == 0 || r->V == 0) return VS_OK;
  if (nm > kFlatMaxLayers) {
    set_error("flat pairs: %zu matrices (at most %u)", nm, kFlatMaxLayers);
    return VS_ERR_UNSUPPORTED;
  }
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  FlatArgs f{};
  CallerOrder co[kFlatMaxLayers];
  for (size_t k = 0; k < nm; ++k) {
    const uint32_t mi = which == UINT32_MAX ? static_cast<uint32_t>(k) : which;
    VS_TRY(co[k].open(r, mi));
    f.m[k] = co[k].d;
    f.dist[k] = static_cast<int32_t>(r->mats[mi].distance);
  }
  f.nm = static_cast<uint32_t>(nm);
  f.S = r->S;
  f.V = r->V;
  f.ntiles = r->ntiles;
  f.nch = (r->V + kFlatChunk - 1) / kFlatChunk;
  const size_t ncnt = static_cast<size_t>(r->S) * f.nch;
  struct Buf { void *p = nullptr; size_t b = 0; };
  std::vector<Buf> bufs;
  struct Back {
    Pool &p; std::vector<Buf> &v;
    ~Back() { for (auto &x : v) if (x.p) p.put(x.p, x.b, false); }
  } back{g->pool, bufs};
  auto get = [&](size_t bytes, void **out) -> vs_status {
    bufs.push_back({nullptr, bytes});
    VS_TRY(g->pool.get(bytes, false, s, &bufs.back().p));
    *out = bufs.back().p;
    return VS_OK;
  };
  void *pc = nullptr, *po = nullptr;
  VS_TRY(get(ncnt * 4, &pc));
  VS_TRY(get((ncnt + 1) * 8, &po));
  f.cnt = static_cast<uint32_t *>(pc);
  f.off = static_cast<const uint64_t *>(po);
  const uint32_t grid = grid_for(static_cast<uint64_t>(r->ntiles) * r->tw * f.nch, g->num_cus);
  launch_flat_tw(r->tw, false, grid, s, f);
  VS_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  VS_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, f.cnt, static_cast<uint64_t *>(po),
                                 static_cast<uint64_t>(0), ncnt, rocprim::plus<uint64_t>(), s));
  void *tmp = nullptr;
  VS_TRY(get(std::max<size_t>(tmp_bytes, 8), &tmp));
  VS_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, f.cnt, static_cast<uint64_t *>(po),
                                 static_cast<uint64_t>(0), ncnt, rocprim::plus<uint64_t>(), s));
  uint64_t last_off = 0;
  uint32_t last_cnt = 0;
  VS_HIP(hipMemcpyAsync(&last_off, static_cast<uint64_t *>(po) + ncnt - 1, 8,
                        hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(&last_cnt, f.cnt + ncnt - 1, 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  const uint64_t total = last_off + last_cnt;
  *n = total;
  if (!want || total == 0) return VS_OK;
  void *rm = nullptr, *cm = nullptr, *ds = nullptr, *dt = nullptr;
  if (!r->row_map.empty()) {
    VS_TRY(get(r->row_map.size() * 4, &rm));
    VS_HIP(hipMemcpyAsync(rm, r->row_map.data(), r->row_map.size() * 4, hipMemcpyHostToDevice, s));
  }
  if (!r->col_map.empty()) {
    VS_TRY(get(r->col_map.size() * 4, &cm));
    VS_HIP(hipMemcpyAsync(cm, r->col_map.data(), r->col_map.size() * 4, hipMemcpyHostToDevice, s));
  }
  // the outputs are the caller's: not in the scoped list
  VS_TRY(g->pool.get(total * 4, false, s, &ds));
  vs_status st = g->pool.get(total * 4, false, s, &dt);
  if (st != VS_OK) {
    g->pool.put(ds, total * 4, false);
    return st;
  }
  *d_start = static_cast<uint32_t *>(ds);
  *d_target = static_cast<uint32_t *>(dt);
  *bytes = total * 4;
  f.start = *d_start;
  f.target = *d_target;
  f.distout = nullptr;
  f.row_map = static_cast<const uint32_t *>(rm);
  f.col_map = static_cast<const uint32_t *>(cm);
  launch_flat_tw(r->tw, true, grid, s, f);
  VS_HIP(hipGetLastError());
  VS_HIP(hipStreamSynchronize(s));  // the maps leave the pool on return
  return VS_OK;
}

}  // namespace vs

using namespace vs;

extern "C" vs_status vs_result_flat(vs_result_t r, uint32_t *start, uint32_t *target,
                                    int32_t *dist, uint64_t cap, uint64_t *n) try {
  clear_error();
  if (!r || !n) {
    set_error("vs_result_flat: null argument");
    return VS_ERR_INVALID;
  }
  *n = 0;
  const size_t nm = r->mats.size();
  if (nm == 0 || r->S == 0 || r->V == 0) return VS_OK;
  if (nm > kFlatMaxLayers) {
    set_error("vs_result_flat: %zu matrices (at most %u)", nm, kFlatMaxLayers);
    return VS_ERR_UNSUPPORTED;
  }
  if (dist && r->mats[0].distance == UINT32_MAX) {
    set_error("vs_result_flat: distances need a result with per-hop layers (with_distance)");
    return VS_ERR_INVALID;
  }
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  hipStream_t s = g->stream;
  FlatArgs f{};
  // the table is ordered by caller row: matrices stored in another row order
  // are read through caller-ordered copies
  CallerOrder co[kFlatMaxLayers];
  for (size_t i = 0; i < nm; ++i) {
    VS_TRY(co[i].open(r, static_cast<uint32_t>(i)));
    f.m[i] = co[i].d;
    f.dist[i] = static_cast<int32_t>(r->mats[i].distance);
  }
  f.nm = static_cast<uint32_t>(nm);
  f.S = r->S;
  f.V = r->V;
  f.ntiles = r->ntiles;
  f.nch = (r->V + kFlatChunk - 1) / kFlatChunk;
  const size_t ncnt = static_cast<size_t>(r->S) * f.nch;
  // device buffers, returned to the graph's pool on every path
  struct Buf { void *p = nullptr; size_t b = 0; };
  std::vector<Buf> bufs;
  struct Back {
    Pool &p; std::vector<Buf> &v;
    ~Back() { for (auto &x : v) if (x.p) p.put(x.p, x.b, false); }
  } back{g->pool, bufs};
  auto get = [&](size_t bytes, void **out) -> vs_status {
    bufs.push_back({nullptr, bytes});
    VS_TRY(g->pool.get(bytes, false, s, &bufs.back().p));
    *out = bufs.back().p;
    return VS_OK;
  };
  void *pc = nullptr, *po = nullptr;
  VS_TRY(get(ncnt * 4, &pc));
  VS_TRY(get((ncnt + 1) * 8, &po));
  f.cnt = static_cast<uint32_t *>(pc);
  f.off = static_cast<const uint64_t *>(po);
  const uint32_t grid = grid_for(static_cast<uint64_t>(r->ntiles) * r->tw * f.nch, g->num_cus);
  launch_flat_tw(r->tw, false, grid, s, f);
  VS_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  VS_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, f.cnt, static_cast<uint64_t *>(po),
                                 static_cast<uint64_t>(0), ncnt, rocprim::plus<uint64_t>(), s));
  void *tmp = nullptr;
  VS_TRY(get(std::max<size_t>(tmp_bytes, 8), &tmp));
  VS_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, f.cnt, static_cast<uint64_t *>(po),
                                 static_cast<uint64_t>(0), ncnt, rocprim::plus<uint64_t>(), s));
  uint64_t last_off = 0;
  uint32_t last_cnt = 0;
  VS_HIP(hipMemcpyAsync(&last_off, static_cast<uint64_t *>(po) + ncnt - 1, 8,
                        hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(&last_cnt, f.cnt + ncnt - 1, 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  const uint64_t total = last_off + last_cnt;
  *n = total;
  if (!start) return VS_OK;
  if (!target) {
    set_error("vs_result_flat: null target");
    return VS_ERR_INVALID;
  }
  if (total > cap) {
    set_error("vs_result_flat: capacity %llu < %llu tuples", static_cast<unsigned long long>(cap),
              static_cast<unsigned long long>(total));
    return VS_ERR_CAPACITY;
  }
  if (total == 0) return VS_OK;
  void *ds = nullptr, *dt = nullptr, *dd = nullptr, *rm = nullptr, *cm = nullptr;
  VS_TRY(get(total * 4, &ds));
  VS_TRY(get(total * 4, &dt));
  if (dist) VS_TRY(get(total * 4, &dd));
  if (!r->row_map.empty()) {
    VS_TRY(get(r->row_map.size() * 4, &rm));
    VS_HIP(hipMemcpyAsync(rm, r->row_map.data(), r->row_map.size() * 4, hipMemcpyHostToDevice, s));
  }
  if (!r->col_map.empty()) {
    VS_TRY(get(r->col_map.size() * 4, &cm));
    VS_HIP(hipMemcpyAsync(cm, r->col_map.data(), r->col_map.size() * 4, hipMemcpyHostToDevice, s));
  }
  f.start = static_cast<uint32_t *>(ds);
  f.target = static_cast<uint32_t *>(dt);
  f.distout = static_cast<int32_t *>(dd);
  f.row_map = static_cast<const uint32_t *>(rm);
  f.col_map = static_cast<const uint32_t *>(cm);
  launch_flat_tw(r->tw, true, grid, s, f);
  VS_HIP(hipGetLastError());
  VS_HIP(hipMemcpyAsync(start, ds, total * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipMemcpyAsync(target, dt, total * 4, hipMemcpyDeviceToHost, s));
  if (dist) VS_HIP(hipMemcpyAsync(dist, dd, total * 4, hipMemcpyDeviceToHost, s));
  VS_HIP(hipStreamSynchronize(s));
  return VS_OK;
} VS_CATCH("vs_result_flat")
