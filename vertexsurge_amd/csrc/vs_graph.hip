// Graph handle (CSR resident in HBM), device buffer pool, raw device API and
// the result handle with its read-out kernels (per-row popcounts, per-row
// digests, reference-layout export, row-major export).
//
// Native matrix layout: a query with S sources has W = ceil(S/64) word-rows;
// they are cut into T tiles of TW words (TW a power of two <= 64).  Word j of
// tile t at vertex v lives at ((t * V) + v) * TW + j and holds rows
// 64 * (t * TW + j) .. +63, bit = row % 64.  With TW = 16 this is exactly the
// reference DenseVarStackMatrix with 1024-row stacks
// (src/include/storage/bitmap.hh:348-356, src/objects/storage/bitmap.cpp:910-920:
// v_per_stack_col = 16, v_per_stack = 16 * C, word(r, c) =
// (r / 1024) * v_per_stack + c * 16 + (r % 1024) / 64).
#include <algorithm>
#include <cstring>
#include <thread>

#include "vs_device.hh"

namespace vs {

vs_status Pool::get(size_t bytes, bool zero, hipStream_t s, void **out) {
  *out = nullptr;
  if (bytes == 0) return VS_OK;
  auto it = free_.find(bytes);
  if (it != free_.end()) {
    void *p = it->second.first;
    const bool clean = it->second.second;
    free_.erase(it);
    if (zero && !clean) VS_HIP(hipMemsetAsync(p, 0, bytes, s));
    *out = p;
    return VS_OK;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // drop every cached buffer and retry once
    VS_HIP(hipStreamSynchronize(s));
    trim();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed: %s", bytes,
                hipGetErrorString(e));
      return VS_ERR_NOMEM;
    }
  }
  if (zero) VS_HIP(hipMemsetAsync(p, 0, bytes, s));
  *out = p;
  return VS_OK;
}

void Pool::put(void *p, size_t bytes, bool clean) {
  if (!p) return;
  free_.emplace(bytes, std::make_pair(p, clean));
}

void Pool::trim() {
  for (auto &kv : free_) (void)hipFree(kv.second.first);
  free_.clear();
}

// On failure nothing stays allocated: d.off / d.adj are freed and reset, so
// a later free_csr (which frees them only when owned) cannot leak them.
static vs_status upload_csr(DevCSR &d, const uint64_t *off, const uint32_t *adj,
                            uint32_t V, uint64_t E) {
  const size_t ob = (static_cast<size_t>(V) + 1) * sizeof(uint64_t);
  hipError_t e = hipMalloc(&d.off, ob);
  if (e == hipSuccess) e = hipMalloc(&d.adj, std::max<uint64_t>(E, 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(d.off, off, ob, hipMemcpyHostToDevice);
  if (e == hipSuccess && E) e = hipMemcpy(d.adj, adj, E * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (d.off) (void)hipFree(d.off);
    if (d.adj) (void)hipFree(d.adj);
    d.off = nullptr;
    d.adj = nullptr;
    set_error("CSR upload (V=%u, E=%llu): %s", V, static_cast<unsigned long long>(E),
              hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? VS_ERR_NOMEM : VS_ERR_HIP;
  }
  d.E = E;
  d.hoff.assign(off, off + static_cast<size_t>(V) + 1);
  d.owned = true;
  return VS_OK;
}

static void free_csr(DevCSR &d) {
  if (d.owned) {
    (void)hipFree(d.off);
    (void)hipFree(d.adj);
  }
  if (d.pitems) (void)hipFree(d.pitems);
  if (d.hwords) (void)hipFree(d.hwords);
  if (d.fence) (void)hipFree(d.fence);
  if (d.hadj) (void)hipFree(d.hadj);
  if (d.hubcnt) (void)hipFree(d.hubcnt);
  d = DevCSR{};
}

// Hubs-first pull lists (see DevCSR::hadj), built on the host in parallel.
static vs_status build_hub_lists(DevCSR &c, uint32_t V, uint32_t hubs) {
  const uint64_t E = c.E;
  std::vector<uint32_t> adj(E);
  if (E) VS_HIP(hipMemcpy(adj.data(), c.adj, E * 4, hipMemcpyDeviceToHost));
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // appearances of each vertex over all lists
  std::vector<std::vector<uint32_t>> part(nt);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      std::vector<uint32_t> &h = part[t];
      h.assign(V, 0);
      for (uint64_t k = E * t / nt; k < E * (t + 1) / nt; ++k) h[adj[k]]++;
    });
  for (auto &x : th) x.join();
  th.clear();
  std::vector<uint64_t> cnt(V, 0);
  for (unsigned t = 0; t < nt; ++t)
    for (uint32_t v = 0; v < V; ++v) cnt[v] += part[t][v];
  std::vector<std::vector<uint32_t>>().swap(part);
  std::vector<uint32_t> order(V);
  for (uint32_t v = 0; v < V; ++v) order[v] = v;
  const uint32_t K = std::min<uint32_t>(hubs, V);
  std::nth_element(order.begin(), order.begin() + K, order.end(), [&](uint32_t x, uint32_t y) {
    return cnt[x] != cnt[y] ? cnt[x] > cnt[y] : x < y;
  });
  std::vector<uint8_t> hub(V, 0);
  for (uint32_t i = 0; i < K; ++i) hub[order[i]] = 1;
  std::vector<uint32_t> hcnt(V, 0);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (uint64_t v = static_cast<uint64_t>(V) * t / nt; v < static_cast<uint64_t>(V) * (t + 1) / nt; ++v) {
        uint32_t *b = adj.data() + c.hoff[v], *e = adj.data() + c.hoff[v + 1];
        uint32_t *mid = std::stable_partition(b, e, [&](uint32_t u) { return hub[u] != 0; });
        // most-listed hubs first: their rows are the densest, so the pull's
        // early exit fires after the fewest gathers
        std::sort(b, mid, [&](uint32_t x, uint32_t y) {
          return cnt[x] != cnt[y] ? cnt[x] > cnt[y] : x < y;
        });
        hcnt[v] = static_cast<uint32_t>(mid - b);
      }
    });
  for (auto &x : th) x.join();
  void *ph = nullptr, *pc = nullptr;
  VS_HIP(hipMalloc(&ph, std::max<uint64_t>(E, 1) * 4));
  hipError_t e = hipMalloc(&pc, std::max<uint32_t>(V, 1) * 4ull);
  if (e == hipSuccess && E) e = hipMemcpy(ph, adj.data(), E * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pc, hcnt.data(), V * 4ull, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(ph);
    (void)hipFree(pc);
    set_error("hub lists upload: %s", hipGetErrorString(e));
    return VS_ERR_HIP;
  }
  c.hadj = static_cast<uint32_t *>(ph);
  c.hubcnt = static_cast<uint32_t *>(pc);
  c.nhubs = hubs;
  return VS_OK;
}

vs_status ensure_pull_items(DevCSR &c, uint32_t V, uint32_t hubs) {
  if (hubs == 0) hubs = kPullHubs;
  if (c.pitems) {
    if (c.hadj && c.nhubs == hubs) return VS_OK;
    VS_HIP(hipDeviceSynchronize());  // no kernel may still read the old lists
    (void)hipFree(c.hadj);
    (void)hipFree(c.hubcnt);
    c.hadj = nullptr;
    c.hubcnt = nullptr;
    return build_hub_lists(c, V, hubs);
  }
  std::vector<unsigned long long> it;
  it.reserve(static_cast<size_t>(V) + c.E / kPullChunk + 1);
  std::vector<uint32_t> hw;
  bool heavy = false;
  // Every first chunk before any later one: items are strided over the
  // waves in list order, so a chunked vertex's first chunk (its hub
  // neighbours) has usually finished before its later chunks start, and
  // they skip the rows it already found.
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t d = c.hoff[v + 1] - c.hoff[v];
    if (d == 0) continue;
    it.push_back(v);
    if (d > kPullChunk) {
      heavy = true;
      if (hw.empty() || hw.back() != v / 64) hw.push_back(v / 64);
    }
  }
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t d = c.hoff[v + 1] - c.hoff[v];
    const uint64_t nch = (d + kPullChunk - 1) / kPullChunk;
    for (uint64_t q = 1; q < nch; ++q) it.push_back(v | (q << 32));
  }
  if (!hw.empty()) {
    void *ph = nullptr;
    VS_HIP(hipMalloc(&ph, hw.size() * 4));
    hipError_t e = hipMemcpy(ph, hw.data(), hw.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(ph);
      set_error("heavy words upload: %s", hipGetErrorString(e));
      return VS_ERR_HIP;
    }
    c.hwords = static_cast<uint32_t *>(ph);
    c.nhwords = static_cast<uint32_t>(hw.size());
  }
  void *p = nullptr;
  VS_HIP(hipMalloc(&p, std::max<size_t>(it.size(), 1) * 8));
  if (!it.empty()) {
    hipError_t e = hipMemcpy(p, it.data(), it.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      set_error("pull items upload: %s", hipGetErrorString(e));
      return VS_ERR_HIP;
    }
  }
  c.pitems = static_cast<unsigned long long *>(p);
  c.npitems = it.size();
  c.has_heavy = heavy;
  return build_hub_lists(c, V, hubs);
}

vs_status graph_csrs(vs_graph_s *g, int direction, const DevCSR **push,
                     const DevCSR **pull) {
  if (direction < VS_DIR_FORWARD || direction > VS_DIR_BOTH) {
    set_error("invalid direction %d", direction);
    return VS_ERR_INVALID;
  }
  if (g->symmetric) {
    *push = *pull = &g->fwd;
    return VS_OK;
  }
  if (!g->rev.owned) {
    std::vector<uint64_t> roff;
    std::vector<uint32_t> radj;
    reverse_csr(g->fwd.hoff.data(), g->hadj.data(), g->V, g->E, roff, radj);
    VS_TRY(upload_csr(g->rev, roff.data(), radj.data(), g->V, g->E));
  }
  if (direction == VS_DIR_FORWARD) {
    *push = &g->fwd;
    *pull = &g->rev;
  } else if (direction == VS_DIR_BACKWARD) {
    *push = &g->rev;
    *pull = &g->fwd;
  } else {
    if (!g->both.owned) {
      std::vector<uint64_t> roff, uoff;
      std::vector<uint32_t> radj, uadj;
      reverse_csr(g->fwd.hoff.data(), g->hadj.data(), g->V, g->E, roff, radj);
      union_csr(g->fwd.hoff.data(), g->hadj.data(), roff.data(), radj.data(),
                g->V, uoff, uadj);
      VS_TRY(upload_csr(g->both, uoff.data(), uadj.data(), g->V, uoff[g->V]));
    }
    *push = *pull = &g->both;
  }
  return VS_OK;
}

void graph_release(vs_graph_s *g) {
  if (--g->refs > 0) return;
  (void)hipSetDevice(g->device);
  (void)hipStreamSynchronize(g->stream);
  g->pool.trim();
  if (g->sched) sched_release(g->sched);
  g->sched = nullptr;
  free_csr(g->fwd);
  free_csr(g->rev);
  free_csr(g->both);
  (void)hipStreamDestroy(g->stream);
  delete g;
}

// --------------------------------------------------------------- kernels

// Per-row popcount with bit-sliced counters: each lane keeps NB counter
// planes c[k]; adding word x ripples a carry through the planes, so one
// 64-bit add counts 64 rows at once.  Planes are unpacked per chunk.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_rowcount(const uint64_t *__restrict__ M,
                                                    uint32_t V, uint32_t S,
                                                    uint32_t ntiles, ull *counts) {
  constexpr int G = 64 / TW;
  constexpr uint32_t CV = 2048;  // vertices per chunk (enough items to fill the GPU)
  constexpr int NB = 12;         // CV / G + 1 <= 2^12
  const int lane = lane_id(), wid = wave_id();
  const int g = lane / TW, j = lane % TW;
  const uint32_t nch = (V + CV - 1) / CV;
  const uint64_t total = static_cast<uint64_t>(ntiles) * nch;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < total; it += nwaves) {
    const uint32_t t = static_cast<uint32_t>(it / nch);
    const uint32_t ch = static_cast<uint32_t>(it % nch);
    const uint32_t v0 = ch * CV;
    const uint32_t v1 = min(V, v0 + CV);
    uint64_t c[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) c[k] = 0;
    const uint64_t *base = M + static_cast<size_t>(t) * V * TW + j;
    for (uint32_t v = v0 + g; v < v1; v += G) {
      uint64_t carry = base[static_cast<size_t>(v) * TW];
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const uint64_t t2 = c[k] & carry;
        c[k] ^= carry;
        carry = t2;
      }
    }
    const uint32_t w = t * TW + j;
    for (int b = 0; b < 64; ++b) {
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < NB; ++k) cnt |= static_cast<uint32_t>((c[k] >> b) & 1) << k;
#pragma unroll
      for (int m = TW; m < 64; m <<= 1) cnt += __shfl_xor(cnt, m);
      const uint64_t row = static_cast<uint64_t>(w) * 64 + b;
      if (g == 0 && cnt && row < S) atomicAdd(&counts[row], static_cast<ull>(cnt));
    }
  }
}

// Per-row digest: digest[r] = sum over set bits (r, c) of digest_hash(c),
// mod 2^64.  Lane b owns row bit b; the 64 words of a 64-vertex block are
// walked through scalar reads.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_rowdigest(const uint64_t *__restrict__ M,
                                                     uint32_t V, uint32_t S,
                                                     uint32_t ntiles, ull *dig) {
  const int lane = lane_id(), wid = wave_id();
  const uint32_t nvb = (V + 63) / 64;
  constexpr uint32_t BPW = 16;  // vertex blocks per work item
  const uint32_t nitem_v = (nvb + BPW - 1) / BPW;
  const uint64_t total = static_cast<uint64_t>(ntiles) * TW * nitem_v;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < total; it += nwaves) {
    const uint32_t iv = static_cast<uint32_t>(it % nitem_v);
    const uint32_t tj = static_cast<uint32_t>(it / nitem_v);
    const uint32_t t = tj / TW, j = tj % TW;
    const uint32_t w = t * TW + j;
    if (static_cast<uint64_t>(w) * 64 >= S) continue;
    const uint64_t *base = M + static_cast<size_t>(t) * V * TW + j;
    uint64_t acc = 0;
    for (uint32_t vb = iv * BPW; vb < min(nvb, (iv + 1) * BPW); ++vb) {
      const uint32_t v = vb * 64 + lane;
      const uint64_t mine = v < V ? base[static_cast<size_t>(v) * TW] : 0;
      uint64_t any = __ballot(mine != 0);
      while (any) {
        const int l = __ffsll(static_cast<long long>(any)) - 1;
        any &= any - 1;
        const uint64_t x = shfl64(mine, l);
        if ((x >> lane) & 1) acc += digest_hash(vb * 64 + l);
      }
    }
    const uint64_t row = static_cast<uint64_t>(w) * 64 + lane;
    if (acc && row < S) atomicAdd(&dig[row], static_cast<ull>(acc));
  }
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_popcount(const uint64_t *__restrict__ M,
                                                    size_t nwords, ull *out) {
  uint64_t acc = 0;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < nwords;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    acc += __popcll(M[i]);
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, static_cast<ull>(acc));
}

// Native (TW words per vertex) -> reference DVSM (16 words per vertex).
template <int TW>
__global__ void k_to_dvsm(const uint64_t *__restrict__ M, uint32_t V, uint32_t W,
                          uint64_t *__restrict__ out, size_t nout) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < nout;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t stack = i / (static_cast<size_t>(V) * 16);
    const size_t rem = i % (static_cast<size_t>(V) * 16);
    const size_t v = rem / 16, jj = rem % 16;
    const size_t w = stack * 16 + jj;
    uint64_t x = 0;
    if (w < W) x = M[((w / TW) * V + v) * TW + (w % TW)];
    out[i] = x;
  }
}

// Native -> row-major: a wave transposes one 64x64 bit block (64 vertices x
// 64 rows of word w) with 64 ballots; lane b keeps row b's word.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_to_rowmajor(const uint64_t *__restrict__ M,
                                                       uint32_t V, uint32_t S,
                                                       uint32_t W, uint64_t *out,
                                                       uint64_t pitch,
                                                       const uint32_t *__restrict__ dest) {
  const int lane = lane_id(), wid = wave_id();
  const uint32_t nvb = (V + 63) / 64;
  const uint64_t total = static_cast<uint64_t>(W) * nvb;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < total; it += nwaves) {
    const uint32_t w = static_cast<uint32_t>(it / nvb);
    const uint32_t vb = static_cast<uint32_t>(it % nvb);
    const uint32_t v = vb * 64 + lane;
    const uint64_t x =
        v < V ? M[(static_cast<size_t>(w / TW) * V + v) * TW + (w % TW)] : 0;
    uint64_t mine = 0;
    for (int b = 0; b < 64; ++b) {
      const uint64_t y = __ballot((x >> b) & 1);
      if (lane == b) mine = y;
    }
    const uint64_t row = static_cast<uint64_t>(w) * 64 + lane;
    if (row < S) out[(dest ? dest[row] : row) * pitch + vb] = mine;
  }
}

// Row packing of the plain engine: the rows of a result are stored sorted,
// within each tile (64 * TW rows), by the source's 2-hop volume, so rows
// that reach the graph at a similar pace share words.  Reading a row-indexed
// output puts them back into the caller's order on a copy: for caller row i
// of tile t, unpack[t * 64 TW + i] is its stored row within the tile.  One
// group of TW lanes per (tile, vertex): the wave stages its 64 words in LDS
// and lane j rebuilds word j from 64 dword reads.  The LDS byte address of
// every read (and the bit within the dword) is computed once per tile and
// kept in registers for the whole loop, so no register that addresses an LDS
// read in flight is rewritten while the read is pending.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_unpack_rows(uint64_t *M, uint32_t V, uint32_t T,
                                                       const uint16_t *__restrict__ unpack) {
  constexpr int G = 64 / TW;
  constexpr uint32_t RT = 64 * TW;
  __shared__ uint32_t s_words[kWavesPerBlock][128];  // each wave's 64 words, as dwords
  const int lane = lane_id(), wid = wave_id(), g = lane / TW, j = lane % TW;
  char *const lds = reinterpret_cast<char *>(&s_words[0][0]);
  const uint64_t vpt = (V + G - 1) / G;  // wave units per tile
  const uint64_t units = vpt * T;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  uint32_t cur = UINT32_MAX;
  uint32_t addr[64];  // byte offset in s_words of the dword holding output bit b
  uint32_t sh[16];    // bit of that dword, four per register
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(kWavesPerBlock) + wid; u < units;
       u += nwaves) {
    const uint32_t t = static_cast<uint32_t>(u / vpt);
    const uint32_t v = static_cast<uint32_t>(u % vpt) * G + g;
    if (t != cur) {  // wave-uniform
      const uint16_t *m = unpack + static_cast<size_t>(t) * RT + j * 64;
#pragma unroll
      for (int q = 0; q < 16; ++q) sh[q] = 0;
#pragma unroll
      for (int b = 0; b < 64; ++b) {
        const uint32_t i = m[b];  // packed row within the tile
        const uint32_t dw = (g * TW + (i >> 6)) * 2 + ((i >> 5) & 1);
        addr[b] = (static_cast<uint32_t>(wid) * 128 + dw) * 4;
        sh[b >> 2] |= (i & 31) << (8 * (b & 3));
      }
      cur = t;
    }
    const bool ok = v < V;
    uint64_t *p = M + (static_cast<size_t>(t) * V + v) * TW + j;
    const uint64_t x = ok ? *p : 0;
    if (__ballot(x != 0) == 0) continue;  // an all-zero wave unit stays as it is
    reinterpret_cast<uint2 *>(s_words[wid])[lane] =
        make_uint2(static_cast<uint32_t>(x), static_cast<uint32_t>(x >> 32));
    wave_sync();
    uint64_t out = 0;
#pragma unroll
    for (int b = 0; b < 64; ++b) {
      const uint32_t d = *reinterpret_cast<const uint32_t *>(lds + addr[b]);
      out |= static_cast<uint64_t>((d >> ((sh[b >> 2] >> (8 * (b & 3))) & 31)) & 1u) << b;
    }
    wave_sync();  // every read is done before the next unit's words land
    if (ok) *p = out;
  }
}



__global__ __launch_bounds__(kBlock) void k_copy_words(uint64_t *__restrict__ dst,
                                                      const uint64_t *__restrict__ src,
                                                      size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

#define VS_TW_DISPATCH(tw, KERNEL, ...)                                      \
  switch (tw) {                                                              \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;               \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;               \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;               \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;               \
    case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;             \
    case 32: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break;             \
    default: hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__); break;             \
  }

vs_status mat_row_counts(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                         uint32_t tw, uint32_t ntiles, uint64_t *d_counts) {
  VS_HIP(hipMemsetAsync(d_counts, 0, static_cast<size_t>(S) * 8, g->stream));
  if (V == 0) return VS_OK;
  const uint64_t items = static_cast<uint64_t>(ntiles) * ((V + 2047) / 2048);
  const uint32_t grid = grid_for(items, g->num_cus);
  VS_TW_DISPATCH(tw, k_rowcount, dim3(grid), dim3(kBlock), 0, g->stream, m.d,
                 V, S, ntiles, reinterpret_cast<ull *>(d_counts));
  VS_HIP(hipGetLastError());
  return VS_OK;
}

vs_status mat_row_digests(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                          uint32_t tw, uint32_t ntiles, uint64_t *d_digests) {
  VS_HIP(hipMemsetAsync(d_digests, 0, static_cast<size_t>(S) * 8, g->stream));
  if (V == 0) return VS_OK;
  const uint64_t items =
      static_cast<uint64_t>(ntiles) * tw * (((V + 63) / 64 + 15) / 16);
  const uint32_t grid = grid_for(items, g->num_cus);
  VS_TW_DISPATCH(tw, k_rowdigest, dim3(grid), dim3(kBlock), 0, g->stream, m.d,
                 V, S, ntiles, reinterpret_cast<ull *>(d_digests));
  VS_HIP(hipGetLastError());
  return VS_OK;
}

vs_status mat_to_rowmajor(vs_graph_s *g, const Mat &m, uint32_t S, uint32_t V,
                          uint32_t tw, uint64_t *out, uint64_t pitch, const uint32_t *d_dest) {
  const uint32_t W = (S + 63) / 64;
  const uint64_t items = static_cast<uint64_t>(W) * ((V + 63) / 64);
  if (items == 0) return VS_OK;
  const uint32_t grid = grid_for(items, g->num_cus);
  VS_TW_DISPATCH(tw, k_to_rowmajor, dim3(grid), dim3(kBlock), 0, g->stream, m.d, V,
                 S, W, out, pitch, d_dest);
  VS_HIP(hipGetLastError());
  return VS_OK;
}

vs_status mat_unpack_rows(vs_graph_s *g, uint64_t *M, uint32_t V, uint32_t T, uint32_t tw,
                          const uint16_t *d_unpack) {
  const uint64_t units = static_cast<uint64_t>(T) * ((V + 64 / tw - 1) / (64 / tw));
  if (units == 0) return VS_OK;
  VS_TW_DISPATCH(tw, k_unpack_rows, dim3(grid_for(units, g->num_cus)), dim3(kBlock), 0,
                 g->stream, M, V, T, d_unpack);
  VS_HIP(hipGetLastError());
  return VS_OK;
}

vs_status CallerOrder::open(vs_result_s *r, uint32_t i) {
  const Mat &m = r->mats[i];
  d = m.d;
  if (r->stored.empty() || !m.d || m.bytes == 0) return VS_OK;
  g_ = r->g;
  const uint32_t RT = 64 * r->tw;
  const size_t nu = static_cast<size_t>(r->ntiles) * RT;
  std::vector<uint16_t> unpack(nu);
  for (size_t q = 0; q < nu; ++q) unpack[q] = static_cast<uint16_t>(q % RT);  // padding rows
  for (uint32_t row = 0; row < r->S; ++row)
    unpack[row] = static_cast<uint16_t>(r->stored[row] % RT);
  void *du = nullptr;
  VS_TRY(g_->pool.get(nu * 2, false, g_->stream, &du));
  struct Back { Pool &p; void *q; size_t b; ~Back() { p.put(q, b, false); } } back{g_->pool, du, nu * 2};
  VS_TRY(g_->pool.get(m.bytes, false, g_->stream, &tmp_));
  bytes_ = m.bytes;
  VS_HIP(hipMemcpyAsync(du, unpack.data(), nu * 2, hipMemcpyHostToDevice, g_->stream));
  const size_t n = m.bytes / 8;
  hipLaunchKernelGGL(k_copy_words, dim3(grid_for((n + 255) / 256, g_->num_cus)), dim3(kBlock), 0,
                     g_->stream, static_cast<uint64_t *>(tmp_), m.d, n);
  VS_HIP(hipGetLastError());
  VS_TRY(mat_unpack_rows(g_, static_cast<uint64_t *>(tmp_), r->V, r->ntiles, r->tw,
                         static_cast<const uint16_t *>(du)));
  // the host table and the pooled one are released when this returns: wait
  VS_HIP(hipStreamSynchronize(g_->stream));
  d = static_cast<const uint64_t *>(tmp_);
  return VS_OK;
}

CallerOrder::~CallerOrder() {
  if (tmp_) g_->pool.put(tmp_, bytes_, false);
}

vs_status DestRows::open(vs_result_s *r) {
  d = nullptr;
  if (r->stored.empty() || r->S == 0) return VS_OK;
  g_ = r->g;
  std::vector<uint32_t> dest(r->S);
  for (uint32_t row = 0; row < r->S; ++row) dest[r->stored[row]] = row;
  bytes_ = static_cast<size_t>(r->S) * 4;
  VS_TRY(g_->pool.get(bytes_, false, g_->stream, &p_));
  VS_HIP(hipMemcpyAsync(p_, dest.data(), bytes_, hipMemcpyHostToDevice, g_->stream));
  VS_HIP(hipStreamSynchronize(g_->stream));  // dest leaves scope
  d = static_cast<const uint32_t *>(p_);
  return VS_OK;
}

DestRows::~DestRows() {
  if (p_) g_->pool.put(p_, bytes_, false);
}

} // namespace vs

using namespace vs;

// ---------------------------------------------------------------- C ABI
extern "C" {

vs_status vs_device_count(int *count) try {
  clear_error();
  if (!count) { set_error("vs_device_count: null"); return VS_ERR_INVALID; }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return VS_OK;
  }
  *count = n;
  return VS_OK;
} VS_CATCH("vs_device_count")

vs_status vs_set_device(int device) try {
  clear_error();
  VS_HIP(hipSetDevice(device));
  return VS_OK;
} VS_CATCH("vs_set_device")

vs_status vs_device_synchronize(void) try {
  clear_error();
  VS_HIP(hipDeviceSynchronize());
  return VS_OK;
} VS_CATCH("vs_device_synchronize")

vs_status vs_dev_alloc(size_t bytes, void **dptr) try {
  clear_error();
  if (!dptr) { set_error("vs_dev_alloc: null"); return VS_ERR_INVALID; }
  *dptr = nullptr;
  if (bytes == 0) return VS_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("vs_dev_alloc(%zu): %s", bytes, hipGetErrorString(e));
    return VS_ERR_NOMEM;
  }
  return VS_OK;
} VS_CATCH("vs_dev_alloc")

vs_status vs_dev_free(void *dptr) try {
  clear_error();
  if (dptr) VS_HIP(hipFree(dptr));
  return VS_OK;
} VS_CATCH("vs_dev_free")

vs_status vs_memcpy_h2d(void *dst, const void *src, size_t bytes) try {
  clear_error();
  if (bytes) VS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return VS_OK;
} VS_CATCH("vs_memcpy_h2d")

vs_status vs_memcpy_d2h(void *dst, const void *src, size_t bytes) try {
  clear_error();
  if (bytes) VS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return VS_OK;
} VS_CATCH("vs_memcpy_d2h")

vs_status vs_device_copy(void *dst, const void *src, size_t bytes) try {
  clear_error();
  if ((bytes % 8) || (bytes && (!dst || !src))) {
    set_error("vs_device_copy: bytes must be a multiple of 8 and pointers non-null");
    return VS_ERR_INVALID;
  }
  if (!bytes) return VS_OK;
  int dev = 0;
  VS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  VS_HIP(hipGetDeviceProperties(&prop, dev));
  const uint32_t grid = static_cast<uint32_t>(prop.multiProcessorCount) * 8;
  hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(kBlock), 0, nullptr,
                     static_cast<uint64_t *>(dst), static_cast<const uint64_t *>(src),
                     bytes / 8);
  VS_HIP(hipGetLastError());
  VS_HIP(hipDeviceSynchronize());
  return VS_OK;
} VS_CATCH("vs_device_copy")

// The kernels index adj through offsets and rows through adj without bounds
// checks of their own, so a CSR is checked once, before it reaches the device:
// offsets non-decreasing and every id below V (threads over ranges).
static vs_status check_csr(const uint64_t *off, const uint32_t *adj, uint32_t V, uint64_t E) {
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<uint64_t> bad_off(nt, UINT64_MAX), bad_id(nt, UINT64_MAX);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const uint64_t v0 = static_cast<uint64_t>(V) * t / nt, v1 = static_cast<uint64_t>(V) * (t + 1) / nt;
      for (uint64_t v = v0; v < v1; ++v)
        if (off[v + 1] < off[v]) { bad_off[t] = v; break; }
      const uint64_t e0 = E * t / nt, e1 = E * (t + 1) / nt;
      for (uint64_t e = e0; e < e1; ++e)
        if (adj[e] >= V) { bad_id[t] = e; break; }
    });
  for (auto &x : th) x.join();
  const uint64_t bo = *std::min_element(bad_off.begin(), bad_off.end());
  const uint64_t bi = *std::min_element(bad_id.begin(), bad_id.end());
  if (bo != UINT64_MAX) {
    set_error("vs_graph_create: offsets decrease at vertex %llu", (unsigned long long)bo);
    return VS_ERR_INVALID;
  }
  if (bi != UINT64_MAX) {
    set_error("vs_graph_create: neighbour id %u >= V=%u at entry %llu", adj[bi], V,
              (unsigned long long)bi);
    return VS_ERR_INVALID;
  }
  return VS_OK;
}

vs_status vs_graph_create(const uint64_t *offsets, const uint32_t *adj,
                          uint32_t num_vertices, uint64_t num_edges,
                          uint32_t flags, vs_graph_t *out) try {
  clear_error();
  if (!out || !offsets || (num_edges > 0 && !adj) || num_vertices == 0) {
    set_error("vs_graph_create: invalid arguments");
    return VS_ERR_INVALID;
  }
  if (offsets[0] != 0 || offsets[num_vertices] != num_edges) {
    set_error("vs_graph_create: offsets[0]=%llu offsets[V]=%llu, E=%llu",
              (unsigned long long)offsets[0],
              (unsigned long long)offsets[num_vertices],
              (unsigned long long)num_edges);
    return VS_ERR_INVALID;
  }
  VS_TRY(check_csr(offsets, adj, num_vertices, num_edges));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    set_error("vs_graph_create: no GPU visible");
    return VS_ERR_NODEVICE;
  }
  std::unique_ptr<vs_graph_s> g(new vs_graph_s);
  VS_HIP(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  VS_HIP(hipGetDeviceProperties(&prop, g->device));
  g->num_cus = prop.multiProcessorCount;
  g->V = num_vertices;
  g->E = num_edges;
  g->symmetric = (flags & VS_GRAPH_SYMMETRIC) != 0;
  VS_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  vs_status s = upload_csr(g->fwd, offsets, adj, num_vertices, num_edges);
  if (s != VS_OK) {
    free_csr(g->fwd);
    (void)hipStreamDestroy(g->stream);
    return s;
  }
  if (!g->symmetric) g->hadj.assign(adj, adj + num_edges);
  // the intersection fence of the forward lists (E / 64 ids) is built here,
  // once, so intersections never allocate it later; if this fails the first
  // vs_mintersect call builds it instead
  if (build_fence(g.get(), g->fwd) != VS_OK) clear_error();
  *out = g.release();
  return VS_OK;
} VS_CATCH("vs_graph_create")

vs_status vs_graph_destroy(vs_graph_t g) try {
  clear_error();
  if (!g) return VS_OK;
  if (g->destroyed) { set_error("vs_graph_destroy: double destroy"); return VS_ERR_INVALID; }
  g->destroyed = true;
  graph_release(g);
  return VS_OK;
} VS_CATCH("vs_graph_destroy")

vs_status vs_graph_info(vs_graph_t g, uint32_t *num_vertices, uint64_t *num_edges,
                        uint64_t *device_bytes) try {
  clear_error();
  if (!g) { set_error("vs_graph_info: null"); return VS_ERR_INVALID; }
  if (num_vertices) *num_vertices = g->V;
  if (num_edges) *num_edges = g->E;
  if (device_bytes) {
    uint64_t b = 0;
    for (const DevCSR *c : {&g->fwd, &g->rev, &g->both})
      if (c->owned) b += (static_cast<uint64_t>(g->V) + 1) * 8 + c->E * 4;
    *device_bytes = b;
  }
  return VS_OK;
} VS_CATCH("vs_graph_info")

vs_status vs_graph_trim(vs_graph_t g) try {
  clear_error();
  if (!g) { set_error("vs_graph_trim: null"); return VS_ERR_INVALID; }
  VS_HIP(hipStreamSynchronize(g->stream));
  g->pool.trim();
  return VS_OK;
} VS_CATCH("vs_graph_trim")

vs_status vs_graph_set_tuning(vs_graph_t g, const vs_tuning *t) try {
  clear_error();
  if (!g || !t) { set_error("vs_graph_set_tuning: null"); return VS_ERR_INVALID; }
  if (t->tile_words > 64 || (t->tile_words & (t->tile_words - 1))) {
    set_error("vs_graph_set_tuning: tile_words %u is not a power of two <= 64",
              t->tile_words);
    return VS_ERR_INVALID;
  }
  g->tuning = *t;
  return VS_OK;
} VS_CATCH("vs_graph_set_tuning")

vs_status vs_graph_get_tuning(vs_graph_t g, vs_tuning *t) try {
  clear_error();
  if (!g || !t) { set_error("vs_graph_get_tuning: null"); return VS_ERR_INVALID; }
  *t = g->tuning;
  return VS_OK;
} VS_CATCH("vs_graph_get_tuning")

// ------------------------------------------------------------- results
static vs_status check_mat(vs_result_t r, uint32_t i) {
  if (!r) { set_error("null result"); return VS_ERR_INVALID; }
  if (i >= r->mats.size()) {
    set_error("matrix index %u out of range (%zu matrices)", i, r->mats.size());
    return VS_ERR_INVALID;
  }
  return VS_OK;
}

vs_status vs_result_info(vs_result_t r, uint32_t *rows, uint32_t *cols,
                         uint32_t *num_matrices, uint32_t *k_actual_max) try {
  clear_error();
  if (!r) { set_error("vs_result_info: null"); return VS_ERR_INVALID; }
  if (rows) *rows = r->S;
  if (cols) *cols = r->V;
  if (num_matrices) *num_matrices = static_cast<uint32_t>(r->mats.size());
  if (k_actual_max) *k_actual_max = r->k_actual_max;
  return VS_OK;
} VS_CATCH("vs_result_info")

vs_status vs_result_distance(vs_result_t r, uint32_t i, uint32_t *distance) try {
  clear_error();
  VS_TRY(check_mat(r, i));
  if (distance) *distance = r->mats[i].distance;
  return VS_OK;
} VS_CATCH("vs_result_distance")

vs_status vs_result_count(vs_result_t r, uint32_t i, uint64_t *count) try {
  clear_error();
  VS_TRY(check_mat(r, i));
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  void *d = nullptr;
  VS_TRY(g->pool.get(8, true, g->stream, &d));
  const size_t nwords = r->mats[i].bytes / 8;
  const uint32_t grid = grid_for((nwords + 63) / 64, g->num_cus);
  hipLaunchKernelGGL(k_popcount<1>, dim3(grid), dim3(kBlock), 0, g->stream,
                     r->mats[i].d, nwords, reinterpret_cast<ull *>(d));
  VS_HIP(hipGetLastError());
  uint64_t h = 0;
  VS_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, g->stream));
  VS_HIP(hipStreamSynchronize(g->stream));
  g->pool.put(d, 8, false);
  if (count) *count = h;
  return VS_OK;
} VS_CATCH("vs_result_count")

static vs_status row_reduce(vs_result_t r, uint32_t i, uint64_t *host, bool digest) {
  VS_TRY(check_mat(r, i));
  if (!host) { set_error("null output"); return VS_ERR_INVALID; }
  if (r->S == 0) return VS_OK;
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  const size_t bytes = static_cast<size_t>(r->S) * 8;
  void *d = nullptr;
  VS_TRY(g->pool.get(bytes, false, g->stream, &d));
  vs_status s = digest ? mat_row_digests(g, r->mats[i], r->S, r->V, r->tw, r->ntiles,
                                         static_cast<uint64_t *>(d))
                       : mat_row_counts(g, r->mats[i], r->S, r->V, r->tw, r->ntiles,
                                        static_cast<uint64_t *>(d));
  std::vector<uint64_t> tmp;
  uint64_t *dst = host;
  if (!r->stored.empty()) {  // stored row order: read, then reorder
    tmp.resize(r->S);
    dst = tmp.data();
  }
  if (s == VS_OK) {
    hipError_t e = hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
      set_error("row readout: %s", hipGetErrorString(e));
      s = VS_ERR_HIP;
    }
  }
  g->pool.put(d, bytes, false);
  if (s == VS_OK && !tmp.empty())
    for (uint32_t row = 0; row < r->S; ++row) host[row] = tmp[r->stored[row]];
  return s;
}

vs_status vs_result_row_counts(vs_result_t r, uint32_t i, uint64_t *counts) try {
  clear_error();
  return row_reduce(r, i, counts, false);
} VS_CATCH("vs_result_row_counts")

vs_status vs_result_row_digests(vs_result_t r, uint32_t i, uint64_t *digests) try {
  clear_error();
  return row_reduce(r, i, digests, true);
} VS_CATCH("vs_result_row_digests")

static size_t export_bytes(vs_result_t r, uint32_t i, int32_t layout) {
  const uint64_t V = r->V;
  switch (layout) {
    case VS_LAYOUT_NATIVE: return r->mats[i].bytes;
    case VS_LAYOUT_DVSM: {
      const uint64_t stacks = (static_cast<uint64_t>(r->S) + 1023) / 1024;
      uint64_t words = stacks * 16 * V;
      words = (words + 7) / 8 * 8;
      return words * 8;
    }
    case VS_LAYOUT_ROWMAJOR: return static_cast<size_t>(r->S) * ((V + 63) / 64) * 8;
    default: return 0;
  }
}

vs_status vs_result_export_size(vs_result_t r, uint32_t i, int32_t layout,
                                size_t *bytes) try {
  clear_error();
  VS_TRY(check_mat(r, i));
  if (layout < VS_LAYOUT_NATIVE || layout > VS_LAYOUT_ROWMAJOR) {
    set_error("unknown layout %d", layout);
    return VS_ERR_INVALID;
  }
  if (bytes) *bytes = export_bytes(r, i, layout);
  return VS_OK;
} VS_CATCH("vs_result_export_size")

vs_status vs_result_export(vs_result_t r, uint32_t i, int32_t layout,
                           void *host_buf, size_t bytes) try {
  clear_error();
  VS_TRY(check_mat(r, i));
  if (layout < VS_LAYOUT_NATIVE || layout > VS_LAYOUT_ROWMAJOR) {
    set_error("unknown layout %d", layout);
    return VS_ERR_INVALID;
  }
  const size_t need = export_bytes(r, i, layout);
  if (bytes < need) {
    set_error("export buffer %zu bytes < %zu", bytes, need);
    return VS_ERR_CAPACITY;
  }
  if (need == 0) return VS_OK;
  if (!host_buf) { set_error("null export buffer"); return VS_ERR_INVALID; }
  vs_graph_s *g = r->g;
  VS_HIP(hipSetDevice(g->device));
  const Mat &m = r->mats[i];
  // word layouts need the rows in caller order; the row-major one writes
  // every row to its caller position directly
  CallerOrder co;
  DestRows dr;
  if (layout == VS_LAYOUT_ROWMAJOR) VS_TRY(dr.open(r));
  else VS_TRY(co.open(r, i));
  if (layout == VS_LAYOUT_NATIVE || (layout == VS_LAYOUT_DVSM && r->tw == 16 &&
                                     m.bytes == need)) {
    VS_HIP(hipMemcpyAsync(host_buf, co.d, need, hipMemcpyDeviceToHost, g->stream));
    VS_HIP(hipStreamSynchronize(g->stream));
    return VS_OK;
  }
  void *d = nullptr;
  VS_TRY(g->pool.get(need, false, g->stream, &d));
  const uint32_t W = (r->S + 63) / 64;
  if (layout == VS_LAYOUT_DVSM) {
    const size_t nout = need / 8;
    const uint32_t grid = grid_for((nout + 63) / 64, g->num_cus);
    VS_TW_DISPATCH(r->tw, k_to_dvsm, dim3(grid), dim3(kBlock), 0, g->stream, co.d,
                   r->V, W, static_cast<uint64_t *>(d), nout);
  } else {
    const vs_status st = mat_to_rowmajor(g, m, r->S, r->V, r->tw,
                                         static_cast<uint64_t *>(d), (r->V + 63) / 64, dr.d);
    if (st != VS_OK) {
      g->pool.put(d, need, false);
      return st;
    }
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host_buf, d, need, hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  g->pool.put(d, need, false);
  if (e != hipSuccess) {
    set_error("export: %s", hipGetErrorString(e));
    return VS_ERR_HIP;
  }
  return VS_OK;
} VS_CATCH("vs_result_export")

vs_status vs_result_get(vs_result_t r, uint32_t i, uint32_t row, uint32_t col,
                        int *bit) try {
  clear_error();
  VS_TRY(check_mat(r, i));
  if (row >= r->S || col >= r->V || !bit) {
    set_error("vs_result_get: (%u,%u) outside %ux%u", row, col, r->S, r->V);
    return VS_ERR_INVALID;
  }
  const uint32_t sr = r->stored.empty() ? row : r->stored[row];  // stored row
  const uint32_t w = sr / 64, t = w / r->tw, j = w % r->tw;
  const size_t idx = (static_cast<size_t>(t) * r->V + col) * r->tw + j;
  uint64_t word = 0;
  VS_HIP(hipSetDevice(r->g->device));
  VS_HIP(hipStreamSynchronize(r->g->stream));
  VS_HIP(hipMemcpy(&word, r->mats[i].d + idx, 8, hipMemcpyDeviceToHost));
  *bit = static_cast<int>((word >> (sr % 64)) & 1);
  return VS_OK;
} VS_CATCH("vs_result_get")

vs_status vs_result_free(vs_result_t r) try {
  clear_error();
  if (!r) return VS_OK;
  vs_graph_s *g = r->g;
  if (g) {
    (void)hipSetDevice(g->device);
    for (auto &m : r->mats) g->pool.put(m.d, m.bytes, false);
    graph_release(g);
  }
  delete r;
  return VS_OK;
} VS_CATCH("vs_result_free")

} // extern "C"
