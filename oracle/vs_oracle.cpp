// CPU oracle for the VExpand / Mintersect hot path.  TEST INFRASTRUCTURE:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// as the checker or the timed CPU baseline, never as the product path.
//
// Two independent restatements of the reference behaviour:
//
//  A. orc_vexpand_rows -- the semantics, one source row at a time: level-
//     synchronous BFS (Shortest) or walk-frontier expansion (Any) over plain
//     per-row state, then the union / per-layer export of VExpand:
//       ExpandStatus::init_first_m_layer   src/include/operator/expand_impl.hh:250-277
//       Expand<DVSM>::expand_dense          src/objects/operator/expand.cpp:210-344
//       dedup_and_update_visit_and_count    src/objects/storage/bitmap.cpp:1054-1096
//       VExpand::after_child (early stop)   src/objects/operator/vexpand.cpp:222-284
//       VExpand::merge_to_MT / copy_to_multi_MT  vexpand.cpp:298-378
//
//  B. orc_vexpand_dvsm -- the reference's shipped CPU algorithm, restated:
//     1024-row stacks in the DenseVarStackMatrix word layout
//     (bitmap.hh:329-425, bitmap.cpp:910-937), edges partitioned by
//     destination into blocks and Hilbert-ordered inside each block
//     (graph_loader.hh:110-139, 239-276), OpenMP schedule(guided) over the
//     blocks with 16-word (128-byte) column ORs per edge per stack
//     (expand.cpp:263-311, bitmap.cpp:1196-1214), the dedup/visit pass and a
//     pairwise union of the layers (runtime_concepts.hh:70-83).  This is the
//     CPU baseline; it must agree with A bit for bit.
//
//  C. orc_mintersect_sorted / orc_mintersect_bitmap -- the multi-way
//     intersection as std::set_intersection chains, and as the reference's
//     bitmap AND of column-major columns (mintersect.cpp:139-241,
//     bitmap.cpp:814-880) for small V.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

inline uint64_t nwords(uint64_t V) { return (V + 63) / 64; }

// Hilbert index of (x, y) on an n x n grid (n a power of two), the classic
// rotate-and-accumulate walk.
uint64_t hilbert_xy2d(uint64_t n, uint64_t x, uint64_t y) {
  uint64_t d = 0;
  for (uint64_t s = n / 2; s > 0; s /= 2) {
    const uint64_t rx = (x & s) > 0;
    const uint64_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

}  // namespace

extern "C" {

// ----------------------------------------------------------------- oracle A
// off/adj: CSR of the expansion direction (u -> adj[u]).
// path: 0 = Any, 1 = Shortest.  Outputs (row-major, nw = ceil(V/64) words per
// row):
//   with_distance == 0: out holds one S x nw matrix (the merged result)
//   with_distance == 1: out holds (k_actual - k_min + 1) matrices, caller
//                       sizes it for k_max - k_min + 1 (unused ones zero)
// te = sum_{i=1..k_actual} sum_r sum_{u in L_{i-1}(r)} deg(u)
// layer_pop[i] = total popcount of L_i over all rows.
int orc_vexpand_rows(const uint64_t *off, const uint32_t *adj, uint32_t V,
                     const uint32_t *src, uint32_t S, uint32_t k_min,
                     uint32_t k_max, int path, int with_distance, uint64_t *out,
                     uint32_t *k_actual_out, uint64_t *te_out,
                     uint64_t *layer_pop /* k_max+1 entries, may be null */) {
  const uint64_t nw = nwords(V);
  const uint32_t L = k_max + 1;
  // layers[r][i] as bitsets
  std::vector<uint64_t> layers(static_cast<size_t>(S) * L * nw, 0);
  auto LAY = [&](uint32_t r, uint32_t i) {
    return layers.data() + (static_cast<size_t>(r) * L + i) * nw;
  };
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t rr = 0; rr < S; ++rr) {
    const uint32_t r = static_cast<uint32_t>(rr);
    uint64_t *l0 = LAY(r, 0);
    l0[src[r] / 64] |= 1ull << (src[r] % 64);
    std::vector<uint64_t> vis;
    if (path == 1) vis.assign(l0, l0 + nw);
    std::vector<uint32_t> frontier{src[r]}, next;
    for (uint32_t i = 1; i <= k_max; ++i) {
      uint64_t *li = LAY(r, i);
      next.clear();
      for (uint32_t u : frontier) {
        for (uint64_t k = off[u]; k < off[u + 1]; ++k) {
          const uint32_t v = adj[k];
          const uint64_t bit = 1ull << (v % 64);
          if (path == 1 && (vis[v / 64] & bit)) continue;
          if (li[v / 64] & bit) continue;
          li[v / 64] |= bit;
          next.push_back(v);
        }
      }
      if (path == 1)
        for (uint32_t v : next) vis[v / 64] |= 1ull << (v % 64);
      frontier.swap(next);
    }
  }
  // k_actual: Shortest stops after the first hop with no new visits
  uint32_t k_actual = k_max;
  std::vector<uint64_t> pop(L, 0);
  for (uint32_t i = 0; i <= k_max; ++i)
    for (uint32_t r = 0; r < S; ++r) {
      const uint64_t *li = LAY(r, i);
      for (uint64_t w = 0; w < nw; ++w) pop[i] += __builtin_popcountll(li[w]);
    }
  if (path == 1) {
    for (uint32_t i = 1; i <= k_max; ++i)
      if (pop[i] == 0) {
        k_actual = i;
        break;
      }
  }
  if (layer_pop)
    for (uint32_t i = 0; i <= k_max; ++i) layer_pop[i] = pop[i];
  uint64_t te = 0;
  for (uint32_t i = 1; i <= k_actual; ++i)
    for (uint32_t r = 0; r < S; ++r) {
      const uint64_t *lp = LAY(r, i - 1);
      for (uint64_t w = 0; w < nw; ++w)
        for (uint64_t m = lp[w]; m; m &= m - 1) {
          const uint64_t u = w * 64 + __builtin_ctzll(m);
          te += off[u + 1] - off[u];
        }
    }
  if (with_distance) {
    const size_t mat = static_cast<size_t>(S) * nw;
    std::memset(out, 0, mat * (k_max >= k_min ? k_max - k_min + 1 : 0) * 8);
    for (uint32_t i = k_min; i <= k_actual; ++i)
      for (uint32_t r = 0; r < S; ++r)
        std::memcpy(out + (i - k_min) * mat + static_cast<size_t>(r) * nw, LAY(r, i), nw * 8);
  } else {
    std::memset(out, 0, static_cast<size_t>(S) * nw * 8);
    for (uint32_t r = 0; r < S; ++r)
      for (uint32_t i = k_min; i <= k_actual; ++i) {
        const uint64_t *li = LAY(r, i);
        uint64_t *o = out + static_cast<size_t>(r) * nw;
        for (uint64_t w = 0; w < nw; ++w) o[w] |= li[w];
      }
  }
  if (k_actual_out) *k_actual_out = k_actual;
  if (te_out) *te_out = te;
  return 0;
}

// Compulsory-traffic bytes of a VExpand run from its layers (row-major,
// (k_actual+1) matrices S x nw) and the final result R (S x nw).
//   B_i = 8 nz(F_i) + 8 nz(L_i) + [Shortest] 8 (nz(Vis_{i-1}) + nz(Vis_i))
//       + [Any and k_min <= i] 8 (nz(R_{i-1}) + nz(R_i))
//       + sum_tiles sum_{u in act(F_i restricted to the tile)} (4 deg(u) + 16)
//   with F_i = L_{i-1};  B = sum_i B_i + 8 nz(R_final).
// nz(X) counts nonzero words of the 64-row word grid; a tile is tile_rows
// consecutive source rows that are expanded together (each tile reads the
// adjacency of its own active vertices).
uint64_t orc_bytes_alg(const uint64_t *off, uint32_t V, uint32_t S,
                       const uint64_t *layers, uint32_t k_actual, uint32_t k_min,
                       int path, const uint64_t *R, uint32_t tile_rows) {
  const uint64_t nw = nwords(V);
  const uint64_t W = (S + 63) / 64;
  const size_t mat = static_cast<size_t>(S) * nw;
  auto nz = [&](const uint64_t *X) {
    uint64_t n = 0;
    std::vector<uint64_t> acc(nw);
    for (uint64_t w = 0; w < W; ++w) {
      std::fill(acc.begin(), acc.end(), 0);
      for (uint64_t r = w * 64; r < std::min<uint64_t>(S, w * 64 + 64); ++r)
        for (uint64_t k = 0; k < nw; ++k) acc[k] |= X[r * nw + k];
      for (uint64_t k = 0; k < nw; ++k) n += __builtin_popcountll(acc[k]);
    }
    return n;
  };
  auto act = [&](const uint64_t *X, uint64_t *cnt, uint64_t *degsum) {
    *cnt = 0;
    *degsum = 0;
    std::vector<uint64_t> acc(nw);
    for (uint64_t r0 = 0; r0 < S; r0 += tile_rows) {
      std::fill(acc.begin(), acc.end(), 0);
      for (uint64_t r = r0; r < std::min<uint64_t>(S, r0 + tile_rows); ++r)
        for (uint64_t k = 0; k < nw; ++k) acc[k] |= X[r * nw + k];
      for (uint64_t k = 0; k < nw; ++k)
        for (uint64_t m = acc[k]; m; m &= m - 1) {
          const uint64_t u = k * 64 + __builtin_ctzll(m);
          *cnt += 1;
          *degsum += off[u + 1] - off[u];
        }
    }
  };
  std::vector<uint64_t> vis(mat, 0), res(mat, 0), tmp(mat);
  auto orin = [&](std::vector<uint64_t> &dst, const uint64_t *X) {
    for (size_t q = 0; q < mat; ++q) dst[q] |= X[q];
  };
  orin(vis, layers);
  if (k_min == 0) orin(res, layers);
  uint64_t nz_vis = nz(vis.data()), nz_res = nz(res.data());
  uint64_t B = 0;
  for (uint32_t i = 1; i <= k_actual; ++i) {
    const uint64_t *F = layers + (i - 1) * mat;
    const uint64_t *Li = layers + i * mat;
    uint64_t c, d;
    act(F, &c, &d);
    B += 8 * nz(F) + 8 * nz(Li) + 4 * d + 16 * c;
    orin(vis, Li);
    const uint64_t nz_vis2 = nz(vis.data());
    if (path == 1) B += 8 * (nz_vis + nz_vis2);
    nz_vis = nz_vis2;
    if (path == 0 && i >= k_min) {
      orin(res, Li);
      const uint64_t nz_res2 = nz(res.data());
      B += 8 * (nz_res + nz_res2);
      nz_res = nz_res2;
    }
  }
  B += 8 * nz(R);
  return B;
}

// ----------------------------------------------------------------- oracle B
// Reference-algorithm restatement.  Output: the result in the DVSM layout
// (ceil(S/1024) stacks x 16*V words, total aligned to 8 words), i.e. exactly
// what DenseVarStackMatrix holds.  with_distance is not supported here (the
// per-layer export is covered by oracle A).
//
// Preprocessing (edge blocks) is not timed, like the reference's offline
// import.  *ms covers layer allocation + zeroing, the hop loop, the dedup
// pass and the union -- VExpand's own run time minus its dependencies.
struct EdgeBlocks {
  std::vector<uint32_t> u, v;       // flat pairs, block-contiguous
  std::vector<uint64_t> bstart;     // block b = [bstart[b], bstart[b+1])
};

void *orc_blocks_build(const uint64_t *off, const uint32_t *adj, uint32_t V,
                       uint32_t nblocks) {
  auto *eb = new EdgeBlocks;
  const uint64_t E = off[V];
  // in-degree per destination
  std::vector<uint64_t> indeg(V + 1, 0);
  for (uint64_t k = 0; k < E; ++k) indeg[adj[k] + 1]++;
  for (uint32_t i = 0; i < V; ++i) indeg[i + 1] += indeg[i];
  // destination ranges with ~E/nblocks edges each (partition_by_dst)
  std::vector<uint32_t> cut{0};
  const uint64_t target = std::max<uint64_t>(1, E / std::max<uint32_t>(nblocks, 1));
  for (uint32_t v = 0; v < V; ++v)
    if (indeg[v + 1] - indeg[cut.back()] >= target) cut.push_back(v + 1);
  if (cut.back() != V) cut.push_back(V);
  const size_t nb = cut.size() - 1;
  eb->u.resize(E);
  eb->v.resize(E);
  eb->bstart.assign(nb + 1, 0);
  for (size_t b = 0; b < nb; ++b) eb->bstart[b + 1] = indeg[cut[b + 1]];
  // scatter by destination
  std::vector<uint64_t> cur(indeg.begin(), indeg.end() - 1);
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t k = off[u]; k < off[u + 1]; ++k) {
      const uint64_t p = cur[adj[k]]++;
      eb->u[p] = u;
      eb->v[p] = adj[k];
    }
  // Hilbert order inside each block (sort_edge_in_peano2)
  uint64_t n = 1;
  while (n < V) n <<= 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < static_cast<int64_t>(nb); ++b) {
    const uint64_t s = eb->bstart[b], e = eb->bstart[b + 1];
    std::vector<std::pair<uint64_t, uint64_t>> key(e - s);
    for (uint64_t q = s; q < e; ++q)
      key[q - s] = {hilbert_xy2d(n, eb->u[q], eb->v[q]),
                    (static_cast<uint64_t>(eb->u[q]) << 32) | eb->v[q]};
    std::sort(key.begin(), key.end());
    for (uint64_t q = s; q < e; ++q) {
      eb->u[q] = static_cast<uint32_t>(key[q - s].second >> 32);
      eb->v[q] = static_cast<uint32_t>(key[q - s].second);
    }
  }
  return eb;
}

void orc_blocks_free(void *p) { delete static_cast<EdgeBlocks *>(p); }

uint64_t orc_blocks_count(void *p) {
  return static_cast<EdgeBlocks *>(p)->bstart.size() - 1;
}

static constexpr uint64_t kStack = 1024;
static constexpr uint64_t kWpc = kStack / 64;  // words per stack column = 16
static constexpr uint64_t kBlockPrefetch = 20;  // BLOCK_PREFETCH_COUNT, common.hh:63

uint64_t orc_dvsm_words(uint32_t S, uint32_t V) {
  const uint64_t stacks = (S + kStack - 1) / kStack;
  return (stacks * kWpc * V + 7) / 8 * 8;
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static void or_col(uint64_t *__restrict o, const uint64_t *__restrict b) {
  for (int q = 0; q < 16; ++q) o[q] |= b[q];
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static uint64_t dedup_range(uint64_t *__restrict vis, uint64_t *__restrict o, size_t n) {
  uint64_t c = 0;
  for (size_t q = 0; q < n; ++q) {
    const uint64_t x = vis[q] & o[q];
    o[q] ^= x;
    vis[q] |= o[q];
    c += __builtin_popcountll(vis[q]);
  }
  return c;
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static void or_range(uint64_t *__restrict d, const uint64_t *__restrict s, size_t n) {
  for (size_t q = 0; q < n; ++q) d[q] |= s[q];
}

int orc_vexpand_dvsm(void *blocks, const uint64_t *off, uint32_t V,
                     const uint32_t *src, uint32_t S, uint32_t k_min,
                     uint32_t k_max, int path, uint64_t *out /* dvsm words */,
                     uint32_t *k_actual_out, uint64_t *visit_count_out,
                     double *ms) {
  (void)off;
  auto *eb = static_cast<EdgeBlocks *>(blocks);
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t words = orc_dvsm_words(S, V);
  const uint64_t stacks = (S + kStack - 1) / kStack;
  const uint64_t v_per_stack = kWpc * V;
  const uint32_t nlayers = std::max<uint32_t>(3, k_max) + 1;
  auto alloc0 = [&]() {
    uint64_t *p = static_cast<uint64_t *>(std::aligned_alloc(64, words * 8));
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < static_cast<int64_t>(words); ++q) p[q] = 0;
    return p;
  };
  std::vector<uint64_t *> layer(nlayers, nullptr);
  for (uint32_t i = 0; i < nlayers; ++i) layer[i] = alloc0();
  uint64_t *vis = path == 1 ? alloc0() : nullptr;
  auto idx = [&](uint64_t r, uint64_t c) {
    return (r / kStack) * v_per_stack + c * kWpc + (r % kStack) / 64;
  };
  for (uint32_t r = 0; r < S; ++r) {
    layer[0][idx(r, src[r])] |= 1ull << (r % 64);
    if (vis) vis[idx(r, src[r])] |= 1ull << (r % 64);
  }
  uint64_t visit_count = S, last_visit = S;
  uint32_t k_actual = k_max;
  const int64_t nb = static_cast<int64_t>(eb->bstart.size() - 1);
  for (uint32_t i = 1; i <= k_max; ++i) {
    uint64_t *base = layer[i - 1], *o = layer[i];
#pragma omp parallel for schedule(guided)
    for (int64_t b = 0; b < nb; ++b) {
      const uint64_t s = eb->bstart[b], e = eb->bstart[b + 1];
      for (uint64_t si = 0; si < stacks; ++si) {
        uint64_t *ob = o + si * v_per_stack;
        const uint64_t *bb = base + si * v_per_stack;
        // BLOCK_PREFETCH (common.hh:52-63, bitmap.cpp:1172-1194): both
        // 128-byte columns of the next 20 edges are prefetched (destination
        // for writing), then ORed
        for (uint64_t q0 = s; q0 < e; q0 += kBlockPrefetch) {
          const uint64_t q1 = std::min<uint64_t>(e, q0 + kBlockPrefetch);
          for (uint64_t q = q0; q < q1; ++q) {
            const uint64_t *po = ob + static_cast<uint64_t>(eb->v[q]) * kWpc;
            const uint64_t *pb = bb + static_cast<uint64_t>(eb->u[q]) * kWpc;
            __builtin_prefetch(po, 1);
            __builtin_prefetch(po + 8, 1);
            __builtin_prefetch(pb, 0);
            __builtin_prefetch(pb + 8, 0);
          }
          for (uint64_t q = q0; q < q1; ++q)
            or_col(ob + static_cast<uint64_t>(eb->v[q]) * kWpc,
                   bb + static_cast<uint64_t>(eb->u[q]) * kWpc);
        }
      }
    }
    if (path == 1) {
      uint64_t cnt = 0;
      const int64_t chunk = 1 << 16;
      const int64_t nchunk = (static_cast<int64_t>(words) + chunk - 1) / chunk;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
      for (int64_t c = 0; c < nchunk; ++c) {
        const size_t s0 = static_cast<size_t>(c) * chunk;
        const size_t n = std::min<size_t>(chunk, words - s0);
        cnt += dedup_range(vis + s0, o + s0, n);
      }
      visit_count = cnt;
      if (visit_count == last_visit) {
        k_actual = i;
        break;
      }
      last_visit = visit_count;
    }
  }
  // union (merge_to_MT): visit itself for Shortest with k_min == 0
  if (path == 1 && k_min == 0) {
    std::memcpy(out, vis, words * 8);
  } else {
    uint64_t *merged = alloc0();
    for (uint32_t i = k_min; i <= k_actual; ++i) {
      const uint64_t *li = layer[i];
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < static_cast<int64_t>((words + 65535) / 65536); ++c) {
        const size_t s0 = static_cast<size_t>(c) * 65536;
        or_range(merged + s0, li + s0, std::min<size_t>(65536, words - s0));
      }
    }
    std::memcpy(out, merged, words * 8);
    std::free(merged);
  }
  const auto t1 = std::chrono::steady_clock::now();
  for (auto p : layer) std::free(p);
  std::free(vis);
  if (k_actual_out) *k_actual_out = k_actual;
  if (visit_count_out) *visit_count_out = visit_count;
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

// Row-major (S x ceil(V/64)) -> DVSM words.
void orc_rowmajor_to_dvsm(const uint64_t *rm, uint32_t S, uint32_t V, uint64_t *out) {
  const uint64_t nw = nwords(V);
  const uint64_t words = orc_dvsm_words(S, V);
  std::memset(out, 0, words * 8);
  const uint64_t v_per_stack = kWpc * V;
  for (uint64_t r = 0; r < S; ++r)
    for (uint64_t k = 0; k < nw; ++k)
      for (uint64_t m = rm[r * nw + k]; m; m &= m - 1) {
        const uint64_t c = k * 64 + __builtin_ctzll(m);
        out[(r / kStack) * v_per_stack + c * kWpc + (r % kStack) / 64] |= 1ull << (r % 64);
      }
}

// Per-row popcounts of a DenseVarStackMatrix of S rows and V columns: row r
// lives in stack r / 1024, word (r % 1024) / 64 of each column's 16, bit
// r % 64 (bitmap.hh:348-356).  One task per (stack, word) position.
void orc_dvsm_row_counts(const uint64_t *dv, uint32_t S, uint32_t V, uint64_t *counts) {
  const uint64_t stacks = (static_cast<uint64_t>(S) + kStack - 1) / kStack;
  const uint64_t v_per_stack = kWpc * static_cast<uint64_t>(V);
  std::fill(counts, counts + S, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t task = 0; task < static_cast<int64_t>(stacks * kWpc); ++task) {
    const uint64_t st = task / kWpc, w = task % kWpc;
    uint64_t cnt[64] = {0};
    const uint64_t *base = dv + st * v_per_stack + w;
    for (uint64_t c = 0; c < V; ++c)
      for (uint64_t x = base[c * kWpc]; x; x &= x - 1) cnt[__builtin_ctzll(x)]++;
    for (uint64_t b = 0; b < 64; ++b) {
      const uint64_t r = st * kStack + w * 64 + b;
      if (r < S) counts[r] = cnt[b];
    }
  }
}

// ----------------------------------------------------------------- oracle C
// deps: arity CSRs; tuples: T x arity row-major.  counts[T]; if lists != null,
// offsets[T+1] must be the exclusive scan of counts and lists gets the sets.
int orc_mintersect_sorted(uint32_t arity, const uint64_t *const *offs,
                          const uint32_t *const *adjs, const uint32_t *Vs,
                          const uint32_t *tuples, uint64_t T, uint32_t *counts,
                          const uint64_t *offsets, uint32_t *lists) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t t = 0; t < static_cast<int64_t>(T); ++t) {
    std::vector<uint32_t> cur, nxt;
    bool empty = false;
    for (uint32_t d = 0; d < arity && !empty; ++d) {
      const uint32_t x = tuples[t * arity + d];
      if (x >= Vs[d]) { empty = true; break; }
      const uint32_t *b = adjs[d] + offs[d][x], *e = adjs[d] + offs[d][x + 1];
      if (d == 0) cur.assign(b, e);
      else {
        nxt.clear();
        std::set_intersection(cur.begin(), cur.end(), b, e, std::back_inserter(nxt));
        cur.swap(nxt);
      }
    }
    if (empty) cur.clear();
    counts[t] = static_cast<uint32_t>(cur.size());
    if (lists) std::copy(cur.begin(), cur.end(), lists + offsets[t]);
  }
  return 0;
}

// The reference's formulation: per tuple an all-ones column of V bits ANDed
// with each dependency's column, then ascending enumeration of set bits.
int orc_mintersect_bitmap(uint32_t arity, const uint64_t *const *offs,
                          const uint32_t *const *adjs, const uint32_t *Vs,
                          uint32_t V, const uint32_t *tuples, uint64_t T,
                          uint32_t *counts, const uint64_t *offsets,
                          uint32_t *lists) {
  const uint64_t nw = nwords(V);
  std::vector<uint64_t> ir(nw), col(nw);
  for (uint64_t t = 0; t < T; ++t) {
    std::fill(ir.begin(), ir.end(), ~0ull);
    if (V % 64) ir[nw - 1] = (1ull << (V % 64)) - 1;
    for (uint32_t d = 0; d < arity; ++d) {
      const uint32_t x = tuples[t * arity + d];
      std::fill(col.begin(), col.end(), 0);
      if (x < Vs[d])
        for (uint64_t k = offs[d][x]; k < offs[d][x + 1]; ++k)
          if (adjs[d][k] < V) col[adjs[d][k] / 64] |= 1ull << (adjs[d][k] % 64);
      for (uint64_t w = 0; w < nw; ++w) ir[w] &= col[w];
    }
    uint32_t n = 0;
    for (uint64_t w = 0; w < nw; ++w) {
      for (uint64_t m = ir[w]; m; m &= m - 1) {
        if (lists) lists[offsets[t] + n] = static_cast<uint32_t>(w * 64 + __builtin_ctzll(m));
        ++n;
      }
    }
    counts[t] = n;
  }
  return 0;
}

// Scalar reference BFS distances from one source (for golden-vector checks).
int orc_bfs_dist(const uint64_t *off, const uint32_t *adj, uint32_t V,
                 uint32_t s, int32_t *dist) {
  std::fill(dist, dist + V, -1);
  std::vector<uint32_t> q{s};
  dist[s] = 0;
  for (size_t h = 0; h < q.size(); ++h) {
    const uint32_t u = q[h];
    for (uint64_t k = off[u]; k < off[u + 1]; ++k)
      if (dist[adj[k]] < 0) {
        dist[adj[k]] = dist[u] + 1;
        q.push_back(adj[k]);
      }
  }
  return 0;
}

}  // extern "C"
