// Host half of libvertexsurge.so: error text, synthetic graph generators,
// CSR construction and seeded samplers.  Built with g++ -fopenmp.
//
// The CSR contract follows the reference's sorted-by-(src, dst) adjacency
// with an offsets index (src/include/storage/graph_concepts.hh:114-137,
// src/include/meta/graph_loader.hh:163-182): offsets[V+1] (uint64),
// adj[E] (uint32), each list ascending.  Unlike the reference importer, which
// removes only adjacent duplicates before sorting (graph_loader.hh:71-77),
// duplicates and (for the synthetic graphs) self-loops are removed fully.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <new>

#include "vs_internal.hh"

namespace vs {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

void clear_error() { g_err.clear(); }

// Builds a CSR from a pair generator: pass 1 counts, pass 2 scatters, pass 3
// sorts + dedups each list, pass 4 compacts.  gen(e, &u, &v) returns false to
// drop a pair.  Result is independent of thread count.
template <class Gen>
static void build_csr(uint32_t V, uint64_t npairs, bool symmetrize,
                      bool drop_self, Gen gen, HostCSR &out) {
  std::vector<uint64_t> deg(static_cast<size_t>(V) + 1, 0);
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < npairs; ++e) {
    uint32_t u, v;
    if (!gen(e, &u, &v)) continue;
    if (drop_self && u == v) continue;
    __atomic_fetch_add(&deg[u], 1, __ATOMIC_RELAXED);
    if (symmetrize && u != v) __atomic_fetch_add(&deg[v], 1, __ATOMIC_RELAXED);
  }
  std::vector<uint64_t> off(static_cast<size_t>(V) + 1, 0);
  for (uint32_t i = 0; i < V; ++i) off[i + 1] = off[i] + deg[i];
  const uint64_t total = off[V];
  std::vector<uint32_t> tmp(total);
  std::vector<uint64_t> cur(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < npairs; ++e) {
    uint32_t u, v;
    if (!gen(e, &u, &v)) continue;
    if (drop_self && u == v) continue;
    tmp[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
    if (symmetrize && u != v)
      tmp[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
  }
  std::vector<uint64_t>().swap(cur);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < V; ++i) {
    uint32_t *b = tmp.data() + off[i];
    uint32_t *e = tmp.data() + off[i + 1];
    std::sort(b, e);
    deg[i] = static_cast<uint64_t>(std::unique(b, e) - b);
  }
  out.V = V;
  out.off.assign(static_cast<size_t>(V) + 1, 0);
  for (uint32_t i = 0; i < V; ++i) out.off[i + 1] = out.off[i] + deg[i];
  out.E = out.off[V];
  out.adj.resize(out.E);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < V; ++i) {
    if (deg[i])  // an edgeless CSR has no adj storage (data() may be null)
      std::memcpy(out.adj.data() + out.off[i], tmp.data() + off[i],
                  deg[i] * sizeof(uint32_t));
  }
}

void reverse_csr(const uint64_t *off, const uint32_t *adj, uint32_t V,
                 uint64_t E, std::vector<uint64_t> &roff,
                 std::vector<uint32_t> &radj) {
  roff.assign(static_cast<size_t>(V) + 1, 0);
  for (uint64_t k = 0; k < E; ++k) roff[adj[k] + 1]++;
  for (uint32_t i = 0; i < V; ++i) roff[i + 1] += roff[i];
  radj.resize(E);
  std::vector<uint64_t> cur(roff.begin(), roff.end() - 1);
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t k = off[u]; k < off[u + 1]; ++k) radj[cur[adj[k]]++] = u;
}

void union_csr(const uint64_t *off, const uint32_t *adj, const uint64_t *roff,
               const uint32_t *radj, uint32_t V, std::vector<uint64_t> &uoff,
               std::vector<uint32_t> &uadj) {
  std::vector<uint64_t> deg(V);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V; ++v) {
    uint64_t a = off[v], ae = off[v + 1], b = roff[v], be = roff[v + 1], n = 0;
    while (a < ae || b < be) {
      uint32_t x;
      if (b >= be || (a < ae && adj[a] < radj[b])) x = adj[a++];
      else if (a >= ae || radj[b] < adj[a]) x = radj[b++];
      else { x = adj[a++]; ++b; }
      (void)x;
      ++n;
    }
    deg[v] = n;
  }
  uoff.assign(static_cast<size_t>(V) + 1, 0);
  for (uint32_t i = 0; i < V; ++i) uoff[i + 1] = uoff[i] + deg[i];
  uadj.resize(uoff[V]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V; ++v) {
    uint64_t a = off[v], ae = off[v + 1], b = roff[v], be = roff[v + 1];
    uint64_t o = uoff[v];
    while (a < ae || b < be) {
      uint32_t x;
      if (b >= be || (a < ae && adj[a] < radj[b])) x = adj[a++];
      else if (a >= ae || radj[b] < adj[a]) x = radj[b++];
      else { x = adj[a++]; ++b; }
      uadj[o++] = x;
    }
  }
}

} // namespace vs

using namespace vs;

struct vs_csr_s {
  HostCSR csr;
};

extern "C" {

const char *vs_last_error(void) { return vs::g_err.c_str(); }
const char *vs_version(void) { return "vertexsurge-amd 0.1 (gfx950)"; }

// R-MAT quadrant draw.  Level l (0 = most significant bit) uses 32 bits of
// draw mix64(seed, e * ceil(scale/2) + l/2): the low half for even l, the
// high half for odd l.  r < A -> (0,0); r < A+B -> (0,1); r < A+B+C -> (1,0);
// else (1,1), with A, A+B, A+B+C the probabilities scaled to 2^32 and rounded.
vs_status vs_rmat_csr(uint32_t scale, uint32_t edge_factor, double a, double b,
                      double c, uint64_t seed, vs_csr_t *out) try {
  clear_error();
  if (!out || scale == 0 || scale > 31 || edge_factor == 0 || a < 0 || b < 0 ||
      c < 0 || a + b + c > 1.0) {
    set_error("vs_rmat_csr: invalid arguments");
    return VS_ERR_INVALID;
  }
  const double two32 = 4294967296.0;
  const uint64_t TA = static_cast<uint64_t>(std::llround(a * two32));
  const uint64_t TAB = static_cast<uint64_t>(std::llround((a + b) * two32));
  const uint64_t TABC = static_cast<uint64_t>(std::llround((a + b + c) * two32));
  const uint32_t V = 1u << scale;
  const uint64_t npairs = static_cast<uint64_t>(edge_factor) << scale;
  const uint64_t nd = (scale + 1) / 2;
  auto gen = [=](uint64_t e, uint32_t *pu, uint32_t *pv) {
    uint32_t u = 0, v = 0;
    uint64_t d = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) d = mix64(seed, e * nd + l / 2);
      const uint64_t r = (l & 1) ? (d >> 32) : (d & 0xffffffffull);
      const uint32_t ub = r >= TAB;
      const uint32_t vb = (r >= TA && r < TAB) || r >= TABC;
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    *pu = u;
    *pv = v;
    return true;
  };
  vs_csr_s *h = new (std::nothrow) vs_csr_s;
  if (!h) { set_error("vs_rmat_csr: out of host memory"); return VS_ERR_NOMEM; }
  try {
    build_csr(V, npairs, true, true, gen, h->csr);
  } catch (const std::bad_alloc &) {
    delete h;
    set_error("vs_rmat_csr: out of host memory");
    return VS_ERR_NOMEM;
  }
  *out = h;
  return VS_OK;
} VS_CATCH("vs_rmat_csr")

vs_status vs_er_csr(uint32_t num_vertices, uint64_t num_pairs, uint64_t seed,
                    vs_csr_t *out) try {
  clear_error();
  if (!out || num_vertices == 0) {
    set_error("vs_er_csr: invalid arguments");
    return VS_ERR_INVALID;
  }
  const uint32_t V = num_vertices;
  auto gen = [=](uint64_t e, uint32_t *pu, uint32_t *pv) {
    const uint64_t d = mix64(seed, e);
    *pu = static_cast<uint32_t>(((d & 0xffffffffull) * V) >> 32);
    *pv = static_cast<uint32_t>(((d >> 32) * V) >> 32);
    return true;
  };
  vs_csr_s *h = new (std::nothrow) vs_csr_s;
  if (!h) { set_error("vs_er_csr: out of host memory"); return VS_ERR_NOMEM; }
  try {
    build_csr(V, num_pairs, true, true, gen, h->csr);
  } catch (const std::bad_alloc &) {
    delete h;
    set_error("vs_er_csr: out of host memory");
    return VS_ERR_NOMEM;
  }
  *out = h;
  return VS_OK;
} VS_CATCH("vs_er_csr")

vs_status vs_csr_from_edges(uint32_t num_vertices, const uint32_t *src,
                            const uint32_t *dst, uint64_t n, int symmetrize,
                            int drop_self_loops, vs_csr_t *out) try {
  clear_error();
  if (!out || (n > 0 && (!src || !dst))) {
    set_error("vs_csr_from_edges: invalid arguments");
    return VS_ERR_INVALID;
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (src[i] >= num_vertices || dst[i] >= num_vertices) {
      set_error("vs_csr_from_edges: edge %llu (%u,%u) out of range V=%u",
                (unsigned long long)i, src[i], dst[i], num_vertices);
      return VS_ERR_INVALID;
    }
  }
  auto gen = [=](uint64_t e, uint32_t *pu, uint32_t *pv) {
    *pu = src[e];
    *pv = dst[e];
    return true;
  };
  // owned until handed out: an exception thrown by build_csr (host OOM)
  // frees the handle on its way to VS_CATCH
  std::unique_ptr<vs_csr_s> h(new vs_csr_s);
  build_csr(num_vertices, n, symmetrize != 0, drop_self_loops != 0, gen, h->csr);
  *out = h.release();
  return VS_OK;
} VS_CATCH("vs_csr_from_edges")

vs_status vs_csr_info(vs_csr_t csr, uint32_t *num_vertices, uint64_t *num_edges,
                      const uint64_t **offsets, const uint32_t **adj) try {
  clear_error();
  if (!csr) { set_error("vs_csr_info: null handle"); return VS_ERR_INVALID; }
  if (num_vertices) *num_vertices = csr->csr.V;
  if (num_edges) *num_edges = csr->csr.E;
  if (offsets) *offsets = csr->csr.off.data();
  if (adj) *adj = csr->csr.adj.data();
  return VS_OK;
} VS_CATCH("vs_csr_info")

vs_status vs_csr_free(vs_csr_t csr) try {
  delete csr;
  return VS_OK;
} VS_CATCH("vs_csr_free")

// Sparse case (2S <= V): draw i gives draw_index(mix64(seed, i), V); keep the
// first S distinct values.  Dense case: partial Fisher-Yates where step i swaps
// position i with i + draw_index(mix64(seed, i), V - i).  Output sorted.
vs_status vs_sample_sources(uint32_t num_vertices, uint32_t S, uint64_t seed,
                            uint32_t *out) try {
  clear_error();
  if (S > num_vertices || (S > 0 && !out)) {
    set_error("vs_sample_sources: S=%u exceeds V=%u", S, num_vertices);
    return VS_ERR_INVALID;
  }
  if (S == 0) return VS_OK;
  const uint32_t V = num_vertices;
  if (2ull * S <= V) {
    std::vector<uint64_t> seen((V + 63) / 64, 0);
    uint32_t n = 0;
    for (uint64_t i = 0; n < S; ++i) {
      const uint32_t x = draw_index(mix64(seed, i), V);
      uint64_t &w = seen[x >> 6];
      const uint64_t bit = 1ull << (x & 63);
      if (w & bit) continue;
      w |= bit;
      out[n++] = x;
    }
  } else {
    std::vector<uint32_t> perm(V);
    for (uint32_t i = 0; i < V; ++i) perm[i] = i;
    for (uint32_t i = 0; i < S; ++i) {
      const uint32_t j = i + draw_index(mix64(seed, i), V - i);
      std::swap(perm[i], perm[j]);
    }
    std::copy(perm.begin(), perm.begin() + S, out);
  }
  std::sort(out, out + S);
  return VS_OK;
} VS_CATCH("vs_sample_sources")

// Triple t uses draws mix64(seed, 3t), mix64(seed, 3t+1), mix64(seed, 3t+2).
vs_status vs_sample_paths3(const uint64_t *offsets, const uint32_t *adj,
                           uint32_t num_vertices, uint64_t T, uint64_t seed,
                           uint32_t *out) try {
  clear_error();
  if (!offsets || (T > 0 && (!adj || !out))) {
    set_error("vs_sample_paths3: invalid arguments");
    return VS_ERR_INVALID;
  }
  // the lists must not overlap or run backwards: every adj index below is
  // then inside [offsets[0], offsets[V])
  for (uint32_t v = 0; v < num_vertices; ++v)
    if (offsets[v + 1] < offsets[v]) {
      set_error("vs_sample_paths3: offsets decrease at vertex %u", v);
      return VS_ERR_INVALID;
    }
  std::vector<uint32_t> nonisolated;
  for (uint32_t v = 0; v < num_vertices; ++v)
    if (offsets[v + 1] > offsets[v]) nonisolated.push_back(v);
  if (nonisolated.empty() && T > 0) {
    set_error("vs_sample_paths3: graph has no edges");
    return VS_ERR_INVALID;
  }
  const uint32_t n = static_cast<uint32_t>(nonisolated.size());
  int bad = 0;  // a neighbour id >= V: the path stops there
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (uint64_t t = 0; t < T; ++t) {
    const uint32_t a = nonisolated[draw_index(mix64(seed, 3 * t), n)];
    const uint32_t da = static_cast<uint32_t>(offsets[a + 1] - offsets[a]);
    const uint32_t b = adj[offsets[a] + draw_index(mix64(seed, 3 * t + 1), da)];
    uint32_t c = b;
    if (b >= num_vertices) {
      bad = 1;
    } else {
      const uint32_t db = static_cast<uint32_t>(offsets[b + 1] - offsets[b]);
      if (db) c = adj[offsets[b] + draw_index(mix64(seed, 3 * t + 2), db)];
      if (c >= num_vertices) bad = 1;
    }
    out[3 * t] = a;
    out[3 * t + 1] = b;
    out[3 * t + 2] = c;
  }
  if (bad) {
    set_error("vs_sample_paths3: a neighbour id is >= V=%u", num_vertices);
    return VS_ERR_INVALID;
  }
  return VS_OK;
} VS_CATCH("vs_sample_paths3")

} // extern "C"
