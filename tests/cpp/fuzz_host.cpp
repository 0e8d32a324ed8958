// Seeded fuzz of the host-side generators and samplers (vs_csr_from_edges,
// vs_sample_sources, vs_sample_paths3) and of the store writer's flat edge
// files (vs_store_write with 1..70 partitions, fewer edges than partitions
// included, into a scratch directory), built for the host only with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_store_fuzz.py.
//
// Edge lists with random sizes, ids (some out of range), symmetrize and
// self-loop flags; CSRs with random and sometimes damaged offsets and ids fed
// to the path sampler; source samples of every size from 0 to V + 1.  Every
// call must return a status; a CSR that comes back must be ascending sets
// with offsets that match its edge count, and sampled paths must follow it.
//
// usage: fuzz_host <iterations> [scratch directory]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/vertexsurge.h"

// The device upload is not part of this build; nothing here calls it.
extern "C" vs_status vs_graph_create(const uint64_t *, const uint32_t *, uint32_t, uint64_t,
                                     uint32_t, vs_graph_t *) {
  return VS_ERR_UNSUPPORTED;
}

static int fail(const char *what, int i) {
  fprintf(stderr, "iteration %d: %s\n", i, what);
  return 3;
}

int main(int argc, char **argv) {
  if (argc != 2 && argc != 3) {
    fprintf(stderr, "usage: %s <iterations> [scratch directory]\n", argv[0]);
    return 2;
  }
  const int iters = atoi(argv[1]);
  std::mt19937_64 rng(11);
  int built = 0, rejected = 0;
  for (int i = 0; i < iters; ++i) {
    const uint32_t V = 1 + static_cast<uint32_t>(rng() % 300);
    const uint64_t n = rng() % 2000;
    std::vector<uint32_t> src(n), dst(n);
    const bool wild = rng() % 5 == 0;  // some ids out of range
    for (uint64_t e = 0; e < n; ++e) {
      src[e] = static_cast<uint32_t>(rng() % (wild ? V + 3 : V));
      dst[e] = static_cast<uint32_t>(rng() % (wild ? V + 3 : V));
    }
    vs_csr_t c = nullptr;
    const vs_status st = vs_csr_from_edges(V, n ? src.data() : nullptr, n ? dst.data() : nullptr,
                                           n, static_cast<int>(rng() % 2),
                                           static_cast<int>(rng() % 2), &c);
    if (st != VS_OK) {
      ++rejected;
    } else {
      ++built;
      uint32_t cv = 0;
      uint64_t ce = 0;
      const uint64_t *off = nullptr;
      const uint32_t *adj = nullptr;
      if (vs_csr_info(c, &cv, &ce, &off, &adj) != VS_OK || cv != V || off[0] != 0 || off[V] != ce)
        return fail("csr info", i);
      for (uint32_t v = 0; v < V; ++v)
        for (uint64_t k = off[v]; k < off[v + 1]; ++k)
          if (adj[k] >= V || (k > off[v] && adj[k] <= adj[k - 1])) return fail("csr lists", i);
      // paths over this CSR, or over a damaged copy of it
      std::vector<uint64_t> o(off, off + V + 1);
      std::vector<uint32_t> a(adj, adj + ce);
      const int damage = static_cast<int>(rng() % 4);
      if (damage == 1 && V > 1) std::swap(o[1 + rng() % (V - 1)], o[V]);  // out of order
      if (damage == 2 && ce) a[rng() % ce] = V + static_cast<uint32_t>(rng() % 5);
      const uint64_t T = rng() % 500;
      std::vector<uint32_t> out(3 * T + 1);
      const vs_status ps =
          vs_sample_paths3(o.data(), a.empty() ? nullptr : a.data(), V, T, rng(), out.data());
      if (ps == VS_OK && damage == 0)
        for (uint64_t t = 0; t < T; ++t) {
          const uint32_t x = out[3 * t], y = out[3 * t + 1], z = out[3 * t + 2];
          if (!std::binary_search(adj + off[x], adj + off[x + 1], y)) return fail("path a-b", i);
          if (off[y + 1] > off[y] && !std::binary_search(adj + off[y], adj + off[y + 1], z))
            return fail("path b-c", i);
        }
      if (argc == 3 && built % 25 == 1) {
        const uint32_t parts = 1 + static_cast<uint32_t>(rng() % 70);
        if (vs_store_write(argv[2], "fuzz", "V", "e", static_cast<int32_t>(rng() % 2), off, adj, V,
                           ce, parts) != VS_OK)
          return fail("store write", i);
      }
      vs_csr_free(c);
    }
    const uint32_t S = static_cast<uint32_t>(rng() % (V + 2));
    std::vector<uint32_t> s(S + 1);
    if (vs_sample_sources(V, S, rng(), s.data()) == VS_OK) {
      for (uint32_t k = 0; k < S; ++k)
        if (s[k] >= V || (k && s[k] <= s[k - 1])) return fail("sources", i);
    } else if (S <= V) {
      return fail("sources rejected", i);
    }
  }
  printf("built %d rejected %d\n", built, rejected);
  return 0;
}
