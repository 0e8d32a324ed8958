// Seeded fuzz of the graph-store reader (vs_store_open and the CSR it
// exposes), built for the host only with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_store_fuzz.py.
//
// Each iteration writes a damaged copy of a store directory -- schema.json
// with random byte changes, inserted JSON punctuation, deletions or a
// truncation; DiskArray headers with extreme counts or element sizes;
// truncated data files -- and opens it.  An open that succeeds must expose
// offsets[0..V] and adj[0..E) that can be read in full.  The sanitizers turn
// any out-of-bounds access, use after free, leak or undefined operation into
// a non-zero exit.
//
// usage: fuzz_store <store dir> <scratch dir> <iterations>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "../../include/vertexsurge.h"

// The device upload is not part of this build; the reader never calls it.
extern "C" vs_status vs_graph_create(const uint64_t *, const uint32_t *, uint32_t, uint64_t,
                                     uint32_t, vs_graph_t *) {
  return VS_ERR_UNSUPPORTED;
}

static bool slurp(const std::string &p, std::string &s) {
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return false;
  char b[4096];
  size_t n;
  while ((n = fread(b, 1, sizeof b, f)) > 0) s.append(b, n);
  fclose(f);
  return true;
}

static bool spit(const std::string &p, const std::string &s) {
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(s.data(), 1, s.size(), f) == s.size();
  return fclose(f) == 0 && ok;
}

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <store dir> <scratch dir> <iterations>\n", argv[0]);
    return 2;
  }
  const std::string src = argv[1], dir = argv[2];
  const int iters = atoi(argv[3]);
  mkdir(dir.c_str(), 0755);
  const char *names[] = {"Person-knows-Person_csr_src_dst",
                         "Person-knows-Person_csr_src_dst_index"};
  std::string files[2], schema;
  if (!slurp(src + "/" + names[0], files[0]) || !slurp(src + "/" + names[1], files[1]) ||
      !slurp(src + "/schema.json", schema)) {
    fprintf(stderr, "cannot read the store under %s\n", src.c_str());
    return 2;
  }
  std::mt19937_64 rng(7);
  const char punct[] = "{}[]\",:\\-0123456789eE.+tfnu\"x ";
  int ok = 0, rejected = 0;
  for (int i = 0; i < iters; ++i) {
    std::string s = schema;
    const int kind = static_cast<int>(rng() % 4);
    const int edits = 1 + static_cast<int>(rng() % 6);
    for (int e = 0; e < edits && !s.empty(); ++e) {
      const size_t p = rng() % s.size();
      if (kind == 0) s[p] = static_cast<char>(rng() & 0xFF);
      else if (kind == 1) s.insert(p, 1, punct[rng() % (sizeof punct - 1)]);
      else if (kind == 2) s.erase(p, 1 + rng() % 8);
      else s.resize(p);
    }
    if (!spit(dir + "/schema.json", s)) return 2;
    for (int f = 0; f < 2; ++f) {
      std::string d = files[f];
      if (rng() % 4 == 0) {
        const int64_t vals[] = {-1, 0, 1, 3, 4, 8, 7, int64_t{1} << 40, INT64_MAX};
        const int64_t v = vals[rng() % 9];
        memcpy(&d[(rng() % 2) * 8], &v, 8);
      }
      if (rng() % 8 == 0) d.resize(rng() % d.size());
      if (!spit(dir + "/" + names[f], d)) return 2;
    }
    vs_store_t st = nullptr;
    if (vs_store_open(dir.c_str(), nullptr, &st) == VS_OK) {
      const uint64_t *o = nullptr;
      const uint32_t *a = nullptr;
      vs_store_info info;
      if (vs_store_csr(st, &o, &a) != VS_OK || vs_store_get_info(st, &info) != VS_OK) return 3;
      uint64_t sum = 0;
      for (uint64_t v = 0; v <= info.num_vertices; ++v) sum += o[v];
      for (uint64_t e = 0; e < info.num_edges; ++e) sum += a[e];
      if (o[0] != 0 || o[info.num_vertices] != info.num_edges) return 4;
      vs_store_close(st);
      ok += sum != ~0ull;  // keeps the reads
    } else {
      ++rejected;
    }
  }
  printf("ok %d rejected %d\n", ok, rejected);
  return 0;
}
