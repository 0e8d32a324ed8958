// Drives the C++ operator shims (vertexsurge_amd/host/vs_operators.hh) the way
// a reference plan would, and prints their outputs as JSON for
// tests/test_cpp_operators.py to compare with the CPU oracle.
#include <cstdio>
#include <string>
#include <vector>

#include "../../vertexsurge_amd/host/vs_operators.hh"

using namespace vertexsurge;

static void print_vec(const char *name, const std::vector<uint64_t> &v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%llu", i ? "," : "", (unsigned long long)v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char **argv) {
  vs_csr_t csr = nullptr;
  check(vs_rmat_csr(12, 16, 0.57, 0.19, 0.19, 1, &csr), "vs_rmat_csr");
  uint32_t V = 0;
  uint64_t E = 0;
  const uint64_t *off = nullptr;
  const uint32_t *adj = nullptr;
  check(vs_csr_info(csr, &V, &E, &off, &adj), "vs_csr_info");
  auto rel = std::make_shared<EdgeRelation>(off, adj, V, E, true);
  GraphStore::instance().add_edge("Person", "knows", "Person", rel);
  std::vector<VID> starts(200);
  check(vs_sample_sources(V, 200, 7, starts.data()), "vs_sample_sources");

  std::printf("{\n");
  // VExpand(Person, Person, knows, Both, 1, 3, Shortest, Dense, "p1.id", {})
  VExpand vx("Person", "Person", "knows", Direction::Both, 1, 3, ExpandPathType::Shortest,
             ExpandStrategy::Dense, "p1.id", std::nullopt);
  vx.run(starts);
  auto tab = vx.results();
  std::printf("\"shortest_k_actual\": %zu,\n", vx.k_actual_max());
  std::printf("\"shortest_tuple_count\": %zu,\n", tab.tuple_count());
  print_vec("shortest_row_counts", tab.row_counts());
  std::printf("\"shortest_get_0_first\": %d,\n",
              (int)tab.get(0, tab.tuples().empty() ? 0 : tab.tuples()[0].second));

  // the CSR strategy: the same pairs as a flat table (merge_to_Flat)
  VExpand vc("Person", "Person", "knows", Direction::Both, 1, 3, ExpandPathType::Shortest,
             ExpandStrategy::CSR, "p1.id", std::nullopt);
  vc.run(starts);
  const FlatTable ft = vc.flat_results();
  std::printf("\"csr_flat_tuples\": %zu,\n", ft.tuple_count());
  std::printf("\"csr_flat_first\": [%u, %u],\n", ft.start.empty() ? 0u : ft.start[0],
              ft.target.empty() ? 0u : ft.target[0]);

  VExpand va("Person", "Person", "knows", Direction::Forward, 0, 2,
             VExpandPathParameters::Any(), ExpandStrategy::Dense, "p1.id", std::nullopt);
  va.run(starts);
  print_vec("any_row_counts", va.results().row_counts());

  VExpand vd("Person", "Person", "knows", Direction::Forward, 1, 3,
             VExpandPathParameters::Shortest().set_with_distance(), ExpandStrategy::Dense,
             "p1.id", std::nullopt);
  vd.run(starts);
  auto multi = vd.multiple_results();
  std::vector<uint64_t> dists, per_layer;
  for (auto &m : multi.mats) {
    dists.push_back(m.first);
    per_layer.push_back(m.second.tuple_count());
  }
  print_vec("layer_distances", dists);
  print_vec("layer_tuple_counts", per_layer);

  std::vector<VID> targets(starts.rbegin(), starts.rend());
  VExpand vf("Person", "Person", "knows", Direction::Forward,
             VExpandPathParameters::FindShortest(), ExpandStrategy::Dense, "p1.id", "p2.id");
  vf.run(starts, &targets);
  std::vector<uint64_t> fd;
  for (int32_t d : vf.distance()) fd.push_back(static_cast<uint64_t>(static_cast<int64_t>(d) + 1));
  print_vec("find_shortest_dist_plus1", fd);

  // empty start column: empty result, no error
  VExpand ve("Person", "Person", "knows", Direction::Forward, 1, 2, ExpandPathType::Shortest,
             ExpandStrategy::Dense, "p1.id", std::nullopt);
  ve.run({});
  std::printf("\"empty_tuple_count\": %zu,\n", ve.results().tuple_count());

  // unknown edge schema -> std::runtime_error
  bool threw = false;
  try {
    VExpand bad("Person", "Person", "likes", Direction::Forward, 1, 2,
                ExpandPathType::Shortest, ExpandStrategy::Dense, "p1.id", std::nullopt);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  std::printf("\"unknown_schema_throws\": %d,\n", threw ? 1 : 0);

  // schema and strategy checks of the reference constructor /
  // finishAddingDependency (vexpand.cpp:88-91, 125-163): a relation between
  // two vertex types allows k_max = 1 only; Adaptive is not an output
  // strategy; each strategy has its own result table kind
  {
    GraphStore::instance().add_edge("Person", "likes", "Post", rel);
    auto throws = [](auto fn) {
      try {
        fn();
      } catch (const std::runtime_error &) {
        return 1;
      }
      return 0;
    };
    const int two_types_k2 = throws([] {
      VExpand v("Person", "Post", "likes", Direction::Forward, 1, 2, ExpandPathType::Any,
                ExpandStrategy::Dense, "p.id", std::nullopt);
    });
    const int two_types_k1 = throws([] {
      VExpand v("Person", "Post", "likes", Direction::Forward, 1, 1, ExpandPathType::Any,
                ExpandStrategy::Dense, "p.id", std::nullopt);
    });
    const int adaptive = throws([] {
      VExpand v("Person", "Person", "knows", Direction::Forward, 1, 2,
                ExpandPathType::Shortest, ExpandStrategy::Adaptive, "p.id", std::nullopt);
    });
    VExpand vcs("Person", "Person", "knows", Direction::Forward, 1, 2, ExpandPathType::Shortest,
                ExpandStrategy::CSR, "p.id", std::nullopt);
    vcs.run(std::vector<VID>(starts.begin(), starts.begin() + 3));
    const int csr_matrix = throws([&] { (void)vcs.results(); });
    VExpand vdn("Person", "Person", "knows", Direction::Forward, 1, 2, ExpandPathType::Shortest,
                ExpandStrategy::Dense, "p.id", std::nullopt);
    vdn.run(std::vector<VID>(starts.begin(), starts.begin() + 3));
    const int dense_flat = throws([&] { (void)vdn.flat_results(); });
    std::printf("\"schema_checks_throw\": [%d, %d, %d, %d, %d],\n", two_types_k2, two_types_k1,
                adaptive, csr_matrix, dense_flat);
  }

  // Mintersect KAT (mintersect.cpp:347-409): 3 result tuples
  {
    // vertices 1..8 have one neighbour each (1,1,1,2,2,3,3,4); 0 and 9 none
    std::vector<uint64_t> koff(11, 0);
    std::vector<uint32_t> kadj = {1, 1, 1, 2, 2, 3, 3, 4};
    for (int k = 2; k <= 10; ++k) koff[k] = k - 1 > 8 ? 8 : k - 1;
    auto dep = std::make_shared<EdgeRelation>(koff.data(), kadj.data(), 10, 8, false);
    Mintersect mi("c", "Person", false);
    mi.add_dependency(dep);
    mi.add_dependency(dep);
    std::vector<VID> base = {1, 1, 2, 1, 3, 1, 4, 2, 5, 2, 6, 3, 7, 3, 8, 4};
    mi.run(base);
    std::vector<uint64_t> flat(mi.flat_tuples().begin(), mi.flat_tuples().end());
    std::printf("\"mintersect_kat_tuples\": %zu,\n", mi.tuple_count());
    print_vec("mintersect_kat_flat", flat);
    Mintersect mc("c", "Person", true);
    mc.add_dependency(dep);
    mc.add_dependency(dep);
    mc.run(base);
    std::printf("\"mintersect_kat_count\": %llu,\n", (unsigned long long)mc.count());
  }

  // MintersectFlatMatMat (mintersect.cpp:411-512): ColMain dependencies, 10x10,
  // columns mapped through the 8 keys 1..8
  {
    const uint32_t keys[8] = {1, 2, 3, 4, 5, 6, 7, 8};
    const uint32_t cs[8] = {1, 1, 1, 2, 2, 3, 3, 4};
    std::vector<uint64_t> words(10 * 8, 0);  // 10 columns x align8(1) words
    for (int i = 0; i < 8; ++i) words[static_cast<size_t>(i) * 8] |= 1ull << cs[i];
    auto a = std::make_shared<ColMainMatrixTable>(
        ColMainMatrixTable::from_words(rel->handle(), 10, 10, words));
    auto b = std::make_shared<ColMainMatrixTable>(
        ColMainMatrixTable::from_words(rel->handle(), 10, 10, words));
    const std::vector<VID> kv(keys, keys + 8);
    a->set_maps(std::nullopt, kv);
    b->set_maps(std::nullopt, kv);
    Mintersect mm("c", "Person", false);
    mm.add_dependency(a);
    mm.add_dependency(b);
    mm.run({1, 1, 2, 1, 3, 1, 4, 2, 5, 2, 6, 3, 7, 3, 8, 4});
    std::vector<uint64_t> flat(mm.flat_tuples().begin(), mm.flat_tuples().end());
    std::printf("\"mintersect_matmat_tuples\": %zu,\n", mm.tuple_count());
    print_vec("mintersect_matmat_flat", flat);
  }

  // triangles through each edge: VExpand(1..1) as the VarRow base, both edge
  // endpoints bound to the same relation (mintersect.cpp:244-333)
  {
    VExpand v1("Person", "Person", "knows", Direction::Forward, 1, 1,
               VExpandPathParameters::Any(), ExpandStrategy::Dense, "a.id", std::nullopt);
    const std::vector<VID> few(starts.begin(), starts.begin() + 40);
    v1.run(few);
    auto edges = v1.results();
    Mintersect tv("c", "Person", false);
    tv.add_dependency(rel);
    tv.add_dependency(rel);
    tv.run(edges, {0, 1});
    std::vector<uint64_t> flat(tv.flat_tuples().begin(), tv.flat_tuples().end());
    print_vec("varrow_triangles_flat", flat);
    Mintersect tc("c", "Person", true);
    tc.add_dependency(rel);
    tc.add_dependency(rel);
    tc.run(edges, {0, 1});
    std::printf("\"varrow_triangles_count\": %llu,\n", (unsigned long long)tc.count());
  }

  // a -[1..2]-> c <-[1..2]- b: VExpand, TableConvert (transposed), Mintersect
  {
    VExpand vp("Person", "Person", "knows", Direction::Forward, 1, 2,
               VExpandPathParameters::Any(), ExpandStrategy::Dense, "a.id", std::nullopt);
    vp.run(starts);
    auto reach = vp.results();
    auto cm = std::make_shared<ColMainMatrixTable>(
        TableConvert(TableType::ColMainMatrix, true).run(reach));
    std::printf("\"convert_rows\": %u, \"convert_cols\": %u,\n", cm->rows(), cm->cols());
    std::vector<VID> base;
    for (size_t i = 0; i + 1 < starts.size(); i += 2) {
      base.push_back(starts[i]);
      base.push_back(starts[i + 1]);
    }
    Mintersect pm("c", "Person", true);
    pm.add_dependency(cm);
    pm.add_dependency(cm);
    pm.run(base);
    std::printf("\"pattern_count\": %llu,\n", (unsigned long long)pm.count());
    // MatrixColumnCount (DESC) and MatrixRowOrSum on the reach matrix
    auto cc = MatrixColumnCount("cnt", SortOrder::DESC).run(reach);
    std::vector<uint64_t> top;
    for (size_t i = 0; i < 5 && i < cc.size(); ++i) {
      top.push_back(cc[i].first);
      top.push_back(static_cast<uint64_t>(cc[i].second));
    }
    print_vec("colcount_desc_top5", top);
    std::printf("\"row_or_size\": %zu,\n", MatrixRowOrSum().run(reach).size());
    // MatrixTableSelect of three columns, then its tuples
    auto sel = MatrixTableSelect("c").run(reach, {starts[0], starts[1], starts[0]});
    std::printf("\"select_tuple_count\": %zu,\n", sel.tuple_count());
  }
  // a graph directory in the reference's storage format (argv[1]), loaded in
  // place of the R-MAT relation
  {
    const std::string dir = argc > 1 ? argv[1] : "tests/golden/store/knows";
    GraphStore::instance().load(dir, "Person", "knows", "Person");
    const std::vector<VID> all = {0, 1, 2, 3, 4};
    VExpand vs1("Person", "Person", "knows", Direction::Both, 1, 3, ExpandPathType::Shortest,
                ExpandStrategy::Dense, "p.id", std::nullopt);
    vs1.run(all);
    print_vec("store_shortest_row_counts", vs1.results().row_counts());
    VExpand va2("Person", "Person", "knows", Direction::Both, 1, 2,
                VExpandPathParameters::Any(), ExpandStrategy::Dense, "p.id", std::nullopt);
    va2.run(all);
    print_vec("store_any_row_counts", va2.results().row_counts(), true);
  }
  std::printf("}\n");
  GraphStore::instance().clear();
  vs_csr_free(csr);
  return 0;
}
