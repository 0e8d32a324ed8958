// Host-side C++ operators over the C ABI: the VExpand and Mintersect
// interfaces of VertexSurge (src/include/operator/big_operators.hh:40-183)
// with their parameter lists, output table kinds and error behaviour, running
// on the MI355X through libvertexsurge.so.
//
// The reference resolves the edge relation from a global schema by
// "<src>_<edge>_<dst>" (vexpand.cpp:57-60); GraphStore keeps the same lookup:
// register each edge relation's CSR under that name once, and operators find
// it by their type names.  Start vertices come from the dependency's flat
// column in the reference (vexpand.cpp:168-176); here run() takes that column.
// Errors become std::runtime_error (the reference asserts or exits).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vertexsurge.h"

namespace vertexsurge {

using VID = uint32_t;
using TypeName = std::string;
using VariableName = std::string;

// src/include/operator/expand.hh:25-29
enum class ExpandPathType { Any, Shortest, FindShortest };
// src/include/common/common.hh:154-159
enum class Direction { Forward, Backward, Both, Undecide };
// src/include/operator/expand.hh:51-55
enum class ExpandStrategy { Adaptive, CSR, Dense };

// src/include/operator/expand.hh:31-49
struct VExpandPathParameters {
  ExpandPathType path_type = ExpandPathType::Shortest;
  static const size_t K_MAX_FIND_SHORTEST = 50;
  bool with_distance = false;
  bool distinct_target = false;
  static VExpandPathParameters Any() { return {ExpandPathType::Any}; }
  static VExpandPathParameters Shortest() { return {ExpandPathType::Shortest}; }
  static VExpandPathParameters FindShortest() {
    VExpandPathParameters p{ExpandPathType::FindShortest};
    p.with_distance = true;
    return p;
  }
  VExpandPathParameters set_with_distance() {
    with_distance = true;
    return *this;
  }
};

inline void check(vs_status s, const char *what) {
  if (s != VS_OK)
    throw std::runtime_error(std::string(what) + ": " + vs_last_error());
}

// One edge relation resident in HBM.
class EdgeRelation {
 public:
  EdgeRelation(const uint64_t *offsets, const uint32_t *adj, uint32_t num_vertices,
               uint64_t num_edges, bool symmetric) {
    check(vs_graph_create(offsets, adj, num_vertices, num_edges,
                          symmetric ? VS_GRAPH_SYMMETRIC : 0u, &g_),
          "vs_graph_create");
  }
  // adopts a graph handle (vs_store_graph)
  explicit EdgeRelation(vs_graph_t g) : g_(g) {}
  ~EdgeRelation() { vs_graph_destroy(g_); }
  EdgeRelation(const EdgeRelation &) = delete;
  EdgeRelation &operator=(const EdgeRelation &) = delete;
  vs_graph_t handle() const { return g_; }
  uint32_t vertex_count() const {
    uint32_t v = 0;
    check(vs_graph_info(g_, &v, nullptr, nullptr), "vs_graph_info");
    return v;
  }

 private:
  vs_graph_t g_ = nullptr;
};

// The global store of edge relations (replaces PhysicalOperator::graph_store).
class GraphStore {
 public:
  static GraphStore &instance() {
    static GraphStore s;
    return s;
  }
  void add_edge(const TypeName &src, const TypeName &edge, const TypeName &dst,
                std::shared_ptr<EdgeRelation> rel) {
    rels_[key(src, edge, dst)] = std::move(rel);
  }
  std::shared_ptr<EdgeRelation> get_edge(const TypeName &src, const TypeName &edge,
                                         const TypeName &dst) const {
    auto it = rels_.find(key(src, edge, dst));
    if (it == rels_.end())
      throw std::runtime_error("no edge schema " + key(src, edge, dst));  // schema.cpp:217
    return it->second;
  }
  // Graph::load (graph_store.cpp:47-73) of one relation src -edge-> dst from
  // a graph directory in the reference's storage format (schema.json +
  // DiskArray files); the relation is registered as add_edge does.
  std::shared_ptr<EdgeRelation> load(const std::string &dir, const TypeName &src,
                                     const TypeName &edge, const TypeName &dst) {
    vs_store_t st = nullptr;
    check(vs_store_open(dir.c_str(), key(src, edge, dst).c_str(), &st), "vs_store_open");
    vs_graph_t g = nullptr;
    const vs_status s = vs_store_graph(st, &g);
    vs_store_close(st);
    check(s, "vs_store_graph");
    auto rel = std::make_shared<EdgeRelation>(g);
    add_edge(src, edge, dst, rel);
    return rel;
  }
  void clear() { rels_.clear(); }

 private:
  static std::string key(const TypeName &s, const TypeName &e, const TypeName &d) {
    return s + "_" + e + "_" + d;
  }
  std::map<std::string, std::shared_ptr<EdgeRelation>> rels_;
};

// VarRowMatrixTable: rows = start vertices (row_map), columns = vertices;
// the bit matrix stays in HBM (matrix_table.hh:44-83).
class VarRowMatrixTable {
 public:
  VarRowMatrixTable(std::shared_ptr<vs_result_s> res, uint32_t index,
                    std::vector<VID> row_map,
                    std::optional<std::vector<VID>> col_map = std::nullopt)
      : res_(std::move(res)), index_(index), row_map_(std::move(row_map)),
        col_map_(std::move(col_map)) {}
  const std::vector<VID> &get_row_map() const { return row_map_; }
  // set by MatrixTableSelect; VExpand output has none (columns = vertices)
  const std::optional<std::vector<VID>> &get_col_map() const { return col_map_; }
  vs_result_s *handle() const { return res_.get(); }
  uint32_t index() const { return index_; }
  // bit_matrix_table.cpp:112
  size_t tuple_count() const {
    uint64_t c = 0;
    check(vs_result_count(res_.get(), index_, &c), "vs_result_count");
    return c;
  }
  // DenseVarStackMatrix::get (bitmap.cpp:953-956)
  bool get(uint32_t r, uint32_t c) const {
    int b = 0;
    check(vs_result_get(res_.get(), index_, r, c, &b), "vs_result_get");
    return b != 0;
  }
  std::vector<uint64_t> row_counts() const {
    std::vector<uint64_t> out(row_map_.size());
    if (!out.empty())
      check(vs_result_row_counts(res_.get(), index_, out.data()), "vs_result_row_counts");
    return out;
  }
  std::vector<uint64_t> export_words(vs_layout layout) const {
    size_t n = 0;
    check(vs_result_export_size(res_.get(), index_, layout, &n), "vs_result_export_size");
    std::vector<uint64_t> out(n / 8);
    if (n) check(vs_result_export(res_.get(), index_, layout, out.data(), n), "vs_result_export");
    return out;
  }
  // (start, target) pairs in the reference iterator's order for one stack:
  // row-major here (row ascending, then column ascending).
  std::vector<std::pair<VID, VID>> tuples() const {
    const auto rm = export_words(VS_LAYOUT_ROWMAJOR);
    std::vector<std::pair<VID, VID>> out;
    if (row_map_.empty()) return out;
    const size_t nw = rm.size() / row_map_.size();
    for (size_t r = 0; r < row_map_.size(); ++r)
      for (size_t w = 0; w < nw; ++w)
        for (uint64_t m = rm[r * nw + w]; m; m &= m - 1)
          out.emplace_back(row_map_[r], col_id(static_cast<VID>(w * 64 + __builtin_ctzll(m))));
    return out;
  }

 private:
  VID col_id(VID c) const { return col_map_ ? (*col_map_)[c] : c; }
  std::shared_ptr<vs_result_s> res_;
  uint32_t index_;
  std::vector<VID> row_map_;
  std::optional<std::vector<VID>> col_map_;
};

// ColMainMatrixTable (matrix_table.hh:85-125): a DenseColMainMatrix in HBM
// with optional row / column maps.
class ColMainMatrixTable {
 public:
  explicit ColMainMatrixTable(vs_colmat_t m)
      : m_(m, [](vs_colmat_s *x) { vs_colmat_free(x); }) {}
  // an R x C matrix from host words in the reference layout (C x align8(ceil(R/64)))
  static ColMainMatrixTable from_words(vs_graph_t g, uint32_t rows, uint32_t cols,
                                       const std::vector<uint64_t> &words) {
    vs_colmat_t m = nullptr;
    check(vs_colmat_create(g, rows, cols, words.empty() ? nullptr : words.data(), &m),
          "vs_colmat_create");
    return ColMainMatrixTable(m);
  }
  void set_maps(const std::optional<std::vector<VID>> &row_map,
                const std::optional<std::vector<VID>> &col_map) {
    check(vs_colmat_set_maps(m_.get(), row_map ? row_map->data() : nullptr,
                             col_map ? col_map->data() : nullptr,
                             col_map ? static_cast<uint32_t>(col_map->size()) : 0),
          "vs_colmat_set_maps");
  }
  uint32_t rows() const { return info().first; }
  uint32_t cols() const { return info().second; }
  bool get(uint32_t r, uint32_t c) const {
    int b = 0;
    check(vs_colmat_get(m_.get(), r, c, &b), "vs_colmat_get");
    return b != 0;
  }
  std::vector<uint64_t> export_words() const {
    uint64_t pitch = 0;
    uint32_t r = 0, c = 0;
    check(vs_colmat_info(m_.get(), &r, &c, &pitch, nullptr, nullptr), "vs_colmat_info");
    std::vector<uint64_t> out(static_cast<size_t>(c) * pitch);
    if (!out.empty())
      check(vs_colmat_export(m_.get(), out.data(), out.size() * 8), "vs_colmat_export");
    return out;
  }
  vs_colmat_s *handle() const { return m_.get(); }

 private:
  std::pair<uint32_t, uint32_t> info() const {
    uint32_t r = 0, c = 0;
    check(vs_colmat_info(m_.get(), &r, &c, nullptr, nullptr, nullptr), "vs_colmat_info");
    return {r, c};
  }
  std::shared_ptr<vs_colmat_s> m_;
};

// src/include/execution/tuple.hh:9-18 (the kinds TableConvert handles here)
enum class TableType { Flat, CSR, VarRowMatrix, ColMainMatrix };
// src/include/operator/big_operators2.hh:100-103
enum class SortOrder { ASC, DESC };

// TableConvert (big_operators2.hh:27-37, big_operators2.cpp:16-86):
// VarRowMatrix -> ColMainMatrix on the device, maps carried along and
// swapped when transposing.
class TableConvert {
 public:
  explicit TableConvert(TableType type, bool transpose = false)
      : convert_to_(type), transpose_(transpose) {
    if (type != TableType::ColMainMatrix)
      throw std::runtime_error("TableConvert: only VarRowMatrix -> ColMainMatrix runs on the GPU");
  }
  ColMainMatrixTable run(const VarRowMatrixTable &in) const {
    vs_colmat_t m = nullptr;
    check(vs_table_convert(in.handle(), in.index(), transpose_ ? 1 : 0, &m), "vs_table_convert");
    return ColMainMatrixTable(m);
  }

 private:
  TableType convert_to_;
  bool transpose_;
};

// MatrixTableSelect (big_operators2.hh:70-80, big_operators2.cpp:138-213):
// keeps the columns named by the key column of a flat dependency.
class MatrixTableSelect {
 public:
  explicit MatrixTableSelect(VariableName key_name) : key_name_(std::move(key_name)) {}
  VarRowMatrixTable run(const VarRowMatrixTable &in, const std::vector<VID> &keys) const {
    vs_result_t r = nullptr;
    check(vs_result_select_cols(in.handle(), in.index(), keys.data(),
                                static_cast<uint32_t>(keys.size()), &r),
          "vs_result_select_cols");
    return VarRowMatrixTable(std::shared_ptr<vs_result_s>(r, [](vs_result_s *x) { vs_result_free(x); }),
                             0, in.get_row_map(), keys);
  }

 private:
  VariableName key_name_;
};

// MatrixColumnCount (big_operators2.hh:112-129, matrix_column_count.cpp):
// a flat table (column id, Int32 count), optionally sorted by (count, id).
class MatrixColumnCount {
 public:
  explicit MatrixColumnCount(VariableName count_name, std::optional<SortOrder> ord = std::nullopt)
      : count_name_(std::move(count_name)), ord_(ord) {}
  std::vector<std::pair<VID, int32_t>> run(const VarRowMatrixTable &in) const {
    uint32_t rows = 0, cols = 0;
    check(vs_result_info(in.handle(), &rows, &cols, nullptr, nullptr), "vs_result_info");
    std::vector<uint32_t> ids(cols);
    std::vector<int32_t> cnt(cols);
    if (cols)
      check(vs_result_col_counts(in.handle(), in.index(), order(), ids.data(), cnt.data()),
            "vs_result_col_counts");
    return zip(ids, cnt);
  }
  std::vector<std::pair<VID, int32_t>> run(const ColMainMatrixTable &in) const {
    const uint32_t cols = in.cols();
    std::vector<uint32_t> ids(cols);
    std::vector<int32_t> cnt(cols);
    if (cols)
      check(vs_colmat_col_counts(in.handle(), order(), ids.data(), cnt.data()),
            "vs_colmat_col_counts");
    return zip(ids, cnt);
  }

 private:
  int32_t order() const {
    return !ord_ ? VS_ORDER_NONE : *ord_ == SortOrder::ASC ? VS_ORDER_ASC : VS_ORDER_DESC;
  }
  static std::vector<std::pair<VID, int32_t>> zip(const std::vector<uint32_t> &a,
                                                  const std::vector<int32_t> &b) {
    std::vector<std::pair<VID, int32_t>> out(a.size());
    for (size_t i = 0; i < a.size(); ++i) out[i] = {a[i], b[i]};
    return out;
  }
  VariableName count_name_;
  std::optional<SortOrder> ord_;
};

// MatrixRowOrSum (big_operators2.hh:131-138, matrix_row_or_sum.cpp): the ids
// of the columns with any set bit.
class MatrixRowOrSum {
 public:
  std::vector<VID> run(const VarRowMatrixTable &in) const {
    uint32_t n = 0;
    check(vs_result_row_or(in.handle(), in.index(), nullptr, 0, &n), "vs_result_row_or");
    std::vector<VID> out(n);
    if (n) check(vs_result_row_or(in.handle(), in.index(), out.data(), n, &n), "vs_result_row_or");
    return out;
  }
  std::vector<VID> run(const ColMainMatrixTable &in) const {
    uint32_t n = 0;
    check(vs_colmat_row_or(in.handle(), nullptr, 0, &n), "vs_colmat_row_or");
    std::vector<VID> out(n);
    if (n) check(vs_colmat_row_or(in.handle(), out.data(), n, &n), "vs_colmat_row_or");
    return out;
  }
};

// MultipleMatrixTable: one matrix per distance (bit_matrix_table.cpp:291-340).
struct MultipleMatrixTable {
  std::vector<std::pair<size_t, VarRowMatrixTable>> mats;
};

// FlatTable of VExpand's CSR strategy: columns (start.id, target.id[,
// distance]) (vexpand.cpp:103-164, 380-446), rows ordered by start row, then
// target.
struct FlatTable {
  std::vector<VID> start, target;
  std::vector<int32_t> distance;  // empty without with_distance
  size_t tuple_count() const { return start.size(); }
};

class VExpand {
 public:
  // big_operators.hh:101-105 / vexpand.cpp:7-14
  VExpand(TypeName src_type, TypeName dst_type, TypeName edge_type, Direction dir,
          size_t k_min, size_t k_max, ExpandPathType path_type, ExpandStrategy strategy,
          VariableName variable_name, std::optional<VariableName> target_variable_name)
      : VExpand(src_type, dst_type, edge_type, dir, k_min, k_max,
                VExpandPathParameters{path_type}, strategy, variable_name,
                target_variable_name) {}

  // For FindShortest, big_operators.hh:107-111 / vexpand.cpp:16-25
  VExpand(TypeName src_type, TypeName dst_type, TypeName edge_type, Direction dir,
          VExpandPathParameters path_params, ExpandStrategy strategy,
          VariableName variable_name, VariableName target_variable_name)
      : VExpand(src_type, dst_type, edge_type, dir, 1,
                VExpandPathParameters::K_MAX_FIND_SHORTEST, path_params, strategy,
                variable_name, std::optional<VariableName>(target_variable_name)) {
    if (path_params.path_type != ExpandPathType::FindShortest || !path_params.with_distance)
      throw std::runtime_error("FindShortest constructor needs FindShortest + with_distance");
  }

  // big_operators.hh:113-118 / vexpand.cpp:27-34
  VExpand(TypeName src_type, TypeName dst_type, TypeName edge_type, Direction dir,
          size_t k_min, size_t k_max, VExpandPathParameters path_params,
          ExpandStrategy strategy, VariableName variable_name,
          std::optional<VariableName> target_variable_name)
      : k_min_(k_min), k_max_(k_max), path_params_(path_params), strategy_(strategy),
        start_variable_name_(std::move(variable_name)),
        target_variable_name_(std::move(target_variable_name)) {
    rel_ = GraphStore::instance().get_edge(src_type, edge_type, dst_type);
    // a walk longer than one hop needs the relation to end where it starts
    // (EdgeSchema::is_self_edge, schema.hh:79; vexpand.cpp:88-91 exits)
    if (k_max > 1 && src_type != dst_type)
      throw std::runtime_error("VExpand: k_max must be 1 when the edge " + src_type + "-" +
                               edge_type + "-" + dst_type + " is not a self edge");
    // finishAddingDependency builds a FlatTable for CSR and a matrix table
    // for Dense; any other strategy asserts (vexpand.cpp:125-163)
    if (strategy != ExpandStrategy::CSR && strategy != ExpandStrategy::Dense)
      throw std::runtime_error("VExpand: the output strategy must be CSR or Dense");
    if (dir == Direction::Undecide) dir = Direction::Forward;
    dir_ = dir;
    // both strategies expand through the bit matrices; CSR hands its result
    // over as a flat table (flat_results), Dense as matrix tables
    if (k_max >= VS_MAX_HOPS) throw std::runtime_error("k_max too large");
  }

  // Runs with the dependency's start column (and target column for
  // FindShortest).  An empty start column gives an empty result, as the
  // reference's early exit (vexpand.cpp:170-173).
  void run(const std::vector<VID> &starts, const std::vector<VID> *targets = nullptr) {
    starts_ = starts;
    if (path_params_.path_type == ExpandPathType::FindShortest) {
      if (!targets || targets->size() != starts.size())
        throw std::runtime_error("FindShortest needs one target per start vertex");
      distance_.assign(starts.size(), -1);
      if (!starts.empty())
        check(vs_find_shortest(rel_->handle(), starts.data(), targets->data(),
                               static_cast<uint32_t>(starts.size()), dir_code(),
                               static_cast<uint32_t>(k_max_), distance_.data(), &stats_),
              "vs_find_shortest");
      k_actual_max_ = stats_.k_actual_max;
      return;
    }
    vs_vexpand_params p{};
    p.k_min = static_cast<uint32_t>(k_min_);
    p.k_max = static_cast<uint32_t>(k_max_);
    p.path_type = path_params_.path_type == ExpandPathType::Any ? VS_PATH_ANY : VS_PATH_SHORTEST;
    p.direction = dir_code();
    p.with_distance = path_params_.with_distance ? 1 : 0;
    p.sched = VS_SCHED_HEURISTIC;
    vs_result_t r = nullptr;
    check(vs_vexpand(rel_->handle(), starts.data(), static_cast<uint32_t>(starts.size()), &p,
                     &r, &stats_),
          "vs_vexpand");
    res_ = std::shared_ptr<vs_result_s>(r, [](vs_result_s *x) { vs_result_free(x); });
    k_actual_max_ = stats_.k_actual_max;
  }

  size_t k_actual_max() const { return k_actual_max_; }
  const vs_vexpand_stats &stats() const { return stats_; }

  // VarRowMatrixTable result (merge_to_MT)
  VarRowMatrixTable results() const {
    if (strategy_ != ExpandStrategy::Dense)
      throw std::runtime_error("VExpand: the CSR strategy's result is a FlatTable");
    if (!res_ || path_params_.with_distance)
      throw std::runtime_error("no merged matrix result");
    return VarRowMatrixTable(res_, 0, starts_);
  }
  // MultipleMatrixTable result (copy_to_multi_MT)
  MultipleMatrixTable multiple_results() const {
    if (strategy_ != ExpandStrategy::Dense)
      throw std::runtime_error("VExpand: the CSR strategy's result is a FlatTable");
    if (!res_ || !path_params_.with_distance)
      throw std::runtime_error("no per-distance result");
    MultipleMatrixTable m;
    uint32_t n = 0;
    check(vs_result_info(res_.get(), nullptr, nullptr, &n, nullptr), "vs_result_info");
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t d = 0;
      check(vs_result_distance(res_.get(), i, &d), "vs_result_distance");
      m.mats.emplace_back(d, VarRowMatrixTable(res_, i, starts_));
    }
    return m;
  }
  // FlatTable result (the CSR strategy's merge_to_Flat, vexpand.cpp:380-446)
  FlatTable flat_results() const {
    if (strategy_ != ExpandStrategy::CSR)
      throw std::runtime_error("VExpand: the Dense strategy's result is a matrix table");
    if (!res_) throw std::runtime_error("no result");
    FlatTable t;
    uint64_t n = 0;
    check(vs_result_flat(res_.get(), nullptr, nullptr, nullptr, 0, &n), "vs_result_flat");
    t.start.resize(n);
    t.target.resize(n);
    if (path_params_.with_distance) t.distance.resize(n);
    if (n)
      check(vs_result_flat(res_.get(), t.start.data(), t.target.data(),
                           path_params_.with_distance ? t.distance.data() : nullptr, n, &n),
            "vs_result_flat");
    return t;
  }
  ExpandStrategy strategy() const { return strategy_; }
  // FindShortest: the dependency's tuples extended with "distance" (-1 = none)
  const std::vector<int32_t> &distance() const { return distance_; }

 private:
  int32_t dir_code() const {
    return dir_ == Direction::Backward ? VS_DIR_BACKWARD
           : dir_ == Direction::Both   ? VS_DIR_BOTH
                                       : VS_DIR_FORWARD;
  }
  size_t k_min_, k_max_, k_actual_max_ = 0;
  VExpandPathParameters path_params_;
  ExpandStrategy strategy_;
  Direction dir_ = Direction::Forward;
  VariableName start_variable_name_;
  std::optional<VariableName> target_variable_name_;
  std::shared_ptr<EdgeRelation> rel_;
  std::vector<VID> starts_;
  std::shared_ptr<vs_result_s> res_;
  std::vector<int32_t> distance_;
  vs_vexpand_stats stats_{};
};

// Mintersect (big_operators.hh:161-183): base tuples of `arity` bound
// variables; dependency d maps x_d to its neighbour set over the new vertex.
class Mintersect {
 public:
  Mintersect(VariableName new_vertex_var_name, TypeName new_vertex_type, bool only_count)
      : new_vertex_var_name_(std::move(new_vertex_var_name)),
        new_vertex_type_(std::move(new_vertex_type)), only_count_(only_count) {}

  // One dependency per bound base column (addDependency order): all edge
  // relations (sorted adjacency) or all ColMain tables (mintersect.cpp:139-156).
  void add_dependency(std::shared_ptr<EdgeRelation> rel) { deps_.push_back(std::move(rel)); }
  void add_dependency(std::shared_ptr<ColMainMatrixTable> m) { mat_deps_.push_back(std::move(m)); }

  // base: T rows of arity = deps.size() vertex ids, row-major.
  void run(const std::vector<VID> &base) {
    if (!deps_.empty() && !mat_deps_.empty())
      throw std::runtime_error("Mintersect: mixed edge-relation and ColMain dependencies");
    const bool mats = !mat_deps_.empty();
    const uint32_t arity = static_cast<uint32_t>(mats ? mat_deps_.size() : deps_.size());
    if (arity == 0 || base.size() % arity) throw std::runtime_error("bad base tuples");
    const uint64_t T = base.size() / arity;
    std::vector<vs_graph_t> hs;
    std::vector<vs_colmat_t> ms;
    for (auto &d : deps_) hs.push_back(d->handle());
    for (auto &d : mat_deps_) ms.push_back(d->handle());
    auto call = [&](uint32_t *lists, uint64_t cap, uint64_t *total) {
      if (mats)
        check(vs_mintersect_colmat(ms.data(), arity, base.data(), T, counts_.data(),
                                   offsets_.data(), lists, cap, total, nullptr),
              "vs_mintersect_colmat");
      else
        check(vs_mintersect(hs.data(), arity, base.data(), T, counts_.data(), offsets_.data(),
                            lists, cap, total, nullptr),
              "vs_mintersect");
    };
    counts_.assign(T, 0);
    offsets_.assign(T + 1, 0);
    uint64_t total = 0;
    call(nullptr, 0, &total);
    total_num_ = total;
    tuples_.clear();
    if (only_count_) return;
    std::vector<uint32_t> lists(total);
    if (total) call(lists.data(), total, &total);
    // base columns + new vertex, tuples in base order, new ids ascending
    for (uint64_t t = 0; t < T; ++t)
      for (uint64_t k = offsets_[t]; k < offsets_[t + 1]; ++k) {
        for (uint32_t d = 0; d < arity; ++d) tuples_.push_back(base[t * arity + d]);
        tuples_.push_back(lists[k]);
      }
  }
  // VarRowMatrixTable base (mintersect.cpp:244-333): every set bit of matrix
  // `which` is a base tuple (start = the row's vertex, target = the column's
  // vertex); dependency d intersects the column bound to base variable
  // var_idx[d] (0 = start, 1 = target; the reference's var_base_idx).  Output
  // tuples (start, target, new vertex), by base tuple then new vertex.
  void run(const VarRowMatrixTable &base, const std::vector<uint32_t> &var_idx,
           uint32_t which = 0) {
    if (!deps_.empty() && !mat_deps_.empty())
      throw std::runtime_error("Mintersect: mixed edge-relation and ColMain dependencies");
    const bool mats = !mat_deps_.empty();
    const uint32_t nd = static_cast<uint32_t>(mats ? mat_deps_.size() : deps_.size());
    if (nd == 0 || var_idx.size() != nd)
      throw std::runtime_error("Mintersect: one base variable per dependency");
    std::vector<vs_graph_t> hs;
    std::vector<vs_colmat_t> ms;
    for (auto &d : deps_) hs.push_back(d->handle());
    for (auto &d : mat_deps_) ms.push_back(d->handle());
    const vs_colmat_t *mp = mats ? ms.data() : nullptr;
    const vs_graph_t *gp = mats ? nullptr : hs.data();
    uint64_t total = 0;
    check(vs_mintersect_varrow(base.handle(), which, mp, gp, nd, var_idx.data(), nullptr,
                               nullptr, nullptr, 0, &total, nullptr),
          "vs_mintersect_varrow");
    total_num_ = total;
    tuples_.clear();
    counts_.clear();
    offsets_.clear();
    if (only_count_ || total == 0) return;
    std::vector<uint32_t> st(total), tg(total), nv(total);
    check(vs_mintersect_varrow(base.handle(), which, mp, gp, nd, var_idx.data(), st.data(),
                               tg.data(), nv.data(), total, &total, nullptr),
          "vs_mintersect_varrow");
    tuples_.reserve(3 * total);
    for (uint64_t k = 0; k < total; ++k) {
      tuples_.push_back(st[k]);
      tuples_.push_back(tg[k]);
      tuples_.push_back(nv[k]);
    }
  }
  // only_count result: the "count" Int64 column (mintersect.cpp:336-340)
  uint64_t count() const { return total_num_; }
  const std::vector<VID> &flat_tuples() const { return tuples_; }
  size_t tuple_count() const { return only_count_ ? 1 : total_num_; }

 private:
  VariableName new_vertex_var_name_;
  TypeName new_vertex_type_;
  bool only_count_;
  std::vector<std::shared_ptr<EdgeRelation>> deps_;
  std::vector<std::shared_ptr<ColMainMatrixTable>> mat_deps_;
  std::vector<uint32_t> counts_;
  std::vector<uint64_t> offsets_;
  std::vector<VID> tuples_;
  uint64_t total_num_ = 0;
};

}  // namespace vertexsurge
