// Graph directories in the reference's storage format: schema.json plus one
// DiskArray file per variable.  Host code only (g++).
//
//   DiskArray file (disk_array.hh:267-354, 537-604): a 64-byte header whose
//   first two int64 are {element count, element size}, then the elements;
//   the file is stretched to 64 + capacity * element size, capacity being the
//   first-page capacity (4096 - 64) / size doubled until it holds the count.
//
//   schema.json: the GraphSchema JSON (schema.hh:10-141): graph_name,
//   graph_type, vertex[] {type, statistics{count}, properties}, edge[]
//   {type, directed, allow_multiple_edge, src, dst, properties,
//   statistics{count}, store_info{flats, partitioned_flats,
//   partition_counts}}.
//
//   An edge relation src -type-> dst is stored as a CSR from its source side
//   (graph_loader.hh:163-182): "<fwd>_csr_src_dst_index" (int64 offsets, one
//   per source vertex + 1) and "<fwd>_csr_src_dst" (uint32 neighbour ids),
//   where <fwd> is "src-type->dst" for directed and "src-type-dst" for
//   undirected relations (schema.cpp:11-49).  A directed relation also
//   stores the reverse CSR under "<bwd>_csr_dst_src[_index]", <bwd> =
//   "dst<-type-src" (schema.cpp:163-205 lists the files Graph::load opens).
//
// The importer removes only adjacent duplicate edges before sorting
// (graph_loader.hh:71-77, graph_concepts.hh:104-107) and keeps multi-edges
// of relations that allow them, so a stored list may repeat an id.  Every
// operator on the path treats a list as a set (VExpand ORs rows, Mintersect
// ANDs bitmaps), so opening a store sorts any list that is not ascending and
// drops repeats; the counts of both are reported.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "vs_internal.hh"

namespace vs {
namespace {

// ------------------------------------------------------------------ JSON
// A small reader for the subset schema.json uses: objects, arrays, strings,
// integers, reals, booleans, null.
struct JValue {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  bool is_int = false;
  uint64_t u = 0;  // integer value when is_int (non-negative)
  double d = 0;
  std::string s;
  std::vector<JValue> a;
  std::vector<std::pair<std::string, JValue>> o;

  const JValue *get(const char *key) const {
    if (kind != Obj) return nullptr;
    for (const auto &kv : o)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JParser {
 public:
  explicit JParser(const std::string &t) : t_(t) {}
  bool parse(JValue &v) {
    ws();
    if (!value(v, 0)) return false;
    ws();
    if (p_ != t_.size()) return fail("trailing characters");
    return true;
  }
  std::string error() const { return err_; }

 private:
  const std::string &t_;
  size_t p_ = 0;
  std::string err_;

  bool fail(const char *what) {
    if (err_.empty()) err_ = std::string(what) + " at byte " + std::to_string(p_);
    return false;
  }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t'))
      ++p_;
  }
  bool lit(const char *w) {
    const size_t n = strlen(w);
    if (t_.compare(p_, n, w) != 0) return fail("bad literal");
    p_ += n;
    return true;
  }
  static void utf8(std::string &out, uint32_t c) {
    if (c < 0x80) {
      out += static_cast<char>(c);
    } else if (c < 0x800) {
      out += static_cast<char>(0xC0 | (c >> 6));
      out += static_cast<char>(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      out += static_cast<char>(0xE0 | (c >> 12));
      out += static_cast<char>(0x80 | ((c >> 6) & 0x3F));
      out += static_cast<char>(0x80 | (c & 0x3F));
    } else {
      out += static_cast<char>(0xF0 | (c >> 18));
      out += static_cast<char>(0x80 | ((c >> 12) & 0x3F));
      out += static_cast<char>(0x80 | ((c >> 6) & 0x3F));
      out += static_cast<char>(0x80 | (c & 0x3F));
    }
  }
  bool hex4(uint32_t &c) {
    if (p_ + 4 > t_.size()) return fail("short \\u escape");
    c = 0;
    for (int i = 0; i < 4; ++i) {
      const char h = t_[p_++];
      c <<= 4;
      if (h >= '0' && h <= '9') c |= h - '0';
      else if (h >= 'a' && h <= 'f') c |= h - 'a' + 10;
      else if (h >= 'A' && h <= 'F') c |= h - 'A' + 10;
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool str(std::string &out) {
    if (t_[p_] != '"') return fail("expected string");
    ++p_;
    while (p_ < t_.size()) {
      const char c = t_[p_++];
      if (c == '"') return true;
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= t_.size()) break;
      const char e = t_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t c1;
          if (!hex4(c1)) return false;
          if (c1 >= 0xD800 && c1 < 0xDC00 && p_ + 1 < t_.size() && t_[p_] == '\\' &&
              t_[p_ + 1] == 'u') {
            p_ += 2;
            uint32_t c2;
            if (!hex4(c2)) return false;
            if (c2 < 0xDC00 || c2 >= 0xE000) return fail("bad surrogate pair");
            c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
          } else if (c1 >= 0xD800 && c1 < 0xE000) {
            return fail("unpaired surrogate");
          }
          utf8(out, c1);
          break;
        }
        default: return fail("bad escape");
      }
    }
    return fail("unterminated string");
  }
  bool num(JValue &v) {
    const size_t b = p_;
    if (t_[p_] == '-') ++p_;
    bool frac = false;
    while (p_ < t_.size()) {
      const char c = t_[p_];
      if (c >= '0' && c <= '9') ++p_;
      else if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') { frac = true; ++p_; }
      else break;
    }
    const std::string tok = t_.substr(b, p_ - b);
    if (tok.empty() || tok == "-") return fail("bad number");
    v.kind = JValue::Num;
    v.d = strtod(tok.c_str(), nullptr);
    if (!frac && tok[0] != '-') {
      errno = 0;
      v.u = strtoull(tok.c_str(), nullptr, 10);
      v.is_int = errno == 0;
    }
    return true;
  }
  bool value(JValue &v, int depth) {
    if (depth > 64) return fail("nesting too deep");
    if (p_ >= t_.size()) return fail("unexpected end");
    const char c = t_[p_];
    if (c == '{') {
      ++p_;
      v.kind = JValue::Obj;
      ws();
      if (p_ < t_.size() && t_[p_] == '}') { ++p_; return true; }
      while (true) {
        ws();
        std::string k;
        if (p_ >= t_.size() || !str(k)) return fail("expected key");
        ws();
        if (p_ >= t_.size() || t_[p_] != ':') return fail("expected ':'");
        ++p_;
        ws();
        v.o.emplace_back(std::move(k), JValue{});
        if (!value(v.o.back().second, depth + 1)) return false;
        ws();
        if (p_ < t_.size() && t_[p_] == ',') { ++p_; continue; }
        if (p_ < t_.size() && t_[p_] == '}') { ++p_; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p_;
      v.kind = JValue::Arr;
      ws();
      if (p_ < t_.size() && t_[p_] == ']') { ++p_; return true; }
      while (true) {
        ws();
        v.a.emplace_back();
        if (!value(v.a.back(), depth + 1)) return false;
        ws();
        if (p_ < t_.size() && t_[p_] == ',') { ++p_; continue; }
        if (p_ < t_.size() && t_[p_] == ']') { ++p_; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') { v.kind = JValue::Str; return str(v.s); }
    if (c == 't') { v.kind = JValue::Bool; v.b = true; return lit("true"); }
    if (c == 'f') { v.kind = JValue::Bool; v.b = false; return lit("false"); }
    if (c == 'n') { v.kind = JValue::Null; return lit("null"); }
    return num(v);
  }
};

void json_escape(std::string &out, const std::string &s) {
  out += '"';
  for (const unsigned char c : s) {
    if (c == '"') out += "\\\"";
    else if (c == '\\') out += "\\\\";
    else if (c == '\n') out += "\\n";
    else if (c == '\t') out += "\\t";
    else if (c < 0x20) {
      char b[8];
      snprintf(b, sizeof b, "\\u%04x", c);
      out += b;
    } else out += static_cast<char>(c);
  }
  out += '"';
}

// --------------------------------------------------------------- schema
struct EdgeRel {
  std::string type, src, dst;
  bool directed = false, multi = false;
  uint64_t count = 0;
  std::string fwd_name() const {  // EdgeSchema::get_forward_edge_name
    return directed ? src + "-" + type + "->" + dst : src + "-" + type + "-" + dst;
  }
  std::string bwd_name() const {  // EdgeSchema::get_backward_edge_name
    return directed ? dst + "<-" + type + "-" + src : dst + "-" + type + "-" + src;
  }
};

bool read_text(const std::string &path, std::string &out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

vs_status load_schema(const std::string &dir, std::map<std::string, uint64_t> &vcount,
                      std::vector<EdgeRel> &edges) {
  std::string text;
  const std::string path = dir + "/schema.json";
  if (!read_text(path, text)) {
    set_error("vs_store_open: cannot read %s: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  JValue root;
  JParser jp(text);
  if (!jp.parse(root) || root.kind != JValue::Obj) {
    set_error("vs_store_open: %s: %s", path.c_str(),
              jp.error().empty() ? "not a JSON object" : jp.error().c_str());
    return VS_ERR_INVALID;
  }
  const JValue *vx = root.get("vertex"), *ex = root.get("edge");
  if (!vx || vx->kind != JValue::Arr || !ex || ex->kind != JValue::Arr) {
    set_error("vs_store_open: %s has no vertex/edge arrays", path.c_str());
    return VS_ERR_INVALID;
  }
  auto count_of = [](const JValue &o, uint64_t &c) {
    const JValue *st = o.get("statistics");
    const JValue *n = st ? st->get("count") : nullptr;
    if (!n || n->kind != JValue::Num || !n->is_int) return false;
    c = n->u;
    return true;
  };
  for (const JValue &v : vx->a) {
    const JValue *t = v.get("type");
    uint64_t c;
    if (!t || t->kind != JValue::Str || !count_of(v, c)) {
      set_error("vs_store_open: a vertex schema lacks type or statistics.count");
      return VS_ERR_INVALID;
    }
    vcount[t->s] = c;
  }
  for (const JValue &e : ex->a) {
    EdgeRel r;
    const JValue *t = e.get("type"), *s = e.get("src"), *d = e.get("dst"),
                 *dir_ = e.get("directed"), *m = e.get("allow_multiple_edge");
    if (!t || t->kind != JValue::Str || !s || s->kind != JValue::Str || !d ||
        d->kind != JValue::Str || !dir_ || dir_->kind != JValue::Bool || !count_of(e, r.count)) {
      set_error("vs_store_open: an edge schema lacks type/src/dst/directed/statistics.count");
      return VS_ERR_INVALID;
    }
    r.type = t->s;
    r.src = s->s;
    r.dst = d->s;
    r.directed = dir_->b;
    r.multi = m && m->kind == JValue::Bool && m->b;
    edges.push_back(r);
  }
  return VS_OK;
}

// ------------------------------------------------------------ DiskArray
struct Mapped {
  void *base = nullptr;
  size_t len = 0;
  void reset() {
    if (base) munmap(base, len);
    base = nullptr;
    len = 0;
  }
  ~Mapped() { reset(); }
};

constexpr size_t kHeader = 64;
constexpr size_t kPage = 4096;

// Maps a DiskArray file read-only and checks its header against the element
// size and count the schema implies.
vs_status map_array(const std::string &path, size_t elem, uint64_t expect, Mapped &m,
                    const void **data) {
  const int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    set_error("vs_store_open: cannot open %s: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || static_cast<size_t>(st.st_size) < kHeader) {
    close(fd);
    set_error("vs_store_open: %s is shorter than the 64-byte DiskArray header", path.c_str());
    return VS_ERR_INVALID;
  }
  m.len = static_cast<size_t>(st.st_size);
  m.base = mmap(nullptr, m.len, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (m.base == MAP_FAILED) {
    m.base = nullptr;
    set_error("vs_store_open: mmap %s: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  int64_t hdr[2];
  memcpy(hdr, m.base, sizeof hdr);
  if (hdr[1] != static_cast<int64_t>(elem)) {
    set_error("vs_store_open: %s holds %lld-byte elements, expected %zu", path.c_str(),
              static_cast<long long>(hdr[1]), elem);
    return VS_ERR_INVALID;
  }
  if (hdr[0] < 0 || static_cast<uint64_t>(hdr[0]) != expect) {
    set_error("vs_store_open: %s holds %lld elements, the schema implies %llu", path.c_str(),
              static_cast<long long>(hdr[0]), static_cast<unsigned long long>(expect));
    return VS_ERR_INVALID;
  }
  if ((m.len - kHeader) / elem < expect) {
    set_error("vs_store_open: %s is truncated (%zu bytes for %llu elements)", path.c_str(),
              m.len, static_cast<unsigned long long>(expect));
    return VS_ERR_INVALID;
  }
  madvise(m.base, m.len, MADV_SEQUENTIAL);
  *data = static_cast<const char *>(m.base) + kHeader;
  return VS_OK;
}

// Writes a DiskArray file the way DiskArray(path, n, true) sizes it.
vs_status write_array(const std::string &path, const void *data, size_t elem, uint64_t n) {
  uint64_t cap = (kPage - kHeader) / elem;
  while (cap < n) cap *= 2;
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    set_error("vs_store_write: cannot create %s: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  unsigned char hdr[kHeader] = {};
  const int64_t h[2] = {static_cast<int64_t>(n), static_cast<int64_t>(elem)};
  memcpy(hdr, h, sizeof h);
  bool ok = fwrite(hdr, 1, kHeader, f) == kHeader;
  const size_t bytes = static_cast<size_t>(n) * elem;
  const char *p = static_cast<const char *>(data);
  for (size_t o = 0; ok && o < bytes; o += (1u << 24)) {
    const size_t k = std::min<size_t>(1u << 24, bytes - o);
    ok = fwrite(p + o, 1, k, f) == k;
  }
  // the stretched tail up to the capacity reads as zeros
  const size_t total = kHeader + static_cast<size_t>(cap) * elem;
  if (ok && total > kHeader + bytes) ok = fseek(f, static_cast<long>(total - 1), SEEK_SET) == 0 &&
                                          fputc(0, f) != EOF;
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    set_error("vs_store_write: writing %s failed: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  return VS_OK;
}

// ----------------------------------------------------------- flat edges
#define FLAT_TRY(call)                  \
  do {                                  \
    const vs_status st_ = (call);       \
    if (st_ != VS_OK) return st_;       \
  } while (0)

// Position of (x, y) on the Peano-Hilbert curve of an n x n grid (n a power
// of two) that starts in the (0, 0) quadrant and turns clockwise (the
// reference's XYToIndex(n, x, y, PeanoDirection::RA), peano.hh:20-89).  A
// curve state is one of four traversals of the quadrants q = 2 * xbit + ybit;
// kVisit[state][q] is the position at which that traversal visits q, and
// kNext[state][position] the traversal inside the quadrant visited there.
// States: 0 clockwise from (0,0), 1 clockwise from (1,1), 2 counter-clockwise
// from (0,0), 3 counter-clockwise from (1,1).
uint64_t peano_key(uint64_t n, uint64_t x, uint64_t y) {
  static constexpr uint8_t kVisit[4][4] = {{0, 1, 3, 2}, {2, 3, 1, 0}, {0, 3, 1, 2},
                                           {2, 1, 3, 0}};
  static constexpr uint8_t kNext[4][4] = {{2, 0, 0, 3}, {3, 1, 1, 2}, {0, 2, 2, 1},
                                          {1, 3, 3, 0}};
  uint64_t key = 0;
  int state = 0;
  for (uint64_t h = n >> 1; h; h >>= 1) {
    const int q = ((x & h) ? 2 : 0) | ((y & h) ? 1 : 0);
    const int pos = kVisit[state][q];
    key += static_cast<uint64_t>(pos) * h * h;
    state = kNext[state][pos];
  }
  return key;
}

struct FlatEdge {
  uint32_t src, dst;
};

// Edge boundaries of `count` destination partitions over edges sorted by
// (dst, src) (partition_by_dst, graph_loader.hh:111-139): partition i takes
// edges until it reaches an even share of what is left and the destination
// changes, so no destination is split.  When fewer edges than partitions
// remain, the last partitions are empty (the reference reads past the end
// there).
std::vector<uint64_t> dst_partitions(const std::vector<FlatEdge> &e, uint32_t count) {
  std::vector<uint64_t> cut{0};
  uint64_t pos = 0;
  const uint64_t n = e.size();
  for (uint32_t i = 0; i < count; ++i) {
    if (pos < n) {
      const uint64_t expected = (n - pos) / (count - i) + pos;
      do {
        ++pos;
      } while (pos < n && !(pos >= expected && e[pos].dst != e[pos - 1].dst));
    }
    cut.push_back(pos);
  }
  return cut;
}

enum FlatOrder { kOrderHilbert = 0, kOrderSrcDst = 1 };  // FlatEdgeOrder values

// Writes edges [b, e) as the DiskArray<VIDPair> file `path`; the backward
// files hold every pair reversed (write_vid_pair_to_disk, graph_loader.hh:86-108).
vs_status write_pairs(const std::string &path, const FlatEdge *b, uint64_t n, bool backward) {
  if (!backward) return write_array(path, b, sizeof(FlatEdge), n);
  std::vector<FlatEdge> r(n);
  for (uint64_t i = 0; i < n; ++i) r[i] = {b[i].dst, b[i].src};
  return write_array(path, r.data(), sizeof(FlatEdge), n);
}

void sort_src_dst(FlatEdge *b, FlatEdge *e) {
  std::sort(b, e, [](const FlatEdge &x, const FlatEdge &y) {
    return x.src != y.src ? x.src < y.src : x.dst < y.dst;
  });
}

// Sorts by curve position; equal positions are equal pairs.
void sort_hilbert(FlatEdge *b, FlatEdge *e, uint64_t n) {
  std::vector<std::pair<uint64_t, FlatEdge>> k(static_cast<size_t>(e - b));
  for (size_t i = 0; i < k.size(); ++i) k[i] = {peano_key(n, b[i].src, b[i].dst), b[i]};
  std::sort(k.begin(), k.end(),
            [](const auto &x, const auto &y) { return x.first < y.first; });
  for (size_t i = 0; i < k.size(); ++i) b[i] = k[i].second;
}

// The flat edge files of one relation (gen_and_write_one_edge,
// graph_loader.hh:186-276, with store_info flats {src_dst, hilbert} and
// partitioned_flats {src_dst, hilbert} x `parts`, as the reference's importers
// set them up, graph_loader.cpp:216-223): for each stored direction,
// <name>_src_dst and <name>_hilbert over all edges, and <name>_src_dst_<i> /
// <name>_hilbert_<i> over the destination partitions.
vs_status write_flats(const std::string &d, const EdgeRel &r, const uint64_t *off,
                      const uint32_t *adj, uint32_t V, uint64_t E, uint32_t parts) {
  std::vector<FlatEdge> e(E);
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t k = off[u]; k < off[u + 1]; ++k) e[k] = {u, adj[k]};
  uint64_t n = 1;
  while (n < V) n <<= 1;  // round_to_2_power(max(src count, dst count))
  std::vector<std::pair<std::string, bool>> dirs{{r.fwd_name(), false}};
  if (r.directed) dirs.push_back({r.bwd_name(), true});
  sort_src_dst(e.data(), e.data() + E);
  for (const auto &[name, bwd] : dirs) FLAT_TRY(write_pairs(d + "/" + name + "_src_dst", e.data(), E, bwd));
  sort_hilbert(e.data(), e.data() + E, n);
  for (const auto &[name, bwd] : dirs) FLAT_TRY(write_pairs(d + "/" + name + "_hilbert", e.data(), E, bwd));
  std::sort(e.begin(), e.end(), [](const FlatEdge &x, const FlatEdge &y) {
    return x.dst != y.dst ? x.dst < y.dst : x.src < y.src;
  });
  const std::vector<uint64_t> cut = dst_partitions(e, parts);
  for (uint32_t i = 0; i < parts; ++i) {
    FlatEdge *b = e.data() + cut[i], *f = e.data() + cut[i + 1];
    const std::string sfx = "_" + std::to_string(i);
    sort_src_dst(b, f);
    for (const auto &[name, bwd] : dirs)
      FLAT_TRY(write_pairs(d + "/" + name + "_src_dst" + sfx, b, f - b, bwd));
    sort_hilbert(b, f, n);
    for (const auto &[name, bwd] : dirs)
      FLAT_TRY(write_pairs(d + "/" + name + "_hilbert" + sfx, b, f - b, bwd));
  }
  return VS_OK;
}

}  // namespace
}  // namespace vs

struct vs_store_s {
  vs::EdgeRel rel;
  uint32_t V = 0;
  vs::Mapped fidx, fadj;
  const uint64_t *off = nullptr;  // file offsets (validated) or owned copy
  const uint32_t *adj = nullptr;
  std::vector<uint64_t> own_off;
  std::vector<uint32_t> own_adj;
  vs_store_info info{};
};

using namespace vs;

extern "C" {

vs_status vs_store_open(const char *dir, const char *edge, vs_store_t *out) try {
  clear_error();
  if (!dir || !out) {
    set_error("vs_store_open: null argument");
    return VS_ERR_INVALID;
  }
  *out = nullptr;
  std::map<std::string, uint64_t> vcount;
  std::vector<EdgeRel> rels;
  vs_status st = load_schema(dir, vcount, rels);
  if (st != VS_OK) return st;

  // GraphSchema::get_edge_schema (schema.cpp:222-232) matches the forward or
  // backward simple name "src_type_dst"; the bare type name is accepted too.
  // With no name, the graph's only relation of a vertex type to itself.
  const bool named = edge && *edge;
  std::vector<const EdgeRel *> hits;
  std::string names;
  for (const EdgeRel &r : rels) {
    names += (names.empty() ? "" : ", ") + r.src + "_" + r.type + "_" + r.dst;
    const bool hit = named ? (r.type == edge || r.src + "_" + r.type + "_" + r.dst == edge ||
                              r.dst + "_" + r.type + "_" + r.src == edge)
                           : r.src == r.dst;
    if (hit) hits.push_back(&r);
  }
  if (hits.size() != 1) {
    if (named)
      set_error("vs_store_open: %s: %zu edge relations match '%s' (relations: %s)", dir,
                hits.size(), edge, names.c_str());
    else
      set_error("vs_store_open: %s: name the edge relation (%zu relations of a vertex type to "
                "itself; relations: %s)", dir, hits.size(), names.c_str());
    return VS_ERR_INVALID;
  }
  const EdgeRel *pick = hits[0];
  if (pick->src != pick->dst) {
    // VExpand walks k > 1 hops only over a relation of a vertex type to
    // itself (vexpand.cpp:88-91); the device graph is one vertex set.
    set_error("vs_store_open: relation %s joins %s to %s; only relations of a vertex type to "
              "itself load as a device graph", pick->type.c_str(), pick->src.c_str(),
              pick->dst.c_str());
    return VS_ERR_UNSUPPORTED;
  }
  const auto vc = vcount.find(pick->src);
  if (vc == vcount.end()) {
    set_error("vs_store_open: vertex type %s of relation %s is not in the schema",
              pick->src.c_str(), pick->type.c_str());
    return VS_ERR_INVALID;
  }
  if (vc->second >= 0xFFFFFFFFull) {
    set_error("vs_store_open: %llu vertices do not fit 32-bit ids",
              static_cast<unsigned long long>(vc->second));
    return VS_ERR_INVALID;
  }

  std::unique_ptr<vs_store_s> s(new (std::nothrow) vs_store_s);
  if (!s) {
    set_error("vs_store_open: out of host memory");
    return VS_ERR_NOMEM;
  }
  s->rel = *pick;
  s->V = static_cast<uint32_t>(vc->second);
  const uint32_t V = s->V;
  const uint64_t E = pick->count;
  const std::string base = std::string(dir) + "/" + pick->fwd_name() + "_csr_src_dst";
  const void *pi = nullptr, *pa = nullptr;
  if ((st = map_array(base + "_index", 8, static_cast<uint64_t>(V) + 1, s->fidx, &pi)) != VS_OK)
    return st;
  if ((st = map_array(base, 4, E, s->fadj, &pa)) != VS_OK) return st;
  const int64_t *idx = static_cast<const int64_t *>(pi);
  const uint32_t *adj = static_cast<const uint32_t *>(pa);

  // offsets: 0 .. E, non-decreasing; ids < V
  if (idx[0] != 0 || static_cast<uint64_t>(idx[V]) != E) {
    set_error("vs_store_open: %s_index runs %lld..%lld, expected 0..%llu", base.c_str(),
              static_cast<long long>(idx[0]), static_cast<long long>(idx[V]),
              static_cast<unsigned long long>(E));
    return VS_ERR_INVALID;
  }
  uint64_t bad_off = 0, bad_id = 0, unsorted = 0, dups = 0, loops = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad_off, bad_id, unsorted, dups, loops)
  for (int64_t u = 0; u < static_cast<int64_t>(V); ++u) {
    const int64_t b = idx[u], e = idx[u + 1];
    if (e < b || e > static_cast<int64_t>(E)) {
      ++bad_off;
      continue;
    }
    bool asc = true;
    for (int64_t i = b; i < e; ++i) {
      const uint32_t x = adj[i];
      if (x >= V) ++bad_id;
      if (x == static_cast<uint32_t>(u)) ++loops;
      if (i > b) {
        if (x < adj[i - 1]) asc = false;
        else if (x == adj[i - 1]) ++dups;
      }
    }
    if (!asc) ++unsorted;
  }
  if (bad_off || bad_id) {
    set_error("vs_store_open: %s: %llu decreasing offsets, %llu neighbour ids >= V=%u",
              base.c_str(), static_cast<unsigned long long>(bad_off),
              static_cast<unsigned long long>(bad_id), V);
    return VS_ERR_INVALID;
  }
  s->info.num_vertices = V;
  s->info.stored_entries = E;
  s->info.directed = pick->directed ? 1 : 0;
  s->info.allow_multiple_edge = pick->multi ? 1 : 0;
  s->info.unsorted_lists = unsorted;
  s->info.self_loops = loops;  // counted before repeats are dropped

  if (!unsorted && !dups) {
    // the files already hold sets: use the mapping as is (int64 offsets that
    // passed the checks above have the same bits as uint64)
    s->off = reinterpret_cast<const uint64_t *>(idx);
    s->adj = adj;
    s->info.num_edges = E;
  } else {
    // sort the lists that need it and drop repeats, in two passes
    try {
      s->own_off.assign(static_cast<size_t>(V) + 1, 0);
      s->own_adj.resize(E);
    } catch (...) {
      set_error("vs_store_open: out of host memory");
      return VS_ERR_NOMEM;
    }
    uint64_t *no = s->own_off.data();
    uint32_t *na = s->own_adj.data();
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t u = 0; u < static_cast<int64_t>(V); ++u) {
      const int64_t b = idx[u], e = idx[u + 1];
      uint32_t *w = na + b;  // each list compacts inside its own input span
      std::copy(adj + b, adj + e, w);
      std::sort(w, w + (e - b));
      no[u + 1] = static_cast<uint64_t>(std::unique(w, w + (e - b)) - w);
    }
    for (uint32_t u = 0; u < V; ++u) no[u + 1] += no[u];
    // move each compacted list down to its final place (ascending order, so a
    // list never overwrites one that has not moved yet)
    for (uint32_t u = 0; u < V; ++u) {
      const uint64_t n = no[u + 1] - no[u];
      if (no[u] != static_cast<uint64_t>(idx[u]))
        memmove(na + no[u], na + idx[u], n * sizeof(uint32_t));
    }
    s->own_adj.resize(no[V]);
    s->own_adj.shrink_to_fit();
    s->off = no;
    s->adj = s->own_adj.data();
    s->info.num_edges = no[V];
    s->info.duplicates = E - no[V];
    s->fidx.reset();
    s->fadj.reset();
  }
  *out = s.release();
  return VS_OK;
} VS_CATCH("vs_store_open")

vs_status vs_store_get_info(vs_store_t s, vs_store_info *info) try {
  if (!s || !info) {
    set_error("vs_store_get_info: null argument");
    return VS_ERR_INVALID;
  }
  *info = s->info;
  return VS_OK;
} VS_CATCH("vs_store_get_info")

vs_status vs_store_csr(vs_store_t s, const uint64_t **offsets, const uint32_t **adj) try {
  if (!s || !offsets || !adj) {
    set_error("vs_store_csr: null argument");
    return VS_ERR_INVALID;
  }
  *offsets = s->off;
  *adj = s->adj;
  return VS_OK;
} VS_CATCH("vs_store_csr")

vs_status vs_store_graph(vs_store_t s, vs_graph_t *out) try {
  if (!s || !out) {
    set_error("vs_store_graph: null argument");
    return VS_ERR_INVALID;
  }
  // an undirected relation is stored with both directions of every edge
  // (add_reverse_edges, graph_concepts.hh:96-103), so its CSR is its own
  // reverse, as Graph::get_csr_edge_refs assumes (schema.cpp:287-300)
  return vs_graph_create(s->off, s->adj, s->V, s->info.num_edges,
                         s->rel.directed ? 0u : VS_GRAPH_SYMMETRIC, out);
} VS_CATCH("vs_store_graph")

vs_status vs_store_close(vs_store_t s) try {
  delete s;
  return VS_OK;
} VS_CATCH("vs_store_close")

vs_status vs_store_write(const char *dir, const char *graph_name, const char *vertex_type,
                         const char *edge_type, int32_t directed, const uint64_t *offsets,
                         const uint32_t *adj, uint32_t V, uint64_t E,
                         uint32_t flat_partitions) try {
  clear_error();
  if (!dir || !vertex_type || !edge_type || !*vertex_type || !*edge_type || !offsets ||
      (E && !adj)) {
    set_error("vs_store_write: invalid arguments");
    return VS_ERR_INVALID;
  }
  if (offsets[0] != 0 || offsets[V] != E) {
    set_error("vs_store_write: offsets run %llu..%llu, expected 0..%llu",
              static_cast<unsigned long long>(offsets[0]),
              static_cast<unsigned long long>(offsets[V]), static_cast<unsigned long long>(E));
    return VS_ERR_INVALID;
  }
  for (uint32_t u = 0; u < V; ++u)
    if (offsets[u + 1] < offsets[u]) {
      set_error("vs_store_write: offsets decrease at vertex %u", u);
      return VS_ERR_INVALID;
    }
  for (uint64_t i = 0; i < E; ++i)
    if (adj[i] >= V) {
      set_error("vs_store_write: neighbour id %u >= V=%u at entry %llu", adj[i], V,
                static_cast<unsigned long long>(i));
      return VS_ERR_INVALID;
    }
  struct stat st;
  if (stat(dir, &st) != 0 || !S_ISDIR(st.st_mode)) {
    set_error("vs_store_write: %s is not a directory", dir);
    return VS_ERR_IO;
  }
  EdgeRel r;
  r.type = edge_type;
  r.src = r.dst = vertex_type;
  r.directed = directed != 0;
  const std::string d(dir);
  vs_status s = write_array(d + "/" + r.fwd_name() + "_csr_src_dst_index", offsets, 8,
                            static_cast<uint64_t>(V) + 1);
  if (s == VS_OK) s = write_array(d + "/" + r.fwd_name() + "_csr_src_dst", adj, 4, E);
  if (s == VS_OK && r.directed) {
    std::vector<uint64_t> roff;
    std::vector<uint32_t> radj;
    try {
      reverse_csr(offsets, adj, V, E, roff, radj);
    } catch (const std::bad_alloc &) {
      set_error("vs_store_write: out of host memory for the reverse CSR");
      return VS_ERR_NOMEM;
    }
    s = write_array(d + "/" + r.bwd_name() + "_csr_dst_src_index", roff.data(), 8,
                    static_cast<uint64_t>(V) + 1);
    if (s == VS_OK) s = write_array(d + "/" + r.bwd_name() + "_csr_dst_src", radj.data(), 4, E);
  }
  if (s != VS_OK) return s;
  if (flat_partitions) {
    try {
      FLAT_TRY(write_flats(d, r, offsets, adj, V, E, flat_partitions));
    } catch (const std::bad_alloc &) {
      set_error("vs_store_write: out of host memory for the flat edge files");
      return VS_ERR_NOMEM;
    }
  }
  // schema.json with the field names of the NLOHMANN_DEFINE_TYPE_INTRUSIVE
  // lists (schema.hh:32, 121-123, 137-138); FlatEdgeOrder values go out as
  // their enum integers (Hilbert 0, SrcDst 1; schema.hh:50-55), the member
  // to_json being no ADL candidate
  std::string flats = "[]", pcounts = "[]";
  if (flat_partitions) {
    flats = "[\n                    1,\n                    0\n                ]";
    const std::string p = std::to_string(flat_partitions);
    pcounts = "[\n                    " + p + ",\n                    " + p + "\n                ]";
  }
  std::string j = "{\n    \"edge\": [\n        {\n            \"allow_multiple_edge\": false,\n"
                  "            \"directed\": ";
  j += r.directed ? "true" : "false";
  j += ",\n            \"dst\": ";
  json_escape(j, r.dst);
  j += ",\n            \"properties\": {},\n            \"src\": ";
  json_escape(j, r.src);
  j += ",\n            \"statistics\": {\n                \"count\": " + std::to_string(E) +
       "\n            },\n            \"store_info\": {\n                \"flats\": " + flats +
       ",\n                \"partition_counts\": " + pcounts +
       ",\n                \"partitioned_flats\": " + flats + "\n            },\n            \"type\": ";
  json_escape(j, r.type);
  j += "\n        }\n    ],\n    \"graph_name\": ";
  json_escape(j, graph_name ? graph_name : "");
  j += ",\n    \"graph_type\": \"\",\n    \"vertex\": [\n        {\n"
       "            \"properties\": {},\n            \"statistics\": {\n"
       "                \"count\": " + std::to_string(V) + "\n            },\n"
       "            \"type\": ";
  json_escape(j, r.src);
  j += "\n        }\n    ]\n}\n";
  const std::string path = d + "/schema.json";
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    set_error("vs_store_write: cannot create %s: %s", path.c_str(), strerror(errno));
    return VS_ERR_IO;
  }
  const bool ok = fwrite(j.data(), 1, j.size(), f) == j.size();
  if ((fclose(f) != 0) || !ok) {
    set_error("vs_store_write: writing %s failed", path.c_str());
    return VS_ERR_IO;
  }
  return VS_OK;
} VS_CATCH("vs_store_write")

}  // extern "C"
