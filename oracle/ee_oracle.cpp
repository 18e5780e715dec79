// CPU oracle for the JV edge test and the k-opt exact cost matrix.
//
// TEST INFRASTRUCTURE ONLY: a restatement of the reference CPU path used to
// check and time the HIP product library.  The product never links it.  Each
// function names the reference file:line whose behaviour it restates.
#include "ee_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

int Fail(const std::string& message) {
  g_error = message;
  return -1;
}

using i128 = __int128;

constexpr std::int64_t kInvalidCost = INT64_MAX;
constexpr int kMaxCandidates = 10;  // jv_elimination.cpp:28

struct Point {
  double x = 0, y = 0;
  std::int64_t ix = 0, iy = 0;
};

struct Edge {
  std::int32_t u = 0, v = 0;
  std::int64_t weight = 0;
  bool active = true;
};

}  // namespace

struct orc_graph {
  std::int32_t dimension = 0;
  int distance_type = 0;  // 0 EUC_2D, 1 CEIL_2D
  bool integer_coordinates = false;
  bool integer_distance_safe = false;
  std::vector<Point> points;
  std::vector<Edge> edges;
  std::vector<std::int32_t> row_offsets, neighbors, csr_edge_ids;
  std::vector<std::int64_t> csr_weights;
};

namespace {

// ---------------------------------------------------------------- distance
// Restoring integer square root: distance.cpp:9-27.
std::uint64_t IsqrtRestoring(std::uint64_t value) {
  std::uint64_t rest = value, root = 0, bit = std::uint64_t{1} << 62;
  while (bit > rest) bit >>= 2;
  for (; bit != 0; bit >>= 2) {
    if (rest >= root + bit) {
      rest -= root + bit;
      root = (root >> 1) + bit;
    } else {
      root >>= 1;
    }
  }
  return root;
}

// distance.cpp:29-68.
bool ExactDistance(std::int64_t ax, std::int64_t ay, std::int64_t bx, std::int64_t by, int type,
                   std::int64_t* out, std::string* error) {
  const i128 dx = static_cast<i128>(ax) - bx;
  const i128 dy = static_cast<i128>(ay) - by;
  const i128 s = dx * dx + dy * dy;
  if (s < 0 || s > static_cast<i128>(UINT64_MAX)) {
    if (error) *error = "squared distance exceeds uint64";
    return false;
  }
  const auto s64 = static_cast<std::uint64_t>(s);
  const std::uint64_t r = IsqrtRestoring(s64);
  std::uint64_t rounded = r;
  if (type == 0) {
    if (s64 - r * r > r) ++rounded;  // S > r^2 + r  <=>  sqrt(S) >= r + 1/2
  } else if (type == 1) {
    if (r * r != s64) ++rounded;
  } else {
    if (error) *error = "unsupported distance type";
    return false;
  }
  if (rounded > static_cast<std::uint64_t>(INT64_MAX)) {
    if (error) *error = "distance exceeds int64";
    return false;
  }
  *out = static_cast<std::int64_t>(rounded);
  return true;
}

// distance.cpp:70-86 (non-integer coordinates).
std::int64_t FloatingDistance(const Point& a, const Point& b, int type) {
  const double raw = std::hypot(a.x - b.x, a.y - b.y);
  if (!std::isfinite(raw) || raw > static_cast<double>(INT64_MAX)) {
    throw std::overflow_error("floating distance invalid or overflows");
  }
  if (type == 0) return static_cast<std::int64_t>(std::floor(raw + 0.5));
  if (type == 1) return static_cast<std::int64_t>(std::ceil(raw));
  throw std::invalid_argument("unsupported distance type");
}

// graph.cpp:300-316.
std::int64_t Distance(const orc_graph& g, std::int32_t u, std::int32_t v) {
  if (u < 0 || v < 0 || u >= g.dimension || v >= g.dimension) {
    throw std::out_of_range("distance query out of range");
  }
  const Point& a = g.points[static_cast<std::size_t>(u)];
  const Point& b = g.points[static_cast<std::size_t>(v)];
  if (g.integer_coordinates) {
    std::int64_t d = 0;
    std::string error;
    if (!ExactDistance(a.ix, a.iy, b.ix, b.iy, g.distance_type, &d, &error)) {
      throw std::overflow_error(error);
    }
    return d;
  }
  return FloatingDistance(a, b, g.distance_type);
}

// ---------------------------------------------------------------- snapshot
bool CanonicalSorted(const std::vector<Edge>& edges) {
  for (std::size_t i = 0; i < edges.size(); ++i) {
    if (edges[i].u >= edges[i].v) return false;
    if (i > 0 && std::tie(edges[i - 1].u, edges[i - 1].v) > std::tie(edges[i].u, edges[i].v)) {
      return false;
    }
  }
  return true;
}

// graph.cpp:230-282: rows are filled in edge-id order from both endpoints;
// a canonically sorted edge list yields ascending rows, any other order falls
// back to sorting each row by (neighbor, edge_id, weight).
void RebuildCsr(orc_graph* g) {
  const auto n = static_cast<std::size_t>(g->dimension);
  g->row_offsets.assign(n + 1, 0);
  for (const Edge& e : g->edges) {
    if (!e.active) continue;
    ++g->row_offsets[static_cast<std::size_t>(e.u) + 1];
    ++g->row_offsets[static_cast<std::size_t>(e.v) + 1];
  }
  for (std::size_t i = 1; i <= n; ++i) g->row_offsets[i] += g->row_offsets[i - 1];
  const auto adjacency = static_cast<std::size_t>(g->row_offsets[n]);
  g->neighbors.assign(adjacency, -1);
  g->csr_edge_ids.assign(adjacency, -1);
  g->csr_weights.assign(adjacency, 0);
  std::vector<std::int32_t> fill(g->row_offsets.begin(), g->row_offsets.end() - 1);
  for (std::size_t id = 0; id < g->edges.size(); ++id) {
    const Edge& e = g->edges[id];
    if (!e.active) continue;
    const std::int32_t ends[2][2] = {{e.u, e.v}, {e.v, e.u}};
    for (const auto& end : ends) {
      const auto slot = static_cast<std::size_t>(fill[static_cast<std::size_t>(end[0])]++);
      g->neighbors[slot] = end[1];
      g->csr_edge_ids[slot] = static_cast<std::int32_t>(id);
      g->csr_weights[slot] = e.weight;
    }
  }
  if (CanonicalSorted(g->edges)) return;
  for (std::size_t vtx = 0; vtx < n; ++vtx) {
    const auto lo = static_cast<std::size_t>(g->row_offsets[vtx]);
    const auto hi = static_cast<std::size_t>(g->row_offsets[vtx + 1]);
    std::vector<std::tuple<std::int32_t, std::int32_t, std::int64_t>> row;
    for (std::size_t i = lo; i < hi; ++i) {
      row.emplace_back(g->neighbors[i], g->csr_edge_ids[i], g->csr_weights[i]);
    }
    std::sort(row.begin(), row.end());
    for (std::size_t i = lo; i < hi; ++i) {
      std::tie(g->neighbors[i], g->csr_edge_ids[i], g->csr_weights[i]) = row[i - lo];
    }
  }
}

std::int32_t Degree(const orc_graph& g, std::int32_t vtx) {
  return g.row_offsets[static_cast<std::size_t>(vtx) + 1] -
         g.row_offsets[static_cast<std::size_t>(vtx)];
}

// graph.cpp:356-363: binary search in the (sorted) CSR row of u.
bool HasActiveEdge(const orc_graph& g, std::int32_t u, std::int32_t v) {
  if (u < 0 || v < 0 || u >= g.dimension || v >= g.dimension) return false;
  const auto first = g.neighbors.begin() + g.row_offsets[static_cast<std::size_t>(u)];
  const auto last = g.neighbors.begin() + g.row_offsets[static_cast<std::size_t>(u) + 1];
  return std::binary_search(first, last, v);
}

// graph.cpp:69-80, 323-346: FNV-1a 64 over the raw bytes of each member.
struct Fnv {
  std::uint64_t h = 14695981039346656037ULL;
  void Bytes(const void* p, std::size_t n) {
    const auto* b = static_cast<const unsigned char*>(p);
    for (std::size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ULL;
    }
  }
  template <typename T> void Value(T v) { Bytes(&v, sizeof(v)); }
};

std::uint64_t ContentHash(const orc_graph& g) {
  Fnv f;
  f.Value<std::int32_t>(g.dimension);
  f.Value<std::uint8_t>(static_cast<std::uint8_t>(g.distance_type));
  f.Value<bool>(g.integer_coordinates);
  for (const Point& p : g.points) {
    if (g.integer_coordinates) {
      f.Value<std::int64_t>(p.ix);
      f.Value<std::int64_t>(p.iy);
    } else {
      std::uint64_t bx, by;
      std::memcpy(&bx, &p.x, 8);
      std::memcpy(&by, &p.y, 8);
      f.Value(bx);
      f.Value(by);
    }
  }
  for (const Edge& e : g.edges) {
    f.Value<std::int32_t>(e.u);
    f.Value<std::int32_t>(e.v);
    f.Value<std::int64_t>(e.weight);
    f.Value<bool>(e.active);
  }
  return f.h;
}

std::string Hex16(std::uint64_t v) {
  std::ostringstream out;
  out << std::hex << std::setfill('0') << std::setw(16) << v;
  return out.str();
}

std::string Trim(const std::string& s) {
  const auto a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return {};
  const auto b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool Integral(double value, std::int64_t* out) {
  if (!std::isfinite(value) || value < static_cast<double>(INT64_MIN) ||
      value > static_cast<double>(INT64_MAX)) {
    return false;
  }
  double whole = 0;
  if (std::modf(value, &whole) != 0.0) return false;
  *out = static_cast<std::int64_t>(whole);
  return true;
}

void ComputeDistanceSafe(orc_graph* g) {
  g->integer_distance_safe = g->integer_coordinates;
  if (!g->integer_coordinates || g->points.empty()) return;
  std::int64_t lox = INT64_MAX, hix = INT64_MIN, loy = INT64_MAX, hiy = INT64_MIN;
  for (const Point& p : g->points) {
    lox = std::min(lox, p.ix);
    hix = std::max(hix, p.ix);
    loy = std::min(loy, p.iy);
    hiy = std::max(hiy, p.iy);
  }
  const i128 dx = static_cast<i128>(hix) - lox;
  const i128 dy = static_cast<i128>(hiy) - loy;
  g->integer_distance_safe = dx * dx + dy * dy <= static_cast<i128>(UINT64_MAX);
}

std::string ReadFile(const char* path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

// graph.cpp:98-228.
orc_graph* Load(const char* tsp_path, const char* edge_path) {
  auto g = std::make_unique<orc_graph>();
  g->integer_coordinates = true;
  std::istringstream tsp(ReadFile(tsp_path));
  std::string line;
  bool coords = false;
  std::vector<bool> seen;
  while (std::getline(tsp, line)) {
    line = Trim(line);
    if (line.empty()) continue;
    if (!coords) {
      if (line == "NODE_COORD_SECTION") {
        if (g->dimension <= 0) throw std::runtime_error("DIMENSION missing before coordinates");
        g->points.resize(static_cast<std::size_t>(g->dimension));
        seen.assign(static_cast<std::size_t>(g->dimension), false);
        coords = true;
        continue;
      }
      const auto colon = line.find(':');
      if (colon == std::string::npos) continue;
      const std::string key = Trim(line.substr(0, colon));
      const std::string value = Trim(line.substr(colon + 1));
      if (key == "DIMENSION") {
        g->dimension = std::stoi(value);
        if (g->dimension <= 0 || g->dimension == INT32_MAX) {
          throw std::runtime_error("DIMENSION out of range");
        }
      } else if (key == "EDGE_WEIGHT_TYPE") {
        if (value == "EUC_2D") {
          g->distance_type = 0;
        } else if (value == "CEIL_2D") {
          g->distance_type = 1;
        } else {
          throw std::runtime_error("unsupported EDGE_WEIGHT_TYPE " + value);
        }
      }
      continue;
    }
    if (line == "EOF") break;
    std::istringstream fields(line);
    std::int32_t id = 0;
    Point p;
    if (!(fields >> id >> p.x >> p.y)) throw std::runtime_error("bad coordinate line: " + line);
    if (id < 1 || id > g->dimension || seen[static_cast<std::size_t>(id - 1)]) {
      throw std::runtime_error("node id out of range or repeated");
    }
    const bool xi = Integral(p.x, &p.ix);
    const bool yi = Integral(p.y, &p.iy);
    g->integer_coordinates = g->integer_coordinates && xi && yi;
    g->points[static_cast<std::size_t>(id - 1)] = p;
    seen[static_cast<std::size_t>(id - 1)] = true;
  }
  if (!coords || std::find(seen.begin(), seen.end(), false) != seen.end()) {
    throw std::runtime_error("incomplete coordinates");
  }
  ComputeDistanceSafe(g.get());

  std::istringstream edg(ReadFile(edge_path));
  std::int32_t dim = 0;
  std::int64_t count = 0;
  if (!(edg >> dim >> count) || dim != g->dimension || count < 0 || count > INT32_MAX) {
    throw std::runtime_error("bad edge file header");
  }
  std::set<std::pair<std::int32_t, std::int32_t>> unique;
  for (std::int64_t i = 0; i < count; ++i) {
    Edge e;
    if (!(edg >> e.u >> e.v >> e.weight)) throw std::runtime_error("edge file ended early");
    if (e.u < 0 || e.v < 0 || e.u >= g->dimension || e.v >= g->dimension || e.u == e.v ||
        e.weight < 0) {
      throw std::runtime_error("edge out of range, self loop or negative weight");
    }
    if (e.u > e.v) std::swap(e.u, e.v);
    if (!unique.insert({e.u, e.v}).second) throw std::runtime_error("duplicate edge");
    if (e.weight != Distance(*g, e.u, e.v)) throw std::runtime_error("weight != distance");
    g->edges.push_back(e);
  }
  std::string extra;
  if (edg >> extra) throw std::runtime_error("extra data after declared edges");
  std::sort(g->edges.begin(), g->edges.end(), [](const Edge& a, const Edge& b) {
    return std::tie(a.u, a.v) < std::tie(b.u, b.v);
  });
  RebuildCsr(g.get());
  return g.release();
}

// ---------------------------------------------------------------------- JV
struct Scored {
  std::int64_t score;
  std::int32_t node;
};

// jv_elimination.cpp:45-84.  Side a (other endpoint b) is visited first, then
// side b, where a neighbour already adjacent to a is skipped; the ten
// smallest (score, node) pairs are kept in ascending order.  Because a node
// is scored at most once, keeping the ten smallest of all scores equals the
// reference's insert-sort-truncate loop.
std::vector<std::int32_t> CandidateNodes(const orc_graph& g, std::int32_t a, std::int32_t b,
                                         std::uint64_t* evaluations) {
  std::vector<Scored> scored;
  const auto visit = [&](std::int32_t from, std::int32_t other, bool second) {
    for (std::int32_t s = g.row_offsets[static_cast<std::size_t>(from)];
         s < g.row_offsets[static_cast<std::size_t>(from) + 1]; ++s) {
      const std::int32_t x = g.neighbors[static_cast<std::size_t>(s)];
      if (x == a || x == b) continue;
      if (second && HasActiveEdge(g, a, x)) continue;
      const i128 score = static_cast<i128>(g.csr_weights[static_cast<std::size_t>(s)]) +
                         Distance(g, x, other);
      ++*evaluations;
      if (score > static_cast<i128>(INT64_MAX)) continue;
      scored.push_back({static_cast<std::int64_t>(score), x});
    }
  };
  visit(a, b, false);
  visit(b, a, true);
  const auto less = [](const Scored& l, const Scored& r) {
    return std::tie(l.score, l.node) < std::tie(r.score, r.node);
  };
  const std::size_t keep = std::min<std::size_t>(scored.size(), kMaxCandidates);
  std::partial_sort(scored.begin(), scored.begin() + static_cast<std::ptrdiff_t>(keep),
                    scored.end(), less);
  std::vector<std::int32_t> out;
  for (std::size_t i = 0; i < keep; ++i) out.push_back(scored[i].node);
  return out;
}

// jv_elimination.cpp:86-117: c witnesses edge (a,b) when no neighbour d of c
// (other than a, b) satisfies w_ab + w_cd <= D(a,c) + D(d,b) or
// w_ab + w_cd <= D(a,d) + D(b,c).
bool IsWitness(const orc_graph& g, const Edge& e, std::int32_t c, std::uint64_t* evaluations) {
  if (c < 0 || c >= g.dimension || c == e.u || c == e.v) return false;
  const std::int64_t cab = e.weight;
  const std::int64_t cac = Distance(g, e.u, c);
  const std::int64_t cbc = Distance(g, e.v, c);
  *evaluations += 2;
  for (std::int32_t s = g.row_offsets[static_cast<std::size_t>(c)];
       s < g.row_offsets[static_cast<std::size_t>(c) + 1]; ++s) {
    const std::int32_t d = g.neighbors[static_cast<std::size_t>(s)];
    if (d == e.u || d == e.v) continue;
    const i128 left = static_cast<i128>(cab) + g.csr_weights[static_cast<std::size_t>(s)];
    const i128 first = static_cast<i128>(cac) + Distance(g, d, e.v);
    const i128 second = static_cast<i128>(Distance(g, e.u, d)) + cbc;
    *evaluations += 2;
    if (left <= first || left <= second) return false;
  }
  return true;
}

bool EdgeEligible(const orc_graph& g, const Edge& e) {
  return e.active && Degree(g, e.u) > 2 && Degree(g, e.v) > 2;
}

// jv_elimination.cpp:209-224.
std::int64_t FindCandidates(const orc_graph& g, std::int32_t* witnesses,
                            std::uint64_t* evaluations) {
  std::int64_t found = 0;
  for (std::size_t id = 0; id < g.edges.size(); ++id) {
    std::int32_t w = -1;
    const Edge& e = g.edges[id];
    if (EdgeEligible(g, e)) {
      for (std::int32_t c : CandidateNodes(g, e.u, e.v, evaluations)) {
        if (IsWitness(g, e, c, evaluations)) {
          w = c;
          break;
        }
      }
    }
    if (w >= 0) ++found;
    if (witnesses) witnesses[id] = w;
  }
  return found;
}

// jv_elimination.cpp:226-248.
bool Verify(const orc_graph& g, std::int32_t edge_id, std::int32_t witness) {
  if (edge_id < 0 || static_cast<std::size_t>(edge_id) >= g.edges.size()) return false;
  const Edge& e = g.edges[static_cast<std::size_t>(edge_id)];
  if (!EdgeEligible(g, e)) return false;
  std::uint64_t unused = 0;
  return IsWitness(g, e, witness, &unused);
}

struct Cand {
  std::int32_t edge_id;
  std::int32_t witness;
};

// elimination_commit.cpp:13-75: snapshot-hash check, sort by
// (u, v, method, witness, edge_id), one candidate per edge, then a greedy
// running degree gate (> 2 at both ends) before deactivating the edges and
// rebuilding the CSR.
std::vector<Cand> Commit(orc_graph* g, std::vector<Cand> cands, std::uint64_t expected_hash) {
  if (ContentHash(*g) != expected_hash) throw std::runtime_error("snapshot hash changed");
  for (const Cand& c : cands) {
    if (c.edge_id < 0 || static_cast<std::size_t>(c.edge_id) >= g->edges.size()) {
      throw std::runtime_error("candidate edge id out of range");
    }
    if (!g->edges[static_cast<std::size_t>(c.edge_id)].active) {
      throw std::runtime_error("candidate is not an active edge");
    }
  }
  const auto key = [&](const Cand& c) {
    const Edge& e = g->edges[static_cast<std::size_t>(c.edge_id)];
    return std::make_tuple(e.u, e.v, 1 /* JV */, c.witness, c.edge_id);
  };
  std::sort(cands.begin(), cands.end(), [&](const Cand& l, const Cand& r) { return key(l) < key(r); });
  cands.erase(std::unique(cands.begin(), cands.end(),
                          [](const Cand& l, const Cand& r) { return l.edge_id == r.edge_id; }),
              cands.end());
  std::vector<std::int32_t> degree(static_cast<std::size_t>(g->dimension));
  for (std::int32_t v = 0; v < g->dimension; ++v) degree[static_cast<std::size_t>(v)] = Degree(*g, v);
  std::vector<Cand> committed;
  for (const Cand& c : cands) {
    const Edge& e = g->edges[static_cast<std::size_t>(c.edge_id)];
    auto& du = degree[static_cast<std::size_t>(e.u)];
    auto& dv = degree[static_cast<std::size_t>(e.v)];
    if (du <= 2 || dv <= 2) continue;
    --du;
    --dv;
    committed.push_back(c);
  }
  if (!committed.empty()) {
    for (const Cand& c : committed) g->edges[static_cast<std::size_t>(c.edge_id)].active = false;
    RebuildCsr(g);
  }
  return committed;
}

double Ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Record {
  std::uint32_t epoch;
  std::uint64_t hash;
  std::int32_t edge_id, u, v, witness;
};
struct Metrics {
  std::uint32_t epoch;
  std::size_t edges_before, proposed, verified, rejected, committed;
  double propose_ms, verify_ms;
};

std::size_t ActiveCount(const orc_graph& g) {
  return static_cast<std::size_t>(
      std::count_if(g.edges.begin(), g.edges.end(), [](const Edge& e) { return e.active; }));
}

// jv_elimination.cpp:250-368 (CPU backend; V1 proof text).
std::string RunJv(orc_graph* g, std::uint32_t max_rounds, const std::string& backend) {
  if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
  const std::uint64_t initial = ContentHash(*g);
  std::vector<Record> records;
  std::vector<Metrics> metrics;
  for (std::uint32_t epoch = 0; epoch < max_rounds; ++epoch) {
    Metrics m{};
    m.epoch = epoch;
    m.edges_before = ActiveCount(*g);
    const std::uint64_t snap = ContentHash(*g);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::int32_t> w(g->edges.size());
    std::uint64_t evals = 0;
    FindCandidates(*g, w.data(), &evals);
    std::vector<Cand> proposed;
    for (std::size_t id = 0; id < w.size(); ++id) {
      if (w[id] >= 0) proposed.push_back({static_cast<std::int32_t>(id), w[id]});
    }
    m.propose_ms = Ms(t0);
    m.proposed = proposed.size();
    t0 = std::chrono::steady_clock::now();
    std::vector<Cand> verified;
    for (const Cand& c : proposed) {
      if (Verify(*g, c.edge_id, c.witness)) {
        verified.push_back(c);
      } else {
        ++m.rejected;
      }
    }
    m.verify_ms = Ms(t0);
    m.verified = verified.size();
    const std::vector<Cand> committed = Commit(g, verified, snap);
    m.committed = committed.size();
    for (const Cand& c : committed) {
      const Edge& e = g->edges[static_cast<std::size_t>(c.edge_id)];
      records.push_back({epoch, snap, c.edge_id, e.u, e.v, c.witness});
    }
    metrics.push_back(m);
    if (committed.empty()) break;
  }
  std::ostringstream out;
  out << "CUDAEE_PROOF_V1\n";
  out << "backend " << backend << '\n';
  out << "initial_hash " << Hex16(initial) << '\n';
  for (const Record& r : records) {
    out << "record " << r.epoch << ' ' << Hex16(r.hash) << ' ' << r.edge_id << ' ' << r.u << ' '
        << r.v << " JV " << r.witness << '\n';
  }
  for (const Metrics& m : metrics) {
    out << "metrics " << m.epoch << ' ' << m.edges_before << ' ' << m.proposed << ' '
        << m.verified << ' ' << m.rejected << ' ' << m.committed << ' ' << m.propose_ms << ' '
        << m.verify_ms << '\n';
  }
  out << "final_hash " << Hex16(ContentHash(*g)) << '\n';
  out << "END\n";
  return out.str();
}

std::uint64_t ParseHex16(const std::string& token) {
  if (token.size() != 16 ||
      token.find_first_not_of("0123456789abcdefABCDEF") != std::string::npos) {
    throw std::runtime_error("bad hash token");
  }
  return std::stoull(token, nullptr, 16);
}

// jv_elimination.cpp:370-523 (V1 subset) and :525-605.
void Replay(orc_graph* g, const std::string& text) {
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line) || line != "CUDAEE_PROOF_V1") throw std::runtime_error("bad magic");
  bool saw_backend = false, saw_initial = false, saw_final = false, saw_end = false;
  std::uint64_t initial = 0, final_hash = 0;
  std::vector<Record> records;
  while (std::getline(in, line)) {
    std::istringstream f(line);
    std::string kind;
    f >> kind;
    std::string extra;
    if (kind == "backend") {
      std::string name;
      if (saw_backend || !(f >> name)) throw std::runtime_error("bad backend line");
      saw_backend = true;
    } else if (kind == "initial_hash") {
      std::string h;
      if (saw_initial || !(f >> h)) throw std::runtime_error("bad initial_hash line");
      initial = ParseHex16(h);
      saw_initial = true;
    } else if (kind == "record") {
      Record r{};
      std::string h, method;
      if (!(f >> r.epoch >> h >> r.edge_id >> r.u >> r.v >> method >> r.witness)) {
        throw std::runtime_error("bad record line");
      }
      if (method != "JV") throw std::runtime_error("V1 proof supports JV records only");
      r.hash = ParseHex16(h);
      records.push_back(r);
    } else if (kind == "metrics") {
      Metrics m{};
      if (!(f >> m.epoch >> m.edges_before >> m.proposed >> m.verified >> m.rejected >>
            m.committed >> m.propose_ms >> m.verify_ms) ||
          !std::isfinite(m.propose_ms) || !std::isfinite(m.verify_ms) || m.propose_ms < 0 ||
          m.verify_ms < 0) {
        throw std::runtime_error("bad metrics line");
      }
    } else if (kind == "final_hash") {
      std::string h;
      if (saw_final || !(f >> h)) throw std::runtime_error("bad final_hash line");
      final_hash = ParseHex16(h);
      saw_final = true;
    } else if (kind == "END") {
      saw_end = true;
      break;
    } else if (!kind.empty()) {
      throw std::runtime_error("unknown proof line " + kind);
    }
    if (f >> extra) throw std::runtime_error("extra field in proof line");
  }
  std::string trailing;
  if (in >> trailing) throw std::runtime_error("data after END");
  if (!saw_backend || !saw_initial || !saw_final || !saw_end) {
    throw std::runtime_error("incomplete proof");
  }
  if (ContentHash(*g) != initial) throw std::runtime_error("initial hash mismatch");
  orc_graph work = *g;
  std::size_t cursor = 0;
  std::uint32_t expected_epoch = 0;
  while (cursor < records.size()) {
    const std::uint32_t epoch = records[cursor].epoch;
    if (epoch != expected_epoch) throw std::runtime_error("epochs not consecutive");
    const std::uint64_t snap = ContentHash(work);
    std::vector<Cand> cands;
    std::vector<const Record*> group;
    while (cursor < records.size() && records[cursor].epoch == epoch) {
      const Record& r = records[cursor++];
      if (r.hash != snap || r.edge_id < 0 ||
          static_cast<std::size_t>(r.edge_id) >= work.edges.size()) {
        throw std::runtime_error("record hash or edge id mismatch");
      }
      const Edge& e = work.edges[static_cast<std::size_t>(r.edge_id)];
      if (e.u != r.u || e.v != r.v) throw std::runtime_error("record endpoints mismatch");
      if (!Verify(work, r.edge_id, r.witness)) throw std::runtime_error("JV record fails verify");
      cands.push_back({r.edge_id, r.witness});
      group.push_back(&r);
    }
    const std::vector<Cand> committed = Commit(&work, cands, snap);
    if (committed.size() != group.size()) throw std::runtime_error("degree gate violated");
    for (std::size_t i = 0; i < committed.size(); ++i) {
      if (committed[i].edge_id != group[i]->edge_id || committed[i].witness != group[i]->witness) {
        throw std::runtime_error("records not in deterministic commit order");
      }
    }
    ++expected_epoch;
  }
  if (ContentHash(work) != final_hash) throw std::runtime_error("final hash mismatch");
  *g = std::move(work);
}

// ------------------------------------------------------------------- k-opt
constexpr std::uint8_t kUnmatched = 0xff;
constexpr int kMaxEndpoints = 14;

struct Matching {
  std::uint8_t count;
  std::array<std::uint8_t, kMaxEndpoints> mate;
};

Matching Empty(std::uint32_t paths) {
  Matching m{};
  m.count = static_cast<std::uint8_t>(2 * paths);
  m.mate.fill(kUnmatched);
  return m;
}

// path_system.cpp:47-71: pair the lowest unmatched endpoint with every later
// unmatched endpoint in ascending order, depth first.
void InsideRecursive(Matching* m, std::vector<Matching>* out) {
  std::uint32_t first = m->count;
  for (std::uint32_t i = 0; i < m->count; ++i) {
    if (m->mate[i] == kUnmatched) {
      first = i;
      break;
    }
  }
  if (first == m->count) {
    out->push_back(*m);
    return;
  }
  for (std::uint32_t second = first + 1; second < m->count; ++second) {
    if (m->mate[second] != kUnmatched) continue;
    m->mate[first] = static_cast<std::uint8_t>(second);
    m->mate[second] = static_cast<std::uint8_t>(first);
    InsideRecursive(m, out);
    m->mate[first] = m->mate[second] = kUnmatched;
  }
}

// path_system.cpp:389-402.
bool Perfect(const Matching& m, std::uint32_t paths) {
  if (paths == 0 || paths > 7 || m.count != 2 * paths) return false;
  for (std::uint32_t i = 0; i < m.count; ++i) {
    const std::uint8_t p = m.mate[i];
    if (p >= m.count || p == i || m.mate[p] != i) return false;
  }
  return true;
}

// path_system.cpp:404-421: walk inside, outside, inside, ... from endpoint 0;
// one cycle through every endpoint.
bool AlternatingHamiltonian(const Matching& outside, const Matching& inside, std::uint32_t paths) {
  if (!Perfect(outside, paths) || !Perfect(inside, paths)) return false;
  std::array<bool, kMaxEndpoints> seen{};
  std::uint32_t at = 0, steps = 0;
  while (!seen[at]) {
    seen[at] = true;
    at = (steps % 2 == 0) ? inside.mate[at] : outside.mate[at];
    ++steps;
  }
  return steps == 2 * paths && at == 0;
}

void HashByte(std::uint64_t* h, std::uint8_t b) {
  *h ^= b;
  *h *= 1099511628211ULL;
}
void HashU32(std::uint64_t* h, std::uint32_t v) {
  for (int s = 0; s < 32; s += 8) HashByte(h, static_cast<std::uint8_t>((v >> s) & 0xff));
}

// kopt_search.cpp:1143-1193.
std::vector<Matching> Templates(std::uint32_t k, std::uint64_t* hash) {
  if (k < 3 || k > 5) throw std::invalid_argument("k must be in [3,5]");
  Matching components = Empty(k);
  const auto pair = [&](std::uint32_t a, std::uint32_t b) {
    components.mate[a] = static_cast<std::uint8_t>(b);
    components.mate[b] = static_cast<std::uint8_t>(a);
  };
  for (std::uint32_t e = 0; e + 1 < k; ++e) pair(2 * e + 1, 2 * e + 2);
  pair(2 * k - 1, 0);
  std::vector<Matching> all;
  Matching scratch = Empty(k);
  InsideRecursive(&scratch, &all);
  std::vector<Matching> out;
  for (const Matching& c : all) {
    if (!AlternatingHamiltonian(components, c, k)) continue;
    bool reuses = false;
    for (std::uint32_t e = 0; e < k; ++e) {
      if (c.mate[2 * e] == 2 * e + 1) reuses = true;
    }
    if (!reuses) out.push_back(c);
  }
  const std::size_t expected[3] = {4, 25, 208};
  if (out.size() != expected[k - 3]) throw std::logic_error("template count mismatch");
  std::uint64_t h = 14695981039346656037ULL;
  for (char ch : std::string("CUDAEE_KOPT_RECONNECT_TEMPLATES_V1")) {
    HashByte(&h, static_cast<std::uint8_t>(ch));
  }
  HashU32(&h, k);
  HashU32(&h, static_cast<std::uint32_t>(out.size()));
  for (const Matching& m : out) {
    for (std::uint32_t i = 0; i < m.count; ++i) HashByte(&h, m.mate[i]);
  }
  *hash = h;
  return out;
}

struct PairKey {
  std::int32_t u, v;
  bool operator==(const PairKey& o) const { return u == o.u && v == o.v; }
};
PairKey Canon(std::int32_t a, std::int32_t b) { return a < b ? PairKey{a, b} : PairKey{b, a}; }

struct Coords {
  const std::int64_t* x;
  const std::int64_t* y;
  std::int32_t n;
  int type;
  std::int64_t Dist(std::int32_t a, std::int32_t b) const {
    std::int64_t d = 0;
    std::string err;
    if (!ExactDistance(x[a], y[a], x[b], y[b], type, &d, &err)) throw std::overflow_error(err);
    return d;
  }
};

// kopt_search.cpp:784-807.
void ValidateTask(std::int32_t n, std::uint32_t k, const orc_kopt_task& t) {
  if (t.deleted_cost < 0) throw std::invalid_argument("k-opt task has negative deleted_cost");
  std::vector<PairKey> deleted;
  for (std::uint32_t e = 0; e < k; ++e) {
    const std::int32_t a = t.port_nodes[2 * e], b = t.port_nodes[2 * e + 1];
    if (a < 0 || b < 0 || a >= n || b >= n || a == b) {
      throw std::invalid_argument("k-opt task has an invalid or repeated deleted edge");
    }
    const PairKey c = Canon(a, b);
    if (std::find(deleted.begin(), deleted.end(), c) != deleted.end()) {
      throw std::invalid_argument("k-opt task has an invalid or repeated deleted edge");
    }
    deleted.push_back(c);
  }
}

// kopt_search.cpp:809-841: template -> ordered list of k port-pair indices
// (port * 10 + partner for port < partner).
std::vector<std::array<std::uint8_t, 5>> Plans(std::uint32_t k, const std::vector<Matching>& tpl) {
  std::vector<std::array<std::uint8_t, 5>> plans;
  for (const Matching& m : tpl) {
    std::array<std::uint8_t, 5> plan{};
    std::uint32_t n = 0;
    for (std::uint32_t p = 0; p < m.count; ++p) {
      const std::uint32_t q = m.mate[p];
      if (p >= q) continue;
      plan[n++] = static_cast<std::uint8_t>(p * 10 + q);
    }
    if (n != k) throw std::logic_error("template pair count mismatch");
    plans.push_back(plan);
  }
  return plans;
}

// kopt_search.cpp:865-960: an optional dense distance table over the distinct
// port nodes of a batch, built only for >= 8192 cells, <= 512 distinct nodes
// and when it at least halves the distance evaluations.  It changes timing,
// never values.
struct BatchCache {
  std::vector<std::int32_t> local;  // node -> local index
  std::vector<std::int64_t> table;
  std::size_t count = 0;
};

bool BuildBatchCache(const Coords& c, std::uint32_t k, const orc_kopt_task* tasks, std::size_t rows,
                     const std::vector<std::array<std::uint8_t, 5>>& plans, std::size_t cells,
                     BatchCache* cache) {
  if (cells < 8192 || rows == 0 || static_cast<std::size_t>(c.n) > 1048576) return false;
  std::vector<std::int32_t> nodes;
  cache->local.assign(static_cast<std::size_t>(c.n), -1);
  for (std::size_t r = 0; r < rows; ++r) {
    for (std::uint32_t p = 0; p < 2 * k; ++p) {
      const std::int32_t node = tasks[r].port_nodes[p];
      auto& slot = cache->local[static_cast<std::size_t>(node)];
      if (slot >= 0) continue;
      if (nodes.size() >= 512) return false;
      slot = static_cast<std::int32_t>(nodes.size());
      nodes.push_back(node);
    }
  }
  if (nodes.empty()) return false;
  std::array<bool, 100> used{};
  std::size_t used_count = 0;
  for (const auto& plan : plans) {
    for (std::uint32_t i = 0; i < k; ++i) {
      if (!used[plan[i]]) {
        used[plan[i]] = true;
        ++used_count;
      }
    }
  }
  const std::size_t full = nodes.size() * (nodes.size() - 1) / 2;
  if (used_count == 0 || full > rows * used_count / 2) return false;
  cache->count = nodes.size();
  cache->table.assign(cache->count * cache->count, 0);
  for (std::size_t i = 0; i < cache->count; ++i) {
    for (std::size_t j = i + 1; j < cache->count; ++j) {
      const std::int64_t d = c.Dist(nodes[i], nodes[j]);
      cache->table[i * cache->count + j] = cache->table[j * cache->count + i] = d;
    }
  }
  return true;
}

// kopt_search.cpp:962-1042: lazily resolved pair costs per row; a pair that
// is a self loop or re-adds a deleted edge is invalid; a template is invalid
// when any pair is invalid, two of its added edges coincide, or the sum
// overflows.
void ScoreRow(const Coords& c, std::uint32_t k, const orc_kopt_task& task,
              const std::vector<std::array<std::uint8_t, 5>>& plans, const BatchCache* cache,
              std::int64_t* out) {
  struct Pair {
    PairKey edge{0, 0};
    std::int64_t cost = -1;
  };
  std::array<Pair, 100> pairs{};
  std::array<PairKey, 5> deleted{};
  for (std::uint32_t e = 0; e < k; ++e) deleted[e] = Canon(task.port_nodes[2 * e], task.port_nodes[2 * e + 1]);
  const auto resolve = [&](std::uint8_t flat) -> const Pair& {
    Pair& p = pairs[flat];
    if (p.cost != -1) return p;
    const int a = flat / 10, b = flat % 10;
    p.edge = Canon(task.port_nodes[a], task.port_nodes[b]);
    if (p.edge.u == p.edge.v ||
        std::find(deleted.begin(), deleted.begin() + k, p.edge) != deleted.begin() + k) {
      p.cost = kInvalidCost;
      return p;
    }
    p.cost = cache ? cache->table[static_cast<std::size_t>(cache->local[static_cast<std::size_t>(task.port_nodes[a])]) * cache->count +
                                  static_cast<std::size_t>(cache->local[static_cast<std::size_t>(task.port_nodes[b])])]
                   : c.Dist(p.edge.u, p.edge.v);
    if (p.cost < 0) p.cost = kInvalidCost;
    return p;
  };
  for (std::size_t t = 0; t < plans.size(); ++t) {
    std::array<PairKey, 5> added{};
    std::uint32_t count = 0;
    std::int64_t total = 0;
    bool valid = true;
    for (std::uint32_t i = 0; i < k && valid; ++i) {
      const Pair& p = resolve(plans[t][i]);
      if (p.cost == kInvalidCost ||
          std::find(added.begin(), added.begin() + count, p.edge) != added.begin() + count ||
          total > INT64_MAX - p.cost) {
        valid = false;
        break;
      }
      added[count++] = p.edge;
      total += p.cost;
    }
    out[t] = (valid && count == k) ? total : kInvalidCost;
  }
}

}  // namespace

// =================================================================== C ABI
namespace {

int Guard(const std::function<void()>& body) {
  try {
    body();
    return 0;
  } catch (const std::exception& e) {
    return Fail(e.what());
  }
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

uint64_t orc_isqrt_floor(uint64_t value) { return IsqrtRestoring(value); }

int orc_exact_distance(int64_t ax, int64_t ay, int64_t bx, int64_t by, int type, int64_t* out) {
  std::string err;
  if (!ExactDistance(ax, ay, bx, by, type, out, &err)) return Fail(err);
  return 0;
}

orc_graph* orc_graph_load(const char* tsp_path, const char* edge_path) {
  orc_graph* g = nullptr;
  if (Guard([&] { g = Load(tsp_path, edge_path); }) != 0) return nullptr;
  return g;
}

orc_graph* orc_graph_from_arrays(int32_t dimension, int distance_type, const int64_t* x,
                                 const int64_t* y, int32_t edge_count, const int32_t* edge_u,
                                 const int32_t* edge_v, const int64_t* edge_weight,
                                 const uint8_t* edge_active) {
  orc_graph* out = nullptr;
  Guard([&] {
    if (dimension <= 0 || edge_count < 0 || (distance_type != 0 && distance_type != 1)) {
      throw std::invalid_argument("bad graph dimensions");
    }
    auto g = std::make_unique<orc_graph>();
    g->dimension = dimension;
    g->distance_type = distance_type;
    g->integer_coordinates = true;
    g->points.resize(static_cast<std::size_t>(dimension));
    for (std::int32_t i = 0; i < dimension; ++i) {
      Point& p = g->points[static_cast<std::size_t>(i)];
      p.ix = x[i];
      p.iy = y[i];
      p.x = static_cast<double>(x[i]);
      p.y = static_cast<double>(y[i]);
    }
    ComputeDistanceSafe(g.get());
    for (std::int32_t i = 0; i < edge_count; ++i) {
      if (edge_u[i] < 0 || edge_v[i] < 0 || edge_u[i] >= dimension || edge_v[i] >= dimension) {
        throw std::invalid_argument("edge endpoint out of range");
      }
      g->edges.push_back({edge_u[i], edge_v[i], edge_weight[i], edge_active ? edge_active[i] != 0 : true});
    }
    RebuildCsr(g.get());
    out = g.release();
  });
  return out;
}

orc_graph* orc_graph_clone(const orc_graph* graph) { return new orc_graph(*graph); }
void orc_graph_free(orc_graph* graph) { delete graph; }

int32_t orc_graph_dimension(const orc_graph* g) { return g->dimension; }
int32_t orc_graph_edge_count(const orc_graph* g) { return static_cast<int32_t>(g->edges.size()); }
int64_t orc_graph_adjacency_count(const orc_graph* g) { return static_cast<int64_t>(g->neighbors.size()); }
int32_t orc_graph_active_edge_count(const orc_graph* g) { return static_cast<int32_t>(ActiveCount(*g)); }

void orc_graph_flags(const orc_graph* g, int* distance_type, int* integer_coordinates,
                     int* integer_distance_safe) {
  if (distance_type) *distance_type = g->distance_type;
  if (integer_coordinates) *integer_coordinates = g->integer_coordinates ? 1 : 0;
  if (integer_distance_safe) *integer_distance_safe = g->integer_distance_safe ? 1 : 0;
}

void orc_graph_export(const orc_graph* g, int64_t* x, int64_t* y, int32_t* edge_u,
                      int32_t* edge_v, int64_t* edge_weight, uint8_t* edge_active,
                      int32_t* row_offsets, int32_t* neighbors, int32_t* csr_edge_ids,
                      int64_t* csr_weights) {
  for (std::size_t i = 0; i < g->points.size(); ++i) {
    if (x) x[i] = g->points[i].ix;
    if (y) y[i] = g->points[i].iy;
  }
  for (std::size_t i = 0; i < g->edges.size(); ++i) {
    if (edge_u) edge_u[i] = g->edges[i].u;
    if (edge_v) edge_v[i] = g->edges[i].v;
    if (edge_weight) edge_weight[i] = g->edges[i].weight;
    if (edge_active) edge_active[i] = g->edges[i].active ? 1 : 0;
  }
  if (row_offsets) std::copy(g->row_offsets.begin(), g->row_offsets.end(), row_offsets);
  if (neighbors) std::copy(g->neighbors.begin(), g->neighbors.end(), neighbors);
  if (csr_edge_ids) std::copy(g->csr_edge_ids.begin(), g->csr_edge_ids.end(), csr_edge_ids);
  if (csr_weights) std::copy(g->csr_weights.begin(), g->csr_weights.end(), csr_weights);
}

int orc_graph_set_active(orc_graph* g, int32_t edge_id, int active) {
  if (edge_id < 0 || static_cast<std::size_t>(edge_id) >= g->edges.size()) {
    return Fail("edge id out of range");
  }
  g->edges[static_cast<std::size_t>(edge_id)].active = active != 0;
  return 0;
}

void orc_graph_rebuild_csr(orc_graph* g) { RebuildCsr(g); }
uint64_t orc_graph_hash(const orc_graph* g) { return ContentHash(*g); }
int orc_graph_has_active_edge(const orc_graph* g, int32_t u, int32_t v) {
  return HasActiveEdge(*g, u, v) ? 1 : 0;
}
int orc_graph_distance(const orc_graph* g, int32_t u, int32_t v, int64_t* out) {
  return Guard([&] { *out = Distance(*g, u, v); });
}

int64_t orc_jv_find_candidates(const orc_graph* g, int32_t* witnesses,
                               uint64_t* distance_evaluations) {
  int64_t found = -1;
  std::uint64_t evals = 0;
  if (Guard([&] { found = FindCandidates(*g, witnesses, &evals); }) != 0) return -1;
  if (distance_evaluations) *distance_evaluations = evals;
  return found;
}

int orc_jv_verify(const orc_graph* g, int32_t edge_id, int32_t witness) {
  int ok = 0;
  if (Guard([&] { ok = Verify(*g, edge_id, witness) ? 1 : 0; }) != 0) return -1;
  return ok;
}

int64_t orc_jv_commit(orc_graph* g, int64_t count, const int32_t* edge_ids,
                      const int32_t* witnesses, uint64_t expected_hash,
                      int32_t* committed_edge_ids, int32_t* committed_witnesses) {
  int64_t n = -1;
  Guard([&] {
    std::vector<Cand> cands;
    for (int64_t i = 0; i < count; ++i) cands.push_back({edge_ids[i], witnesses[i]});
    const std::vector<Cand> done = Commit(g, cands, expected_hash);
    for (std::size_t i = 0; i < done.size(); ++i) {
      if (committed_edge_ids) committed_edge_ids[i] = done[i].edge_id;
      if (committed_witnesses) committed_witnesses[i] = done[i].witness;
    }
    n = static_cast<int64_t>(done.size());
  });
  return n;
}

int64_t orc_jv_run(orc_graph* g, uint32_t max_rounds, const char* backend_name, char* proof_out,
                   size_t capacity) {
  std::string proof;
  if (Guard([&] { proof = RunJv(g, max_rounds, backend_name ? backend_name : "cpu"); }) != 0) {
    return -1;
  }
  if (proof_out && capacity > 0) {
    const std::size_t n = std::min(proof.size(), capacity - 1);
    std::memcpy(proof_out, proof.data(), n);
    proof_out[n] = '\0';
  }
  return static_cast<int64_t>(proof.size());
}

int orc_jv_replay(orc_graph* g, const char* proof_text) {
  return Guard([&] { Replay(g, proof_text); });
}

int orc_kopt_templates(uint32_t k, uint8_t* mates, uint32_t capacity, uint32_t* count,
                       uint64_t* generator_hash) {
  return Guard([&] {
    std::uint64_t h = 0;
    const std::vector<Matching> t = Templates(k, &h);
    if (count) *count = static_cast<uint32_t>(t.size());
    if (generator_hash) *generator_hash = h;
    if (mates) {
      if (capacity < t.size()) throw std::invalid_argument("template capacity too small");
      for (std::size_t i = 0; i < t.size(); ++i) {
        std::memcpy(mates + i * kMaxEndpoints, t[i].mate.data(), kMaxEndpoints);
      }
    }
  });
}

int orc_kopt_costs(const int64_t* x, const int64_t* y, int32_t dimension, int distance_type,
                   uint32_t k, const orc_kopt_task* tasks, size_t task_count, int64_t* costs,
                   int max_threads, uint32_t* threads_used, uint32_t* distance_cache_nodes) {
  return Guard([&] {
    // kopt_search.cpp:1279-1302.
    if (k < 3 || k > 5) throw std::invalid_argument("k must be in [3,5]");
    if (dimension <= 0) throw std::invalid_argument("empty graph");
    orc_graph probe;
    probe.dimension = dimension;
    probe.integer_coordinates = true;
    probe.points.resize(static_cast<std::size_t>(dimension));
    for (std::int32_t i = 0; i < dimension; ++i) {
      probe.points[static_cast<std::size_t>(i)].ix = x[i];
      probe.points[static_cast<std::size_t>(i)].iy = y[i];
    }
    ComputeDistanceSafe(&probe);
    if (!probe.integer_distance_safe) {
      throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
    }
    for (std::size_t r = 0; r < task_count; ++r) ValidateTask(dimension, k, tasks[r]);
    std::uint64_t h = 0;
    const std::vector<Matching> tpl = Templates(k, &h);
    const auto plans = Plans(k, tpl);
    const Coords c{x, y, dimension, distance_type};
    const std::size_t cells = task_count * plans.size();
    BatchCache cache;
    const bool cached = BuildBatchCache(c, k, tasks, task_count, plans, cells, &cache);
    if (distance_cache_nodes) *distance_cache_nodes = cached ? static_cast<uint32_t>(cache.count) : 0;
    std::uint32_t used = 1;
#ifdef _OPENMP
    // kopt_search.cpp:1069-1086: OpenMP static schedule over rows, team size
    // min(cap, omp_get_max_threads(), omp_get_num_procs()), only >= 8192 cells.
    int limit = std::min(omp_get_max_threads(), omp_get_num_procs());
    if (max_threads > 0) limit = std::min(limit, max_threads);
    if (limit > 1 && cells >= 8192) {
#pragma omp parallel num_threads(limit)
      {
#pragma omp single
        used = static_cast<std::uint32_t>(omp_get_num_threads());
#pragma omp for schedule(static)
        for (std::size_t r = 0; r < task_count; ++r) {
          ScoreRow(c, k, tasks[r], plans, cached ? &cache : nullptr, costs + r * plans.size());
        }
      }
      if (threads_used) *threads_used = used;
      return;
    }
#endif
    for (std::size_t r = 0; r < task_count; ++r) {
      ScoreRow(c, k, tasks[r], plans, cached ? &cache : nullptr, costs + r * plans.size());
    }
    if (threads_used) *threads_used = used;
  });
}


int64_t orc_kopt_cursor_tasks(const int32_t* tour, const int64_t* edge_costs, uint32_t tour_length,
                              const uint64_t* selectable, uint32_t selectable_count, uint32_t k,
                              uint64_t first, uint64_t count, orc_kopt_task* tasks,
                              uint64_t* positions) {
  int64_t written = 0;
  const int status = Guard([&] {
    // PrepareNextK (kopt_search.cpp:2258-2271): k in [3, max_k] and k < |selectable|
    if (k < 3 || k > 5 || k >= selectable_count) return;
    std::vector<std::size_t> combination(k);
    for (std::size_t i = 0; i < k; ++i) combination[i] = i;
    // AdvanceCombination (kopt_search.cpp:736-752), restated: the last index
    // that can still move moves up by one and the ones after it follow; index 0
    // never moves.
    const auto advance = [&]() -> bool {
      std::size_t pos = combination.size() - 1;
      while (pos >= 1 && combination[pos] == selectable_count - (combination.size() - pos)) {
        if (pos == 1) return false;
        --pos;
      }
      ++combination[pos];
      for (std::size_t next = pos + 1; next < combination.size(); ++next) {
        combination[next] = combination[next - 1] + 1;
      }
      return true;
    };
    uint64_t index = 0;
    bool more = true;
    while (more && static_cast<uint64_t>(written) < count) {
      if (index >= first) {
        std::size_t deleted[5];
        for (uint32_t e = 0; e < k; ++e) deleted[e] = selectable[combination[e]];
        std::sort(deleted, deleted + k);
        // BuildKOptCostTask (kopt_search.cpp:561-584)
        orc_kopt_task t{};
        __int128 cost = 0;
        for (uint32_t e = 0; e < k; ++e) {
          const std::size_t at = deleted[e];
          if (at >= tour_length || edge_costs[at] < 0) {
            throw std::runtime_error("deletion position or cached edge cost is invalid");
          }
          t.port_nodes[2 * e] = tour[at];
          t.port_nodes[2 * e + 1] = tour[(at + 1) % tour_length];
          cost += edge_costs[at];
        }
        if (cost > std::numeric_limits<int64_t>::max()) {
          throw std::runtime_error("k-opt deleted cost overflows int64_t");
        }
        t.deleted_cost = static_cast<int64_t>(cost);
        tasks[written] = t;
        if (positions) {
          for (uint32_t e = 0; e < k; ++e) positions[written * k + e] = deleted[e];
        }
        ++written;
      }
      ++index;
      more = advance();
    }
  });
  return status == 0 ? written : status;
}


int64_t orc_ht_hamilton_replies(const orc_graph* graph, int32_t target_u, int32_t target_v,
                                const int32_t* centers, size_t count, uint64_t* offsets,
                                int32_t* replies, uint64_t capacity) {
  uint64_t total = 0;
  const int status = Guard([&] {
    const orc_graph& g = *graph;
    const int32_t tu = std::min(target_u, target_v), tv = std::max(target_u, target_v);
    if (tu < 0 || tv >= g.dimension || tu == tv || !HasActiveEdge(g, tu, tv)) {
      throw std::invalid_argument("HT reply target is not an active edge");
    }
    for (size_t i = 0; i < count; ++i) {
      if (centers[i] < 0 || centers[i] >= g.dimension || centers[i] == tu || centers[i] == tv) {
        throw std::invalid_argument("HT reply center is invalid");
      }
    }
    // 2-opt survival of edge (a, b) against the target (hamilton_tutte.cpp:79-87)
    const auto survives = [&](int32_t a, int32_t b) {
      const int32_t lo = std::min(a, b), hi = std::max(a, b);
      const __int128 original = static_cast<__int128>(Distance(g, tu, tv)) + Distance(g, lo, hi);
      const __int128 first = static_cast<__int128>(Distance(g, tu, hi)) + Distance(g, tv, lo);
      const __int128 second = static_cast<__int128>(Distance(g, tu, lo)) + Distance(g, tv, hi);
      return first >= original || second >= original;
    };
    offsets[0] = 0;
    for (size_t i = 0; i < count; ++i) {
      const int32_t c = centers[i];
      const int32_t begin = g.row_offsets[c], end = g.row_offsets[c + 1];
      const __int128 center_target = static_cast<__int128>(Distance(g, tu, c)) + Distance(g, tv, c);
      for (int32_t a = begin; a < end; ++a) {
        for (int32_t b = a + 1; b < end; ++b) {
          const int32_t f = g.neighbors[a], s = g.neighbors[b];
          if (std::min(f, s) == tu && std::max(f, s) == tv) continue;
          if (!survives(c, f) || !survives(c, s)) continue;
          const __int128 original =
              static_cast<__int128>(Distance(g, tu, tv)) + Distance(g, c, f) + Distance(g, c, s);
          const __int128 replacement = static_cast<__int128>(Distance(g, f, s)) + center_target;
          if (original <= replacement) {
            if (total < capacity) {
              replies[3 * total] = c;
              replies[3 * total + 1] = f;
              replies[3 * total + 2] = s;
            }
            ++total;
          }
        }
      }
      offsets[i + 1] = total;
    }
  });
  return status == 0 ? static_cast<int64_t>(total) : status;
}

int64_t orc_ht_end_replies(const orc_graph* graph, const int32_t* tasks, size_t count,
                           uint64_t* offsets, int32_t* replies, uint64_t capacity) {
  uint64_t total = 0;
  const int status = Guard([&] {
    const orc_graph& g = *graph;
    for (size_t i = 0; i < count; ++i) {
      const int32_t e = tasks[2 * i], in = tasks[2 * i + 1];
      if (e < 0 || e >= g.dimension || in < 0 || in >= g.dimension || e == in ||
          !HasActiveEdge(g, e, in)) {
        throw std::invalid_argument("HT end reply task must name an active path-internal edge");
      }
    }
    offsets[0] = 0;
    for (size_t i = 0; i < count; ++i) {
      const int32_t e = tasks[2 * i], in = tasks[2 * i + 1];
      for (int32_t a = g.row_offsets[e]; a < g.row_offsets[e + 1]; ++a) {
        const int32_t nb = g.neighbors[a];
        if (nb == in) continue;
        if (total < capacity) {
          replies[2 * total] = std::min(e, nb);
          replies[2 * total + 1] = std::max(e, nb);
        }
        ++total;
      }
      offsets[i + 1] = total;
    }
  });
  return status == 0 ? static_cast<int64_t>(total) : status;
}

}  // extern "C"
