// Test harness, not part of the product: the backend's host-only units
// (ee_host.cpp, ee_gen.cpp, ee_jvhost.cpp, ee_certify.cpp) and the oracle,
// built with AddressSanitizer and UndefinedBehaviorSanitizer by plain g++
// (oracle/Makefile, target sanitize; run by tests/test_host.py), as the
// reference builds its host code with sanitizers (CMakeLists.txt:100-106).
//
// Every step is also checked against the oracle:
//  - seeded points (uniform, clustered, lattice), k-NN graphs, CSR and the
//    snapshot content hash;
//  - the canonical reconnect templates and their generator hash;
//  - k-opt task rows from random walks, certified against the oracle's CPU
//    matrix, and a planted cell difference found at its index;
//  - a JV host run (re-verification, commit rule, CSR rebuild, proof text,
//    snapshot hashes on helper threads) fed each epoch with the oracle's
//    witnesses, whose proof equals the oracle's own run;
//  - the host's two exact distance forms (the integer root, and the FP64
//    form for extents below 2^25 that re-verification and certification
//    use) against the oracle's restoring square root;
//  - the certification pool: every iteration exactly once and the first
//    exception rethrown after all of them, for pools of 1 to 16 threads,
//    two of them driven at once.
#include <atomic>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "ee_certify.h"
#include "ee_exact.h"
#include "ee_gen.h"
#include "ee_hip.h"
#include "ee_jvhost.h"
#include "ee_oracle.h"

namespace {

int g_failures = 0;

void Expect(bool ok, const std::string& what) {
  if (!ok) {
    std::cout << "FAIL " << what << '\n';
    ++g_failures;
  }
}

struct Graph {
  int32_t n = 0;
  uint8_t type = 0;
  std::vector<int64_t> x, y, w;
  std::vector<int32_t> u, v, ro, nb, eid;
  std::vector<uint8_t> active;

  ee_graph_view View() const {
    ee_graph_view g{};
    g.dimension = n;
    g.distance_type = type;
    g.integer_coordinates = 1;
    g.x = x.data();
    g.y = y.data();
    g.edge_count = static_cast<int32_t>(u.size());
    g.edge_u = u.data();
    g.edge_v = v.data();
    g.edge_weight = w.data();
    g.edge_active = active.data();
    g.row_offsets = ro.data();
    g.neighbors = nb.data();
    g.csr_edge_ids = eid.data();
    return g;
  }
};

Graph KnnGraph(std::vector<int64_t> x, std::vector<int64_t> y, uint8_t type, int32_t k) {
  Graph g;
  g.n = static_cast<int32_t>(x.size());
  g.type = type;
  g.x = std::move(x);
  g.y = std::move(y);
  int64_t m = 0;
  Expect(ee_gen_knn_graph(g.x.data(), g.y.data(), g.n, type, k, nullptr, nullptr, nullptr, 0, &m) == EE_OK && m > 0,
         "k-NN graph size query");
  g.u.resize(static_cast<size_t>(m));
  g.v.resize(static_cast<size_t>(m));
  g.w.resize(static_cast<size_t>(m));
  Expect(ee_gen_knn_graph(g.x.data(), g.y.data(), g.n, type, k, g.u.data(), g.v.data(), g.w.data(), m, &m) == EE_OK,
         "k-NN graph");
  g.active.assign(static_cast<size_t>(m), 1);
  g.ro.resize(static_cast<size_t>(g.n) + 1);
  g.nb.resize(2 * static_cast<size_t>(m));
  g.eid.resize(2 * static_cast<size_t>(m));
  Expect(ee_graph_build_csr(g.n, static_cast<int32_t>(m), g.u.data(), g.v.data(), g.active.data(), g.ro.data(),
                            g.nb.data(), g.eid.data()) == EE_OK,
         "CSR");
  return g;
}

orc_graph* OracleGraph(const Graph& g) {
  return orc_graph_from_arrays(g.n, g.type, g.x.data(), g.y.data(), static_cast<int32_t>(g.u.size()), g.u.data(),
                               g.v.data(), g.w.data(), g.active.data());
}

void CheckSnapshot(const Graph& g, const std::string& name) {
  orc_graph* og = OracleGraph(g);
  Expect(og != nullptr, name + ": oracle graph");
  if (og == nullptr) return;
  const uint64_t hash = ee_graph_content_hash(g.n, g.type, g.x.data(), g.y.data(), static_cast<int32_t>(g.u.size()),
                                              g.u.data(), g.v.data(), g.w.data(), g.active.data());
  Expect(hash == orc_graph_hash(og), name + ": content hash equals the oracle's");
  std::vector<int32_t> ro(static_cast<size_t>(g.n) + 1), nb(g.nb.size()), eid(g.eid.size());
  orc_graph_export(og, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ro.data(), nb.data(), eid.data(),
                   nullptr);
  Expect(ro == g.ro && nb == g.nb && eid == g.eid, name + ": CSR equals the oracle's");
  orc_graph_free(og);
}

std::string Canonical(const std::string& proof) {
  std::istringstream in(proof);
  std::ostringstream out;
  std::string line;
  while (std::getline(in, line)) {
    if (line.rfind("backend ", 0) == 0) {
      out << "backend *\n";
    } else if (line.rfind("metrics ", 0) == 0) {
      std::istringstream f(line);
      std::string word;
      for (int i = 0; i < 7 && f >> word; ++i) out << (i ? " " : "") << word;
      out << '\n';
    } else {
      out << line << '\n';
    }
  }
  return out.str();
}

void JvRun(const Graph& g, const std::string& name) {
  ee::JvHost host;
  host.async_hash = true;
  const ee_graph_view view = g.View();
  host.Load(&view);
  orc_graph* og = OracleGraph(g);
  host.StartRun();
  host.BeginEpoch(0);
  std::vector<int32_t> witnesses(g.u.size());
  for (uint32_t epoch = 0;; ++epoch) {
    orc_jv_find_candidates(og, witnesses.data(), nullptr);
    const std::vector<int32_t>& commit = host.VerifyAndGate(witnesses.data(), 0.0);
    for (const int32_t e : commit) orc_graph_set_active(og, e, 0);
    orc_graph_rebuild_csr(og);
    const bool more = !commit.empty() && epoch + 1 < 100;
    host.ApplyCommit();
    if (!more) break;
    host.BeginEpoch(epoch + 1);
  }
  host.EndRun(0.0);
  orc_graph_free(og);
  orc_graph* fresh = OracleGraph(g);
  std::vector<char> buf(1 << 22);
  const int64_t size = orc_jv_run(fresh, 100, "hip", buf.data(), buf.size());
  orc_graph_free(fresh);
  Expect(size > 0 && static_cast<size_t>(size) < buf.size(), name + ": oracle JV run");
  Expect(Canonical(host.proof) == Canonical(std::string(buf.data())), name + ": JV host proof equals the oracle's");
  Expect(host.stats.committed > 0, name + ": the run commits deletions");
}

void KoptChecks(const Graph& g) {
  for (uint32_t k = 3; k <= 5; ++k) {
    std::vector<ee_endpoint_matching> tpl(945);
    uint32_t count = 0;
    uint64_t hash = 0;
    Expect(ee_kopt_build_templates(k, tpl.data(), 945, &count, &hash) == EE_OK, "templates");
    std::vector<uint8_t> mates(945 * 14);
    uint32_t ocount = 0;
    uint64_t ohash = 0;
    orc_kopt_templates(k, mates.data(), 945, &ocount, &ohash);
    bool same = count == ocount && hash == ohash;
    for (uint32_t t = 0; same && t < count; ++t) same = std::memcmp(tpl[t].mate, &mates[t * 14], 2 * k) == 0;
    Expect(same, "k=" + std::to_string(k) + " templates and hash equal the oracle's");
    const uint64_t rows = 700;
    std::vector<ee_kopt_task> tasks(rows);
    Expect(ee_gen_kopt_tasks(0xA5A5 + k, g.x.data(), g.y.data(), g.n, g.type, g.ro.data(), g.nb.data(), k, nullptr,
                             rows, tasks.data()) == EE_OK,
           "walk task rows");
    std::vector<int64_t> cells(rows * count);
    uint32_t threads = 0, nodes = 0;
    orc_kopt_costs(g.x.data(), g.y.data(), g.n, g.type, k, reinterpret_cast<const orc_kopt_task*>(tasks.data()), rows,
                   cells.data(), 8, &threads, &nodes);
    std::vector<uint32_t> improving(rows * ((count + 31) / 32));
    uint64_t mismatch = 0, better = 0;
    Expect(ee_kopt_certify(g.x.data(), g.y.data(), g.n, g.type, k, tpl.data(), count, tasks.data(), rows,
                           cells.data(), improving.data(), &mismatch, &better) == EE_OK,
           "k=" + std::to_string(k) + " oracle matrix certifies");
    const uint64_t at = cells.size() / 2 + k;
    cells[at] ^= 1;
    Expect(ee_kopt_certify(g.x.data(), g.y.data(), g.n, g.type, k, tpl.data(), count, tasks.data(), rows,
                           cells.data(), improving.data(), &mismatch, &better) == EE_ERR_RUNTIME &&
               mismatch == at,
           "k=" + std::to_string(k) + " planted difference found at its cell");
  }
}

// Both host distance forms against the oracle, both rounding types: random
// differences, squares and their neighbours, roots just below an integer
// (s = r^2 - 1 for large r), right triangles, and the extremes.
void DistanceChecks() {
  const int64_t lim = (int64_t{1} << 25) - 1;
  uint64_t bad = 0;
  const auto check = [&](int64_t dx, int64_t dy) {
    for (int t = 0; t < 2; ++t) {
      int64_t want = 0;
      if (orc_exact_distance(0, 0, dx, dy, t, &want) != 0) {
        ++bad;
        continue;
      }
      const int64_t exact = ee::ExactDistance(3, -5, 3 + dx, -5 + dy, t);
      const int64_t near = ee::ExactDistanceNear(static_cast<int32_t>(dx), static_cast<int32_t>(dy), t);
      if ((exact != want || near != want) && bad++ < 5) {
        std::cout << "distance dx=" << dx << " dy=" << dy << " type=" << t << ": oracle " << want
                  << ", integer form " << exact << ", FP64 form " << near << '\n';
      }
    }
  };
  std::mt19937_64 rng(0xEE5D);
  std::uniform_int_distribution<int64_t> u(-lim, lim);
  for (int i = 0; i < 200000; ++i) check(u(rng), u(rng));
  for (int64_t r = lim; r > lim - 2000; --r) {
    check(r, 0);
    check(r, 1);   // s = r^2 + 1
    check(r - 1, 0);
  }
  for (int64_t r = 2; r < 3000; ++r) {
    for (int64_t d = -2; d <= 2; ++d) check(r, d);  // s within 4 of r^2
  }
  // EUC_2D's half-way points: s - m^2 = m (m = q^2, dy = q; not rounded up)
  // and s - m^2 = m + 1 (m = q^2 - 1, dy = q; rounded up)
  for (int64_t q = 1; q * q <= lim; ++q) {
    check(q * q, q);
    check(q * q - 1, q);
  }
  for (int64_t k = 1; 5 * k <= lim; k += 4099) {
    check(3 * k, 4 * k);
    check(3 * k, 4 * k - 1);
    check(3 * k + 1, 4 * k);
  }
  check(0, 0);
  check(lim, lim);
  check(-lim, lim);
  check(lim, -lim);
  Expect(bad == 0, "host distance forms equal the oracle's");
}

void PoolChecks() {
  std::atomic<int> bad{0};
  const auto drive = [&](unsigned threads, unsigned seed, int runs) {
    ee::HostPool pool(threads);
    std::mt19937 rng(seed);
    for (int r = 0; r < runs; ++r) {
      const size_t n = rng() % 200;
      std::vector<std::atomic<int>> hits(n);
      for (auto& h : hits) h.store(0);
      const bool plant = n > 0 && rng() % 5 == 0;
      const size_t at = plant ? rng() % n : 0;
      bool caught = false;
      try {
        pool.Run(n, [&](size_t i) {
          hits[i].fetch_add(1);
          if (plant && i == at) throw std::runtime_error("planted");
        });
      } catch (const std::runtime_error&) {
        caught = true;
      }
      for (size_t i = 0; i < n; ++i) bad += hits[i].load() != 1;
      bad += caught != plant;
    }
  };
  std::thread other(drive, 4u, 0xEE61u, 300);
  drive(16, 0xEE62, 300);
  other.join();
  drive(1, 0xEE63, 100);
  Expect(bad.load() == 0, "certification pool runs every iteration once and rethrows");
}

}  // namespace

int main() {
  DistanceChecks();
  PoolChecks();
  // uniform points, EUC_2D, 10-NN
  {
    std::vector<int64_t> x(600), y(600);
    Expect(ee_gen_points_uniform(0xEE51, 600, 0, 100000, x.data(), y.data()) == EE_OK, "uniform points");
    const Graph g = KnnGraph(x, y, EE_EUC_2D, 10);
    CheckSnapshot(g, "uniform");
    JvRun(g, "uniform");
    KoptChecks(g);
  }
  // clustered points, CEIL_2D, 8-NN
  {
    std::vector<int64_t> x(500), y(500);
    Expect(ee_gen_points_clustered(0xEE52, 500, 6, 3000.0, 0, 100000, x.data(), y.data()) == EE_OK,
           "clustered points");
    const Graph g = KnnGraph(x, y, EE_CEIL_2D, 8);
    CheckSnapshot(g, "clustered");
    JvRun(g, "clustered");
  }
  // jittered lattice, EUC_2D, 10-NN (ties on the exact lattice spacing)
  {
    std::vector<int64_t> x(30 * 30), y(30 * 30);
    int32_t n = 0;
    Expect(ee_gen_points_lattice(0xEE53, 30, 30, 25, 0.17, 3, &n, x.data(), y.data()) == EE_OK, "lattice points");
    x.resize(static_cast<size_t>(n));
    y.resize(static_cast<size_t>(n));
    const Graph g = KnnGraph(x, y, EE_EUC_2D, 10);
    CheckSnapshot(g, "lattice");
    JvRun(g, "lattice");
  }
  std::cout << (g_failures == 0 ? "sanitized host checks passed" : "sanitized host checks failed") << '\n';
  return g_failures == 0 ? 0 : 1;
}
