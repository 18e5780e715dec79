// Seeded synthetic workloads and host graph helpers (see ee_gen.h).
#include "ee_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <unordered_set>
#include <vector>

#include "ee_exact.h"
#include "ee_runtime.h"

namespace ee {
namespace {

// ------------------------------------------------------------------ PRNG
class Rng {
 public:
  explicit Rng(uint64_t seed) {
    uint64_t sm = seed;
    for (auto& word : s_) {
      sm += 0x9E3779B97F4A7C15ull;
      uint64_t z = sm;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      word = z ^ (z >> 31);
    }
  }
  // xoshiro256**
  uint64_t Next() {
    const uint64_t result = Rotl(s_[1] * 5, 7) * 9;
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = Rotl(s_[3], 45);
    return result;
  }
  // Uniform integer in [0, range) (Lemire: multiply, reject the biased low zone).
  uint64_t Below(uint64_t range) {
    if (range == 0) throw std::invalid_argument("empty random range");
    unsigned __int128 m = static_cast<unsigned __int128>(Next()) * range;
    uint64_t low = static_cast<uint64_t>(m);
    if (low < range) {
      const uint64_t threshold = (0 - range) % range;
      while (low < threshold) {
        m = static_cast<unsigned __int128>(Next()) * range;
        low = static_cast<uint64_t>(m);
      }
    }
    return static_cast<uint64_t>(m >> 64);
  }
  int64_t Between(int64_t lo, int64_t hi) {  // inclusive
    return lo + static_cast<int64_t>(Below(static_cast<uint64_t>(hi - lo) + 1));
  }
  double Unit() { return static_cast<double>(Next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double Gaussian() {
    const double u1 = 1.0 - Unit();  // (0, 1]
    const double u2 = Unit();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }

 private:
  static uint64_t Rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t s_[4];
};

uint64_t PointKey(int64_t x, int64_t y) {
  return static_cast<uint64_t>(x) * 0x9E3779B97F4A7C15ull ^ static_cast<uint64_t>(y);
}

// Distinct points: a repeated coordinate pair is redrawn.
template <typename Draw>
void DistinctPoints(int32_t n, int64_t* x, int64_t* y, Draw draw) {
  struct Hash {
    size_t operator()(const std::pair<int64_t, int64_t>& p) const { return PointKey(p.first, p.second); }
  };
  std::unordered_set<std::pair<int64_t, int64_t>, Hash> seen;
  seen.reserve(static_cast<size_t>(n) * 2);
  for (int32_t i = 0; i < n; ++i) {
    for (int attempt = 0;; ++attempt) {
      if (attempt > 1000000) throw std::runtime_error("cannot draw distinct points");
      std::pair<int64_t, int64_t> p = draw();
      if (seen.insert(p).second) {
        x[i] = p.first;
        y[i] = p.second;
        break;
      }
    }
  }
}

// ------------------------------------------------------------------ kNN
struct Neighbor {
  uint64_t s;
  int32_t id;
  bool operator<(const Neighbor& o) const { return s < o.s || (s == o.s && id < o.id); }
};

uint64_t Squared(int64_t ax, int64_t ay, int64_t bx, int64_t by) {
  const uint64_t dx = AbsDiff(ax, bx), dy = AbsDiff(ay, by);
  return dx * dx + dy * dy;
}

// Exact k nearest neighbours of every point with a uniform grid: rings of
// cells are scanned outward until the k-th best squared distance is strictly
// below the squared gap to the next ring, so ties are resolved by id exactly.
std::vector<std::vector<int32_t>> KNearest(const int64_t* x, const int64_t* y, int32_t n, int32_t k) {
  int64_t lx = x[0], hx = x[0], ly = y[0], hy = y[0];
  for (int32_t i = 1; i < n; ++i) {
    lx = std::min(lx, x[i]);
    hx = std::max(hx, x[i]);
    ly = std::min(ly, y[i]);
    hy = std::max(hy, y[i]);
  }
  const int64_t extent = std::max<int64_t>(std::max(hx - lx, hy - ly), 1) + 1;
  const int64_t side = std::max<int64_t>(1, static_cast<int64_t>(std::sqrt(n / 2.0)));
  const int64_t cell = (extent + side - 1) / side;
  const int64_t gx = (hx - lx) / cell + 1, gy = (hy - ly) / cell + 1;
  std::vector<int32_t> head(static_cast<size_t>(gx * gy) + 1, 0), order(static_cast<size_t>(n));
  std::vector<int64_t> cx(static_cast<size_t>(n)), cy(static_cast<size_t>(n));
  for (int32_t i = 0; i < n; ++i) {
    cx[static_cast<size_t>(i)] = (x[i] - lx) / cell;
    cy[static_cast<size_t>(i)] = (y[i] - ly) / cell;
    ++head[static_cast<size_t>(cy[static_cast<size_t>(i)] * gx + cx[static_cast<size_t>(i)]) + 1];
  }
  for (size_t c = 1; c < head.size(); ++c) head[c] += head[c - 1];
  {
    std::vector<int32_t> fill(head.begin(), head.end() - 1);
    for (int32_t i = 0; i < n; ++i) {
      order[static_cast<size_t>(fill[static_cast<size_t>(cy[static_cast<size_t>(i)] * gx + cx[static_cast<size_t>(i)])]++)] = i;
    }
  }
  const int32_t want = std::min(k, n - 1);
  std::vector<std::vector<int32_t>> out(static_cast<size_t>(n));
  std::vector<Neighbor> best;
  for (int32_t p = 0; p < n; ++p) {
    best.clear();
    const int64_t px = cx[static_cast<size_t>(p)], py = cy[static_cast<size_t>(p)];
    const auto visit = [&](int64_t qx, int64_t qy) {
      if (qx < 0 || qy < 0 || qx >= gx || qy >= gy) return;
      const size_t c = static_cast<size_t>(qy * gx + qx);
      for (int32_t slot = head[c]; slot < head[c + 1]; ++slot) {
        const int32_t q = order[static_cast<size_t>(slot)];
        if (q == p) continue;
        const Neighbor cand{Squared(x[p], y[p], x[q], y[q]), q};
        if (static_cast<int32_t>(best.size()) < want) {
          best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
        } else if (cand < best.back()) {
          best.pop_back();
          best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
        }
      }
    };
    for (int64_t r = 0;; ++r) {
      if (r == 0) {
        visit(px, py);
      } else {
        for (int64_t d = -r; d <= r; ++d) {
          visit(px + d, py - r);
          visit(px + d, py + r);
        }
        for (int64_t d = -r + 1; d <= r - 1; ++d) {
          visit(px - r, py + d);
          visit(px + r, py + d);
        }
      }
      const bool covered = px - r <= 0 && py - r <= 0 && px + r >= gx - 1 && py + r >= gy - 1;
      if (covered) break;
      if (static_cast<int32_t>(best.size()) == want) {
        const unsigned __int128 gap = static_cast<unsigned __int128>(r) * static_cast<uint64_t>(cell);
        if (static_cast<unsigned __int128>(best.back().s) < gap * gap) break;
      }
    }
    for (const Neighbor& nb : best) out[static_cast<size_t>(p)].push_back(nb.id);
  }
  return out;
}

template <typename F> ee_status Guarded(F&& body) {
  try {
    body();
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

}  // namespace
}  // namespace ee

using ee::Guarded;

namespace ee {

uint64_t Fnv1a(uint64_t h, const void* p, size_t len) {
  const auto* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < len; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

// ContentHash (src/common/graph.cpp:323-346): FNV-1a-64 over the raw bytes
// of dimension (int32), distance type (uint8), the integer-coordinates flag,
// every point's int64 x, y, then every edge's int32 u, int32 v, int64
// weight and bool active.  The prefix covers everything before the edges.
uint64_t ContentHashPrefix(int32_t n, uint8_t distance_type, const int64_t* x, const int64_t* y) {
  uint64_t h = 14695981039346656037ULL;
  const uint8_t integer = 1;
  h = Fnv1a(h, &n, 4);
  h = Fnv1a(h, &distance_type, 1);
  h = Fnv1a(h, &integer, 1);
  for (int32_t i = 0; i < n; ++i) {
    h = Fnv1a(h, &x[i], 8);
    h = Fnv1a(h, &y[i], 8);
  }
  return h;
}

uint64_t ContentHashEdges(uint64_t h, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                          const int64_t* edge_weight, const uint8_t* edge_active) {
  for (int32_t e = 0; e < m; ++e) {
    const uint8_t active = edge_active == nullptr || edge_active[e] != 0 ? 1 : 0;
    h = Fnv1a(h, &edge_u[e], 4);
    h = Fnv1a(h, &edge_v[e], 4);
    h = Fnv1a(h, &edge_weight[e], 8);
    h = Fnv1a(h, &active, 1);
  }
  return h;
}

}  // namespace ee

extern "C" {

ee_status ee_gen_points_uniform(uint64_t seed, int32_t n, int64_t lo, int64_t hi, int64_t* x,
                                int64_t* y) {
  return Guarded([&] {
    if (n <= 0 || x == nullptr || y == nullptr || lo > hi) throw std::invalid_argument("bad point request");
    ee::Rng rng(seed);
    ee::DistinctPoints(n, x, y, [&] {
      const int64_t px = rng.Between(lo, hi);
      const int64_t py = rng.Between(lo, hi);
      return std::make_pair(px, py);
    });
  });
}

ee_status ee_gen_points_lattice(uint64_t seed, int32_t side_x, int32_t side_y, int64_t spacing,
                                double drop_fraction, int64_t jitter, int32_t* n_out, int64_t* x,
                                int64_t* y) {
  return Guarded([&] {
    if (side_x <= 0 || side_y <= 0 || jitter < 0 || drop_fraction < 0 || drop_fraction >= 1 ||
        n_out == nullptr || x == nullptr || y == nullptr) {
      throw std::invalid_argument("bad lattice request");
    }
    ee::Rng rng(seed);
    const int64_t sites = static_cast<int64_t>(side_x) * side_y;
    const int64_t drop = std::llround(drop_fraction * static_cast<double>(sites));
    std::vector<int64_t> perm(static_cast<size_t>(sites));
    std::iota(perm.begin(), perm.end(), 0);
    for (int64_t i = 0; i < drop; ++i) {  // partial Fisher-Yates: first `drop` are dropped
      const int64_t j = i + static_cast<int64_t>(rng.Below(static_cast<uint64_t>(sites - i)));
      std::swap(perm[static_cast<size_t>(i)], perm[static_cast<size_t>(j)]);
    }
    std::vector<bool> dropped(static_cast<size_t>(sites), false);
    for (int64_t i = 0; i < drop; ++i) dropped[static_cast<size_t>(perm[static_cast<size_t>(i)])] = true;
    int32_t n = 0;
    for (int64_t site = 0; site < sites; ++site) {
      if (dropped[static_cast<size_t>(site)]) continue;
      const int64_t row = site / side_x, col = site % side_x;
      x[n] = col * spacing + rng.Between(-jitter, jitter);
      y[n] = row * spacing + rng.Between(-jitter, jitter);
      ++n;
    }
    *n_out = n;
  });
}

ee_status ee_gen_points_clustered(uint64_t seed, int32_t n, int32_t clusters, double sigma,
                                  int64_t lo, int64_t hi, int64_t* x, int64_t* y) {
  return Guarded([&] {
    if (n <= 0 || clusters <= 0 || !(sigma > 0) || lo > hi || x == nullptr || y == nullptr) {
      throw std::invalid_argument("bad clustered point request");
    }
    ee::Rng rng(seed);
    std::vector<std::pair<double, double>> centre(static_cast<size_t>(clusters));
    for (auto& c : centre) {
      c.first = static_cast<double>(rng.Between(lo, hi));
      c.second = static_cast<double>(rng.Between(lo, hi));
    }
    ee::DistinctPoints(n, x, y, [&] {
      const auto& c = centre[static_cast<size_t>(rng.Below(static_cast<uint64_t>(clusters)))];
      const int64_t px = std::clamp<int64_t>(std::llround(c.first + sigma * rng.Gaussian()), lo, hi);
      const int64_t py = std::clamp<int64_t>(std::llround(c.second + sigma * rng.Gaussian()), lo, hi);
      return std::make_pair(px, py);
    });
  });
}

ee_status ee_gen_knn_graph(const int64_t* x, const int64_t* y, int32_t n, uint8_t distance_type,
                           int32_t k, int32_t* edge_u, int32_t* edge_v, int64_t* edge_weight,
                           int64_t capacity, int64_t* edge_count) {
  return Guarded([&] {
    if (n <= 1 || k <= 0 || x == nullptr || y == nullptr || edge_count == nullptr) {
      throw std::invalid_argument("bad kNN request");
    }
    if (!ee::DistanceSafe(x, y, n)) throw std::invalid_argument("coordinates are not distance safe");
    const auto nearest = ee::KNearest(x, y, n, k);
    std::vector<uint64_t> keys;
    for (int32_t p = 0; p < n; ++p) {
      for (const int32_t q : nearest[static_cast<size_t>(p)]) {
        const uint32_t a = static_cast<uint32_t>(std::min(p, q)), b = static_cast<uint32_t>(std::max(p, q));
        keys.push_back((static_cast<uint64_t>(a) << 32) | b);
      }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    *edge_count = static_cast<int64_t>(keys.size());
    if (edge_u == nullptr) return;
    if (capacity < static_cast<int64_t>(keys.size())) throw std::invalid_argument("edge capacity too small");
    for (size_t i = 0; i < keys.size(); ++i) {
      const int32_t a = static_cast<int32_t>(keys[i] >> 32), b = static_cast<int32_t>(keys[i] & 0xffffffffu);
      edge_u[i] = a;
      edge_v[i] = b;
      edge_weight[i] = ee::ExactDistance(x[a], y[a], x[b], y[b], distance_type);
    }
  });
}

ee_status ee_graph_build_csr(int32_t n, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                             const uint8_t* edge_active, int32_t* row_offsets, int32_t* neighbors,
                             int32_t* csr_edge_ids) {
  return Guarded([&] {
    if (n <= 0 || m < 0 || row_offsets == nullptr) throw std::invalid_argument("bad CSR request");
    std::fill(row_offsets, row_offsets + n + 1, 0);
    bool canonical = true;
    for (int32_t e = 0; e < m; ++e) {
      if (edge_u[e] < 0 || edge_v[e] < 0 || edge_u[e] >= n || edge_v[e] >= n) {
        throw std::invalid_argument("edge endpoint out of range");
      }
      if (edge_u[e] >= edge_v[e] ||
          (e > 0 && (edge_u[e - 1] > edge_u[e] || (edge_u[e - 1] == edge_u[e] && edge_v[e - 1] > edge_v[e])))) {
        canonical = false;
      }
      if (edge_active != nullptr && edge_active[e] == 0) continue;
      ++row_offsets[edge_u[e] + 1];
      ++row_offsets[edge_v[e] + 1];
    }
    for (int32_t v = 0; v < n; ++v) row_offsets[v + 1] += row_offsets[v];
    std::vector<int32_t> fill(row_offsets, row_offsets + n);
    for (int32_t e = 0; e < m; ++e) {
      if (edge_active != nullptr && edge_active[e] == 0) continue;
      int32_t s = fill[static_cast<size_t>(edge_u[e])]++;
      neighbors[s] = edge_v[e];
      csr_edge_ids[s] = e;
      s = fill[static_cast<size_t>(edge_v[e])]++;
      neighbors[s] = edge_u[e];
      csr_edge_ids[s] = e;
    }
    if (canonical) return;
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int32_t v = 0; v < n; ++v) {
      row.clear();
      for (int32_t s = row_offsets[v]; s < row_offsets[v + 1]; ++s) row.emplace_back(neighbors[s], csr_edge_ids[s]);
      std::sort(row.begin(), row.end());
      for (int32_t s = row_offsets[v]; s < row_offsets[v + 1]; ++s) {
        neighbors[s] = row[static_cast<size_t>(s - row_offsets[v])].first;
        csr_edge_ids[s] = row[static_cast<size_t>(s - row_offsets[v])].second;
      }
    }
  });
}

uint64_t ee_graph_content_hash(int32_t n, uint8_t distance_type, const int64_t* x,
                               const int64_t* y, int32_t m, const int32_t* edge_u,
                               const int32_t* edge_v, const int64_t* edge_weight,
                               const uint8_t* edge_active) {
  return ee::ContentHashEdges(ee::ContentHashPrefix(n, distance_type, x, y), m, edge_u, edge_v,
                              edge_weight, edge_active);
}

ee_status ee_gen_kopt_tasks(uint64_t seed, const int64_t* x, const int64_t* y, int32_t n,
                            uint8_t distance_type, const int32_t* row_offsets,
                            const int32_t* neighbors, uint32_t k, const uint8_t* ks, uint64_t count,
                            ee_kopt_task* tasks) {
  return Guarded([&] {
    if (n <= 1 || x == nullptr || y == nullptr || row_offsets == nullptr || tasks == nullptr) {
      throw std::invalid_argument("bad task request");
    }
    ee::Rng rng(seed);
    std::vector<int32_t> path;
    std::vector<uint8_t> on_path(static_cast<size_t>(n), 0);
    std::vector<int32_t> options;
    for (uint64_t r = 0; r < count; ++r) {
      const uint32_t kr = ks != nullptr ? ks[r] : k;
      if (kr < 3 || kr > 5) throw std::invalid_argument("task k must be in [3,5]");
      const uint32_t walk_edges = kr + static_cast<uint32_t>(rng.Below(kr + 1));
      for (int attempt = 0;; ++attempt) {
        if (attempt > 100000) throw std::runtime_error("graph has no simple walk of the needed length");
        for (const int32_t v : path) on_path[static_cast<size_t>(v)] = 0;
        path.clear();
        int32_t at = static_cast<int32_t>(rng.Below(static_cast<uint64_t>(n)));
        path.push_back(at);
        on_path[static_cast<size_t>(at)] = 1;
        while (path.size() < walk_edges + 1) {
          options.clear();
          for (int32_t s = row_offsets[at]; s < row_offsets[at + 1]; ++s) {
            if (!on_path[static_cast<size_t>(neighbors[s])]) options.push_back(neighbors[s]);
          }
          if (options.empty()) break;
          at = options[static_cast<size_t>(rng.Below(options.size()))];
          path.push_back(at);
          on_path[static_cast<size_t>(at)] = 1;
        }
        if (path.size() == walk_edges + 1) break;
      }
      ee_kopt_task t{};
      uint32_t need = kr, e = 0;
      for (uint32_t i = 0; i < walk_edges && need > 0; ++i) {  // selection sampling, in order
        if (rng.Below(walk_edges - i) < need) {
          const int32_t a = path[i], b = path[i + 1];
          t.port_nodes[2 * e] = a;
          t.port_nodes[2 * e + 1] = b;
          t.deleted_cost += ee::ExactDistance(x[a], y[a], x[b], y[b], distance_type);
          ++e;
          --need;
        }
      }
      tasks[r] = t;
    }
  });
}

ee_status ee_gen_shuffled_ks(uint64_t seed, const uint64_t quota[3], uint8_t* ks) {
  return Guarded([&] {
    if (quota == nullptr || ks == nullptr) throw std::invalid_argument("bad k request");
    const uint64_t total = quota[0] + quota[1] + quota[2];
    uint64_t at = 0;
    for (int i = 0; i < 3; ++i) {
      for (uint64_t j = 0; j < quota[i]; ++j) ks[at++] = static_cast<uint8_t>(3 + i);
    }
    ee::Rng rng(seed);
    for (uint64_t i = total; i > 1; --i) {
      const uint64_t j = rng.Below(i);
      std::swap(ks[i - 1], ks[j]);
    }
  });
}

}  // extern "C"
