// Resident multi-epoch JV elimination: the epoch loop of RunJvElimination
// (src/cpu/jv_elimination.cpp:250-322) with the propose step on the device.
//
// Per epoch: snapshot hash, device proposal over every edge slot, host
// re-verification of every candidate (jv_elimination.cpp:226-248; a rejected
// device candidate stops the run before anything is committed, as the
// reference does at :301-303), the deterministic commit with the running
// degree gate (src/cpu/elimination_commit.cpp:13-75), CSR rebuild and one
// proof record per committed edge.  The static part of the snapshot stays on
// the device; active flags and the CSR are refreshed each epoch.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ee_exact.h"
#include "ee_gen.h"
#include "ee_runtime.h"

namespace ee {
namespace {

struct Session {
  explicit Session(Context* c) : ctx(c) {}

  Context* ctx;
  int32_t n = 0, m = 0;
  uint8_t type = 0;
  bool near = false;  // coordinate extents below 2^25: FP64 distances on the device
  std::vector<int64_t> x, y, weight;
  std::vector<int32_t> u, v;
  std::vector<uint8_t> active;
  std::vector<int32_t> row_offsets, neighbors, edge_ids, active32, witnesses;
  // device
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  DeviceArray<int32_t> d_active, d_row_offsets, d_neighbors, d_edge_ids, d_witnesses;
  std::string proof;

  int32_t Degree(int32_t vtx) const { return row_offsets[vtx + 1] - row_offsets[vtx]; }

  int64_t Dist(int32_t a, int32_t b) const { return ExactDistance(x[a], y[a], x[b], y[b], type); }

  void RebuildCsr() {
    row_offsets.assign(static_cast<size_t>(n) + 1, 0);
    int64_t live = 0;
    for (int32_t e = 0; e < m; ++e) live += active[e] != 0;
    neighbors.assign(static_cast<size_t>(2 * live), 0);
    edge_ids.assign(static_cast<size_t>(2 * live), 0);
    CheckStatus(ee_graph_build_csr(n, m, u.data(), v.data(), active.data(), row_offsets.data(),
                                   neighbors.data(), edge_ids.data()));
  }

  static void CheckStatus(ee_status s) {
    if (s != EE_OK) throw std::runtime_error(LastError());
  }

  uint64_t Hash() const {
    return ee_graph_content_hash(n, type, x.data(), y.data(), m, u.data(), v.data(), weight.data(),
                                 active.data());
  }

  // IsJvWitness (jv_elimination.cpp:86-117) with 128-bit sums.
  bool Verify(int32_t e, int32_t c) const {
    if (e < 0 || e >= m || !active[e]) return false;
    const int32_t a = u[e], b = v[e];
    if (Degree(a) <= 2 || Degree(b) <= 2) return false;
    if (c < 0 || c >= n || c == a || c == b) return false;
    const __int128 cab = weight[e], cac = Dist(a, c), cbc = Dist(b, c);
    for (int32_t s = row_offsets[c]; s < row_offsets[c + 1]; ++s) {
      const int32_t d = neighbors[s];
      if (d == a || d == b) continue;
      const __int128 left = cab + weight[edge_ids[s]];
      if (left <= cac + Dist(d, b) || left <= Dist(a, d) + cbc) return false;
    }
    return true;
  }

  void Upload(double* h2d_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    active32.resize(static_cast<size_t>(m));
    for (int32_t e = 0; e < m; ++e) active32[e] = active[e] ? 1 : 0;
    d_active.Upload(active32.data(), static_cast<size_t>(m), ctx->stream);
    d_row_offsets.Upload(row_offsets.data(), row_offsets.size(), ctx->stream);
    d_neighbors.Upload(neighbors.data(), neighbors.size(), ctx->stream);
    d_edge_ids.Upload(edge_ids.data(), edge_ids.size(), ctx->stream);
    ctx->Sync();
    *h2d_ms += MsSince(t0);
  }

  void Run(uint32_t max_rounds, ee_jv_run_stats* st) {
    if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
    ctx->Activate();
    const auto t_total = std::chrono::steady_clock::now();
    ee_jv_run_stats s{};
    s.initial_hash = Hash();
    struct Record {
      uint32_t epoch;
      uint64_t hash;
      int32_t edge, a, b, witness;
    };
    struct Metrics {
      uint32_t epoch;
      size_t before, proposed, verified, rejected, committed;
      double propose_ms, verify_ms;
    };
    std::vector<Record> records;
    std::vector<Metrics> metrics;
    witnesses.resize(static_cast<size_t>(m));
    for (uint32_t epoch = 0; epoch < max_rounds; ++epoch) {
      Metrics mt{};
      mt.epoch = epoch;
      for (int32_t e = 0; e < m; ++e) mt.before += active[e] != 0;
      const uint64_t snap = Hash();

      // propose on the device
      auto t0 = std::chrono::steady_clock::now();
      Upload(&s.h2d_ms);
      JvLaunch a{};
      a.edge_count = m;
      a.distance_type = type;
      a.near = near ? 1 : 0;
      a.edge_uv = d_uv.get();
      a.edge_weight = d_weight.get();
      a.edge_active = d_active.get();
      a.row_offsets = d_row_offsets.get();
      a.neighbors = d_neighbors.get();
      a.csr_edge_ids = d_edge_ids.get();
      a.xy = d_xy.get();
      a.witnesses = d_witnesses.get();
      CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
      LaunchJv(a, ctx->jv_grid, ctx->stream);
      CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
      CheckHip(hipEventSynchronize(ctx->ev_end), "event synchronize");
      float kernel = 0.0f;
      CheckHip(hipEventElapsedTime(&kernel, ctx->ev_begin, ctx->ev_end), "event elapsed");
      s.device_ms += kernel;
      const auto t_d2h = std::chrono::steady_clock::now();
      d_witnesses.Download(witnesses.data(), static_cast<size_t>(m), ctx->stream);
      ctx->Sync();
      s.d2h_ms += MsSince(t_d2h);
      std::vector<std::pair<int32_t, int32_t>> proposed;
      for (int32_t e = 0; e < m; ++e) {
        if (witnesses[e] >= 0) proposed.emplace_back(e, witnesses[e]);
      }
      mt.propose_ms = MsSince(t0);
      mt.proposed = proposed.size();

      // host re-verification on the same snapshot
      t0 = std::chrono::steady_clock::now();
      std::vector<std::pair<int32_t, int32_t>> verified;
      for (const auto& c : proposed) {
        if (Verify(c.first, c.second)) {
          verified.push_back(c);
        } else {
          ++mt.rejected;
        }
      }
      mt.verify_ms = MsSince(t0);
      mt.verified = verified.size();
      s.propose_ms += mt.propose_ms;
      s.verify_ms += mt.verify_ms;
      if (mt.rejected != 0) {
        throw std::runtime_error("HIP JV candidate failed CPU re-verification; epoch stopped, nothing committed");
      }

      // deterministic commit: sort by (u, v, method, witness, edge), one per
      // edge, running degree gate > 2 at both ends
      t0 = std::chrono::steady_clock::now();
      if (Hash() != snap) throw std::runtime_error("snapshot changed before commit");
      std::sort(verified.begin(), verified.end(), [&](const auto& l, const auto& r) {
        return std::make_tuple(u[l.first], v[l.first], l.second, l.first) <
               std::make_tuple(u[r.first], v[r.first], r.second, r.first);
      });
      verified.erase(std::unique(verified.begin(), verified.end(),
                                 [](const auto& l, const auto& r) { return l.first == r.first; }),
                     verified.end());
      std::vector<int32_t> degree(static_cast<size_t>(n));
      for (int32_t vtx = 0; vtx < n; ++vtx) degree[vtx] = Degree(vtx);
      std::vector<std::pair<int32_t, int32_t>> committed;
      for (const auto& c : verified) {
        int32_t& du = degree[u[c.first]];
        int32_t& dv = degree[v[c.first]];
        if (du <= 2 || dv <= 2) continue;
        --du;
        --dv;
        committed.push_back(c);
      }
      if (!committed.empty()) {
        for (const auto& c : committed) active[c.first] = 0;
        RebuildCsr();
      }
      s.commit_ms += MsSince(t0);
      mt.committed = committed.size();
      for (const auto& c : committed) {
        records.push_back({epoch, snap, c.first, u[c.first], v[c.first], c.second});
      }
      metrics.push_back(mt);
      s.edges_scanned += mt.before;
      s.proposed += mt.proposed;
      s.committed += mt.committed;
      ++s.epochs;
      if (committed.empty()) break;
    }
    s.final_hash = Hash();
    s.total_ms = MsSince(t_total);
    // V1 proof text (src/cpu/jv_elimination.cpp:324-368)
    std::ostringstream out;
    const auto hex = [](uint64_t h) {
      char buf[17];
      std::snprintf(buf, sizeof(buf), "%016llx", static_cast<unsigned long long>(h));
      return std::string(buf);
    };
    out << "CUDAEE_PROOF_V1\nbackend hip\ninitial_hash " << hex(s.initial_hash) << '\n';
    for (const Record& r : records) {
      out << "record " << r.epoch << ' ' << hex(r.hash) << ' ' << r.edge << ' ' << r.a << ' '
          << r.b << " JV " << r.witness << '\n';
    }
    for (const Metrics& mt : metrics) {
      out << "metrics " << mt.epoch << ' ' << mt.before << ' ' << mt.proposed << ' ' << mt.verified
          << ' ' << mt.rejected << ' ' << mt.committed << ' ' << mt.propose_ms << ' '
          << mt.verify_ms << '\n';
    }
    out << "final_hash " << hex(s.final_hash) << "\nEND\n";
    proof = out.str();
    if (st) *st = s;
  }
};

template <typename F> ee_status Guarded(F&& body) {
  try {
    body();
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

}  // namespace

void CheckGraphView(const ee_graph_view* g);

}  // namespace ee

using ee::Guarded;

extern "C" {

ee_status ee_jv_session_create(ee_ctx* handle, const ee_graph_view* g, ee_jv_session** out) {
  return Guarded([&] {
    if (handle == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    ee::CheckGraphView(g);
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    auto s = std::make_unique<ee::Session>(ctx);
    s->n = g->dimension;
    s->m = g->edge_count;
    s->type = g->distance_type;
    const size_t n = static_cast<size_t>(s->n), m = static_cast<size_t>(s->m);
    s->x.assign(g->x, g->x + n);
    s->y.assign(g->y, g->y + n);
    if (!ee::DistanceSafe(s->x.data(), s->y.data(), s->n)) {
      throw std::runtime_error("HIP JV needs integer coordinates whose squared distances fit uint64");
    }
    s->near = ee::NearExtent(s->x.data(), s->y.data(), s->n);
    if (m > 0) {
      s->u.assign(g->edge_u, g->edge_u + m);
      s->v.assign(g->edge_v, g->edge_v + m);
      s->weight.assign(g->edge_weight, g->edge_weight + m);
      s->active.assign(g->edge_active, g->edge_active + m);
    }
    for (size_t e = 0; e < m; ++e) {
      if (s->u[e] < 0 || s->u[e] >= s->n || s->v[e] < 0 || s->v[e] >= s->n) {
        throw std::invalid_argument("JV edge endpoint out of range");
      }
      if (s->weight[e] < 0 || s->weight[e] > (int64_t{1} << 62)) {
        throw std::invalid_argument("JV edge weight outside [0, 2^62]");
      }
      s->active[e] = s->active[e] ? 1 : 0;
    }
    s->RebuildCsr();
    std::vector<int2> uv(m);
    std::vector<longlong2> xy(n);
    for (size_t e = 0; e < m; ++e) uv[e] = make_int2(s->u[e], s->v[e]);
    for (size_t i = 0; i < n; ++i) xy[i] = longlong2{s->x[i], s->y[i]};
    const size_t adjacency = std::max<size_t>(2 * m, 1);
    s->d_uv.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_weight.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_xy.Allocate(n, ctx->device);
    s->d_active.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_row_offsets.Allocate(n + 1, ctx->device);
    s->d_neighbors.Allocate(adjacency, ctx->device);
    s->d_edge_ids.Allocate(adjacency, ctx->device);
    s->d_witnesses.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_uv.Upload(uv.data(), m, ctx->stream);
    s->d_weight.Upload(s->weight.data(), m, ctx->stream);
    s->d_xy.Upload(xy.data(), n, ctx->stream);
    ctx->Sync();
    *out = reinterpret_cast<ee_jv_session*>(s.release());
  });
}

void ee_jv_session_destroy(ee_jv_session* session) {
  if (session == nullptr) return;
  auto* s = reinterpret_cast<ee::Session*>(session);
  static_cast<void>(hipSetDevice(s->ctx->device));
  static_cast<void>(hipStreamSynchronize(s->ctx->stream));
  delete s;
}

ee_status ee_jv_session_run(ee_jv_session* session, uint32_t max_rounds, ee_jv_run_stats* stats) {
  return Guarded([&] {
    if (session == nullptr) throw std::invalid_argument("session is NULL");
    reinterpret_cast<ee::Session*>(session)->Run(max_rounds, stats);
  });
}

ee_status ee_jv_session_proof(const ee_jv_session* session, char* buffer, size_t capacity,
                              size_t* length) {
  return Guarded([&] {
    if (session == nullptr) throw std::invalid_argument("session is NULL");
    const std::string& p = reinterpret_cast<const ee::Session*>(session)->proof;
    if (length) *length = p.size();
    if (buffer && capacity > 0) {
      const size_t len = std::min(p.size(), capacity - 1);
      std::memcpy(buffer, p.data(), len);
      buffer[len] = '\0';
    }
  });
}

ee_status ee_jv_session_active(const ee_jv_session* session, uint8_t* active) {
  return Guarded([&] {
    if (session == nullptr || active == nullptr) throw std::invalid_argument("NULL argument");
    const auto& s = *reinterpret_cast<const ee::Session*>(session);
    std::copy(s.active.begin(), s.active.end(), active);
  });
}

}  // extern "C"
