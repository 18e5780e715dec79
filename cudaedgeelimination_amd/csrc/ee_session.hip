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
#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {
namespace {

struct Session {
  explicit Session(Context* c) : ctx(c) {}

  Context* ctx;
  JvHost host;
  std::vector<int32_t> active32, witnesses;
  // device
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  DeviceArray<int32_t> d_active, d_row_offsets, d_neighbors, d_edge_ids, d_witnesses;

  void Upload(double* h2d_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    const int32_t m = host.m;
    active32.resize(static_cast<size_t>(m));
    for (int32_t e = 0; e < m; ++e) active32[e] = host.active[e] ? 1 : 0;
    d_active.Upload(active32.data(), static_cast<size_t>(m), ctx->stream);
    d_row_offsets.Upload(host.row_offsets.data(), host.row_offsets.size(), ctx->stream);
    d_neighbors.Upload(host.neighbors.data(), host.neighbors.size(), ctx->stream);
    d_edge_ids.Upload(host.edge_ids.data(), host.edge_ids.size(), ctx->stream);
    ctx->Sync();
    *h2d_ms += MsSince(t0);
  }

  void Run(uint32_t max_rounds, ee_jv_run_stats* st) {
    if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
    ctx->Activate();
    const auto t_total = std::chrono::steady_clock::now();
    host.StartRun();
    const int32_t m = host.m;
    witnesses.resize(static_cast<size_t>(m));
    for (uint32_t epoch = 0; epoch < max_rounds && !host.done; ++epoch) {
      host.BeginEpoch(epoch);
      const auto t0 = std::chrono::steady_clock::now();
      Upload(&host.stats.h2d_ms);
      JvLaunch a{};
      a.edge_count = m;
      a.distance_type = host.type;
      a.near = host.near ? 1 : 0;
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
      host.stats.device_ms += kernel;
      const auto t_d2h = std::chrono::steady_clock::now();
      d_witnesses.Download(witnesses.data(), static_cast<size_t>(m), ctx->stream);
      ctx->Sync();
      host.stats.d2h_ms += MsSince(t_d2h);
      host.FinishEpoch(witnesses.data(), MsSince(t0));
    }
    host.EndRun(MsSince(t_total));
    if (st) *st = host.stats;
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
    s->host.async_hash = true;  // one graph: hash beside the device step
    s->host.Load(g);
    const ee::JvHost& h = s->host;
    const size_t n = static_cast<size_t>(h.n), m = static_cast<size_t>(h.m);
    std::vector<int2> uv(m);
    std::vector<longlong2> xy(n);
    for (size_t e = 0; e < m; ++e) uv[e] = make_int2(h.u[e], h.v[e]);
    for (size_t i = 0; i < n; ++i) xy[i] = longlong2{h.x[i], h.y[i]};
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
    s->d_weight.Upload(h.weight.data(), m, ctx->stream);
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
    const std::string& p = reinterpret_cast<const ee::Session*>(session)->host.proof;
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
    std::copy(s.host.active.begin(), s.host.active.end(), active);
  });
}

}  // extern "C"
