// Resident multi-epoch JV elimination: the epoch loop of RunJvElimination
// (src/cpu/jv_elimination.cpp:250-322) with the propose step and the CSR
// refresh of the commit on the device.
//
// Per epoch: device proposal over every edge slot, host re-verification of
// every candidate (jv_elimination.cpp:226-248; a rejected device candidate
// stops the run before anything is committed, as the reference does at
// :301-303) and the deterministic commit rule (src/cpu/elimination_commit.cpp:
// 13-75).  The committed edge slots go to the device, which clears their
// active flags and compacts its CSR (ee_csr.h) and proposes the next epoch,
// while the host applies the same commit to its own snapshot (needed by the
// next re-verification) and the snapshot hashes run on a helper thread.  The
// static part of the snapshot, the active flags and the CSR stay on the
// device; only the first epoch of a run uploads them.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ee_csr.h"
#include "ee_exact.h"
#include "ee_gen.h"
#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {
namespace {

struct Session {
  explicit Session(Context* c) : ctx(c) {}
  ~Session() {
    if (ev_compact != nullptr) static_cast<void>(hipEventDestroy(ev_compact));
  }

  Context* ctx;
  JvHost host;
  std::vector<int32_t> active32;
  Pinned<int32_t> witnesses;
  hipEvent_t ev_compact = nullptr;
  // device
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  DeviceArray<int32_t> d_witnesses;
  DeviceCsr csr;

  void UploadFull(double* h2d_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    const int32_t m = host.m;
    active32.resize(static_cast<size_t>(m));
    for (int32_t e = 0; e < m; ++e) active32[e] = host.active[e] ? 1 : 0;
    csr.Upload(active32.data(), host.row_offsets.data(), host.neighbors.data(), host.edge_ids.data(),
               host.neighbors.size(), ctx->stream);
    ctx->Sync();
    *h2d_ms += MsSince(t0);
  }

  // The device side of one epoch on the context's stream: [commit of the
  // previous epoch,] JV kernel, witness download.  Events bracket the commit
  // and the kernel.
  void Enqueue(bool compact) {
    if (compact) {
      csr.UploadCommits(ctx->stream);
      CheckHip(hipEventRecord(ev_compact, ctx->stream), "event record");
      csr.EnqueueCompact(ctx->stream);
    }
    JvLaunch a{};
    a.edge_count = host.m;
    a.distance_type = host.type;
    a.near = host.near ? 1 : 0;
    a.edge_uv = d_uv.get();
    a.edge_weight = d_weight.get();
    a.edge_active = csr.active();
    a.row_offsets = csr.row_offsets();
    a.neighbors = csr.neighbors();
    a.csr_edge_ids = csr.edge_ids();
    a.xy = d_xy.get();
    a.witnesses = d_witnesses.get();
    a.one_per_wave = ctx->jv_edges_per_wave == 1 ? 1 : 0;
    CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
    LaunchJv(a, ctx->jv_grid, ctx->stream);
    CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
    d_witnesses.Download(witnesses.get(), static_cast<size_t>(host.m), ctx->stream);
  }

  void Wait(bool compacted) {
    ctx->Sync();
    float kernel = 0.0f;
    CheckHip(hipEventElapsedTime(&kernel, ctx->ev_begin, ctx->ev_end), "event elapsed");
    host.stats.device_ms += kernel;
    if (compacted) {
      float compact = 0.0f;
      CheckHip(hipEventElapsedTime(&compact, ev_compact, ctx->ev_begin), "event elapsed");
      host.stats.compact_ms += compact;
    }
  }

  void Run(uint32_t max_rounds, ee_jv_run_stats* st) {
    if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
    ctx->Activate();
    const auto t_total = std::chrono::steady_clock::now();
    host.StartRun();
    host.BeginEpoch(0);
    if (host.m == 0) {  // nothing to propose: one empty epoch
      std::vector<int32_t> none(1, -1);
      host.FinishEpoch(none.data(), 0.0);
    } else {
      // the first epoch uploads the host's snapshot (a run may follow another)
      UploadFull(&host.stats.h2d_ms);
      auto t0 = std::chrono::steady_clock::now();
      Enqueue(false);
      bool compacted = false;
      for (uint32_t epoch = 0;; ++epoch) {
        Wait(compacted);
        const std::vector<int32_t>& commit = host.VerifyAndGate(witnesses.get(), MsSince(t0));
        const bool next = !commit.empty() && epoch + 1 < max_rounds;
        if (next) {
          // the device commits and proposes the next epoch while the host
          // applies the same commit to its snapshot
          int32_t* staged = csr.StageCommits();
          staged[0] = static_cast<int32_t>(commit.size());
          std::copy(commit.begin(), commit.end(), staged + 1);
          t0 = std::chrono::steady_clock::now();
          Enqueue(true);
          compacted = true;
        }
        host.ApplyCommit();
        if (!next) break;
        host.BeginEpoch(epoch + 1);
      }
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
    s->d_uv.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_weight.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->d_xy.Allocate(n, ctx->device);
    s->d_witnesses.Allocate(std::max<size_t>(m, 1), ctx->device);
    s->witnesses.Reserve(std::max<size_t>(m, 1));
    // at most every slot active (2 CSR entries each), at most every slot per commit
    s->csr.Allocate(n, m, 2 * m, m, ctx->device);
    ee::CheckHip(hipEventCreate(&s->ev_compact), "hipEventCreate");
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

ee_status ee_jv_session_device_snapshot(const ee_jv_session* session, uint8_t* active,
                                        int32_t* row_offsets, int32_t* neighbors, int32_t* edge_ids,
                                        size_t capacity) {
  return Guarded([&] {
    if (session == nullptr || active == nullptr || row_offsets == nullptr ||
        (capacity > 0 && (neighbors == nullptr || edge_ids == nullptr))) {
      throw std::invalid_argument("NULL argument");
    }
    auto& s = *const_cast<ee::Session*>(reinterpret_cast<const ee::Session*>(session));
    s.ctx->Activate();
    s.ctx->Sync();
    std::vector<int32_t> flags(static_cast<size_t>(std::max(s.host.m, 1)));
    if (s.host.m > 0) s.csr.DownloadActive(flags.data(), s.ctx->stream);
    for (int32_t e = 0; e < s.host.m; ++e) active[e] = flags[static_cast<size_t>(e)] != 0 ? 1 : 0;
    s.csr.Download(row_offsets, neighbors, edge_ids, capacity, s.ctx->stream);
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
