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
// next re-verification).  Each snapshot's content hash is computed on the
// device (ee_hash.h) on a second stream, beside the proposal, from the
// device's active flags; the next commit waits for it.  The static part of
// the snapshot, the active flags and the CSR stay on the device; only the
// first epoch of a run uploads them.
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
#include "ee_hash.h"
#include "ee_gen.h"
#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {
namespace {

struct Session {
  explicit Session(Context* c) : ctx(c) {}
  ~Session() {
    if (hash_stream != nullptr) static_cast<void>(hipStreamSynchronize(hash_stream));
    if (ev_compact != nullptr) static_cast<void>(hipEventDestroy(ev_compact));
    if (ev_flags != nullptr) static_cast<void>(hipEventDestroy(ev_flags));
    if (ev_hashed != nullptr) static_cast<void>(hipEventDestroy(ev_hashed));
    if (hash_stream != nullptr) static_cast<void>(hipStreamDestroy(hash_stream));
  }

  Context* ctx;
  JvHost host;
  std::vector<int32_t> active32;
  Pinned<int32_t> witnesses;
  hipEvent_t ev_compact = nullptr;
  // snapshot hashes on the device: one slot per generation of a run
  DeviceContentHash dhash;
  DeviceArray<uint64_t> d_hashes;
  std::vector<uint64_t> h_hashes;
  hipStream_t hash_stream = nullptr;
  hipEvent_t ev_flags = nullptr, ev_hashed = nullptr;
  // a fault for tests of the fail-closed path: one witness of the next
  // epoch's host copy replaced before re-verification
  bool inject = false;
  int32_t inject_edge = 0, inject_witness = -1;
  size_t hash_slots = 0;      // enqueued this run
  bool hash_pending = false;  // the newest slot is not yet handed to the host
  bool hashes_fetched = false;
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

  // The hash of the device's current snapshot into the run's next slot, on
  // the hash stream once the context's stream has the flags in place.
  void EnqueueHash() {
    if (hash_slots >= d_hashes.count()) throw std::logic_error("JV session: hash slots exhausted");
    CheckHip(hipEventRecord(ev_flags, ctx->stream), "event record");
    CheckHip(hipStreamWaitEvent(hash_stream, ev_flags, 0), "stream wait");
    dhash.Enqueue(d_uv.get(), d_weight.get(), csr.active(), d_hashes.get() + hash_slots, hash_stream);
    CheckHip(hipEventRecord(ev_hashed, hash_stream), "event record");
    ++hash_slots;
    hash_pending = true;
  }

  // The host's request for the hash of a new generation: the slot enqueued
  // for it, read when the proof is written.
  bool HashOfNewGeneration(std::shared_future<uint64_t>* job) {
    if (!hash_pending) return false;  // a generation the device has not hashed
    hash_pending = false;
    const size_t slot = hash_slots - 1;
    *job = std::async(std::launch::deferred, [this, slot] { return FetchHash(slot); }).share();
    return true;
  }

  uint64_t FetchHash(size_t slot) {
    if (!hashes_fetched) {
      CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
      h_hashes.resize(hash_slots);
      d_hashes.Download(h_hashes.data(), hash_slots, hash_stream);
      CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
      hashes_fetched = true;
    }
    return h_hashes.at(slot);
  }

  // The device side of one epoch on the context's stream: [commit of the
  // previous epoch and the hash of its result,] JV kernel, witness download.
  // Events bracket the commit and the kernel.
  void Enqueue(bool compact) {
    if (compact) {
      // the previous snapshot's hash reads the flags this commit changes
      CheckHip(hipStreamWaitEvent(ctx->stream, ev_hashed, 0), "stream wait");
      csr.UploadCommits(ctx->stream);
      CheckHip(hipEventRecord(ev_compact, ctx->stream), "event record");
      csr.EnqueueCompact(ctx->stream);
      EnqueueHash();
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
    host.StartRun();  // settles the previous run's hashes first
    // a run that stopped on an error may have left hashes in flight
    if (hash_stream != nullptr) CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
    hash_slots = 0;
    hash_pending = false;
    hashes_fetched = false;
    if (host.m == 0) {  // nothing to propose: one empty epoch
      host.BeginEpoch(0);
      std::vector<int32_t> none(1, -1);
      host.FinishEpoch(none.data(), 0.0);
    } else {
      // one hash per generation: the first, and one after each commit
      const size_t slots = static_cast<size_t>(std::min<uint64_t>(max_rounds, static_cast<uint64_t>(host.m))) + 1;
      if (d_hashes.count() < slots) d_hashes.Allocate(slots, ctx->device);
      // the first epoch uploads the host's snapshot (a run may follow another)
      UploadFull(&host.stats.h2d_ms);
      EnqueueHash();
      host.BeginEpoch(0);
      auto t0 = std::chrono::steady_clock::now();
      Enqueue(false);
      bool compacted = false;
      for (uint32_t epoch = 0;; ++epoch) {
        Wait(compacted);
        if (inject) {
          inject = false;
          if (inject_edge >= 0 && inject_edge < host.m) witnesses.get()[inject_edge] = inject_witness;
        }
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
    s->host.async_hash = true;  // a generation the device does not hash: on a helper thread
    s->host.verify_threads = 0;  // one graph: a large epoch's re-verification over the host cores
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
    if (m > 0) {
      s->dhash.Allocate({0, m}, ctx->device, ctx->stream);
      const uint64_t prefix = s->host.prefix_hash();
      s->dhash.SetPrefixes(&prefix, ctx->stream);
      ee::CheckHip(hipStreamCreateWithFlags(&s->hash_stream, hipStreamNonBlocking), "hipStreamCreate");
      ee::CheckHip(hipEventCreateWithFlags(&s->ev_flags, hipEventDisableTiming), "hipEventCreate");
      ee::CheckHip(hipEventCreateWithFlags(&s->ev_hashed, hipEventDisableTiming), "hipEventCreate");
      ee::Session* raw = s.get();
      s->host.hash_source = [raw](std::shared_future<uint64_t>* job) { return raw->HashOfNewGeneration(job); };
    }
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

ee_status ee_jv_session_inject_witness(ee_jv_session* session, int32_t edge, int32_t witness) {
  return Guarded([&] {
    if (session == nullptr) throw std::invalid_argument("session is NULL");
    auto* s = reinterpret_cast<ee::Session*>(session);
    s->inject = true;
    s->inject_edge = edge;
    s->inject_witness = witness;
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
