// Many independent resident JV elimination runs driven in lockstep.
//
// Each graph keeps its own host run (ee_jvhost.h: re-verification, commit,
// proof) exactly as a standalone session does, so its proof equals the
// standalone one (src/cpu/jv_elimination.cpp:250-368 per graph).  The device
// side is shared: the graphs' snapshots sit back to back in one union (node
// ids, edge slots and adjacency offset per graph).  The first epoch uploads
// the union's active flags and CSR; every later epoch uploads only the
// union's commit list (the slots every graph committed, plus every live slot
// of a graph that reached its fixed point, so the kernel skips it), and the
// device clears those flags, compacts the union CSR (ee_csr.h), runs one JV
// launch over all edge slots and downloads the witnesses -- a sequence
// replayable as a captured HIP graph (one per CSR buffer).  Meanwhile the
// host applies each graph's commit to its own snapshot.  After every device
// commit, one device content hash of the union (ee_hash.h, one segment per
// graph) on a second stream gives every graph's new snapshot hash; the next
// commit waits for it.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ee_csr.h"
#include "ee_hash.h"
#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {

void CheckGraphView(const ee_graph_view* g);

namespace {

struct Fleet {
  explicit Fleet(Context* c) : ctx(c) {}
  ~Fleet() {
    DropGraphs();
    if (hash_stream != nullptr) static_cast<void>(hipStreamSynchronize(hash_stream));
    if (ev_flags != nullptr) static_cast<void>(hipEventDestroy(ev_flags));
    if (ev_hashed != nullptr) static_cast<void>(hipEventDestroy(ev_hashed));
    if (hash_stream != nullptr) static_cast<void>(hipStreamDestroy(hash_stream));
  }

  Context* ctx;
  std::vector<JvHost> hosts;
  std::vector<size_t> node_base, edge_base;  // per graph; entry count = totals
  size_t nodes = 0, slots = 0, adjacency = 0;  // adjacency: capacity, 2 x initially active edges
  // device: static union
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  DeviceArray<uint8_t> d_mode;
  DeviceArray<int32_t> d_witnesses;
  // device: active flags and CSR, updated by commits
  DeviceCsr csr;
  Pinned<int32_t> h_active, h_row_offsets, h_neighbors, h_edge_ids, h_witnesses;
  // one captured epoch (commit, launch, download) per CSR buffer
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<uint8_t> retired;  // graph done and its slots deactivated on the device
  // snapshot hashes on the device: slot e holds every graph's hash after the
  // device's e-th commit of the run (slot 0: the uploaded snapshots)
  DeviceContentHash dhash;
  DeviceArray<uint64_t> d_hashes;
  std::vector<uint64_t> h_hashes;
  hipStream_t hash_stream = nullptr;
  hipEvent_t ev_flags = nullptr, ev_hashed = nullptr;
  size_t hash_slots = 0;      // enqueued this run
  bool hash_pending = false;  // new generations of this step take the newest slot
  bool hashes_fetched = false;
  std::mutex hash_mu;

  // The hash of the device's union snapshot into the run's next slot, on the
  // hash stream once the context's stream has the flags in place.
  void EnqueueHash() {
    const size_t n = hosts.size();
    if ((hash_slots + 1) * n > d_hashes.count()) throw std::logic_error("JV fleet: hash slots exhausted");
    CheckHip(hipEventRecord(ev_flags, ctx->stream), "event record");
    CheckHip(hipStreamWaitEvent(hash_stream, ev_flags, 0), "stream wait");
    dhash.Enqueue(d_uv.get(), d_weight.get(), csr.active(), d_hashes.get() + hash_slots * n, hash_stream);
    CheckHip(hipEventRecord(ev_hashed, hash_stream), "event record");
    ++hash_slots;
  }

  // A graph's request for the hash of its new generation (from the host
  // threads of a parallel step): the newest slot, if the device holds it.
  bool HashOfNewGeneration(size_t graph, std::shared_future<uint64_t>* job) {
    if (!hash_pending || hash_slots == 0) return false;
    const size_t slot = hash_slots - 1;
    *job = std::async(std::launch::deferred, [this, slot, graph] { return FetchHash(slot, graph); }).share();
    return true;
  }

  uint64_t FetchHash(size_t slot, size_t graph) {
    std::lock_guard<std::mutex> lk(hash_mu);
    if (!hashes_fetched) {
      ctx->Activate();
      CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
      h_hashes.resize(hash_slots * hosts.size());
      d_hashes.Download(h_hashes.data(), h_hashes.size(), hash_stream);
      CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
      hashes_fetched = true;
    }
    return h_hashes.at(slot * hosts.size() + graph);
  }

  void DropGraphs() {
    for (int p = 0; p < 2; ++p) {
      if (exec[p] != nullptr) static_cast<void>(hipGraphExecDestroy(exec[p]));
      if (graph[p] != nullptr) static_cast<void>(hipGraphDestroy(graph[p]));
      exec[p] = nullptr;
      graph[p] = nullptr;
    }
  }

  void Load(const ee_graph_view* graphs, uint32_t count) {
    hosts.resize(count);
    node_base.assign(count + 1, 0);
    edge_base.assign(count + 1, 0);
    size_t adj = 0;
    for (uint32_t i = 0; i < count; ++i) CheckGraphView(&graphs[i]);
    ParallelFor(count, [&](size_t i) { hosts[i].Load(&graphs[i]); });
    for (uint32_t i = 0; i < count; ++i) {
      node_base[i + 1] = node_base[i] + static_cast<size_t>(hosts[i].n);
      edge_base[i + 1] = edge_base[i] + static_cast<size_t>(hosts[i].m);
      adj += hosts[i].neighbors.size();
    }
    nodes = node_base[count];
    slots = edge_base[count];
    adjacency = std::max<size_t>(adj, 1);
    if (nodes >= (size_t{1} << 31) || slots >= (size_t{1} << 31) - 1 || adjacency >= (size_t{1} << 31)) {
      throw std::invalid_argument("JV fleet union exceeds int32 ids");
    }
    std::vector<int2> uv(std::max<size_t>(slots, 1));
    std::vector<int64_t> weight(std::max<size_t>(slots, 1));
    std::vector<uint8_t> mode(std::max<size_t>(slots, 1));
    std::vector<longlong2> xy(std::max<size_t>(nodes, 1));
    ParallelFor(count, [&](size_t i) {
      const JvHost& h = hosts[i];
      const int32_t nb = static_cast<int32_t>(node_base[i]);
      const uint8_t md = static_cast<uint8_t>((h.type & 1u) | (h.near ? 2u : 0u));
      for (int32_t e = 0; e < h.m; ++e) {
        const size_t s = edge_base[i] + static_cast<size_t>(e);
        uv[s] = make_int2(nb + h.u[e], nb + h.v[e]);
        weight[s] = h.weight[e];
        mode[s] = md;
      }
      for (int32_t vtx = 0; vtx < h.n; ++vtx) xy[node_base[i] + vtx] = longlong2{h.x[vtx], h.y[vtx]};
    });
    const int dev = ctx->device;
    d_uv.Allocate(uv.size(), dev);
    d_weight.Allocate(weight.size(), dev);
    d_mode.Allocate(mode.size(), dev);
    d_xy.Allocate(xy.size(), dev);
    d_witnesses.Allocate(uv.size(), dev);
    // a commit deactivates each slot at most once
    csr.Allocate(nodes, slots, adjacency, slots, dev);
    h_active.Reserve(uv.size());
    h_witnesses.Reserve(uv.size());
    h_row_offsets.Reserve(nodes + 1);
    h_neighbors.Reserve(adjacency);
    h_edge_ids.Reserve(adjacency);
    d_uv.Upload(uv.data(), uv.size(), ctx->stream);
    d_weight.Upload(weight.data(), weight.size(), ctx->stream);
    d_mode.Upload(mode.data(), mode.size(), ctx->stream);
    d_xy.Upload(xy.data(), xy.size(), ctx->stream);
    ctx->Sync();
    // one hash segment per graph, each continuing from its own prefix
    std::vector<int32_t> node_offset(count);
    for (uint32_t i = 0; i < count; ++i) node_offset[i] = static_cast<int32_t>(node_base[i]);
    dhash.Allocate(edge_base, dev, ctx->stream, node_offset);
    std::vector<uint64_t> prefixes(count);
    for (uint32_t i = 0; i < count; ++i) prefixes[i] = hosts[i].prefix_hash();
    dhash.SetPrefixes(prefixes.data(), ctx->stream);
    CheckHip(hipStreamCreateWithFlags(&hash_stream, hipStreamNonBlocking), "hipStreamCreate");
    CheckHip(hipEventCreateWithFlags(&ev_flags, hipEventDisableTiming), "hipEventCreate");
    CheckHip(hipEventCreateWithFlags(&ev_hashed, hipEventDisableTiming), "hipEventCreate");
    for (uint32_t i = 0; i < count; ++i) {
      hosts[i].async_hash = true;  // a generation the device does not hash: on a helper thread
      hosts[i].hash_source = [this, i](std::shared_future<uint64_t>* job) { return HashOfNewGeneration(i, job); };
    }
  }

  // The union's active flags and CSR from the host snapshots (first epoch).
  size_t AssembleFull() {
    std::vector<size_t> adj_base(hosts.size() + 1, 0);
    for (size_t i = 0; i < hosts.size(); ++i) adj_base[i + 1] = adj_base[i] + hosts[i].neighbors.size();
    const size_t adj = adj_base[hosts.size()];
    if (adj > adjacency) throw std::logic_error("JV fleet adjacency grew");
    ParallelFor(hosts.size(), [&](size_t i) {
      const JvHost& h = hosts[i];
      const size_t eb = edge_base[i], nb = node_base[i], ab = adj_base[i];
      for (int32_t e = 0; e < h.m; ++e) h_active.get()[eb + e] = h.active[e] ? 1 : 0;
      for (int32_t vtx = 0; vtx < h.n; ++vtx) {
        h_row_offsets.get()[nb + vtx] = static_cast<int32_t>(ab) + h.row_offsets[vtx];
      }
      for (size_t s = 0; s < h.neighbors.size(); ++s) {
        h_neighbors.get()[ab + s] = static_cast<int32_t>(nb) + h.neighbors[s];
        h_edge_ids.get()[ab + s] = static_cast<int32_t>(eb) + h.edge_ids[s];
      }
    });
    h_row_offsets.get()[nodes] = static_cast<int32_t>(adj);
    return adj;
  }

  // The union commit list of this epoch into the CSR's pinned staging: each
  // graph's committed slots, and every live slot of a graph that stopped.
  void StageCommits(const std::vector<const std::vector<int32_t>*>& commits) {
    std::vector<size_t> at(hosts.size() + 1, 0);
    std::vector<uint8_t> retire(hosts.size(), 0);
    for (size_t i = 0; i < hosts.size(); ++i) {
      size_t c = 0;
      if (!retired[i]) {
        if (commits[i] != nullptr && !commits[i]->empty()) {
          c = commits[i]->size();
        } else {
          retire[i] = 1;
          for (int32_t e = 0; e < hosts[i].m; ++e) c += hosts[i].active[e] != 0;
        }
      }
      at[i + 1] = at[i] + c;
    }
    int32_t* staged = csr.StageCommits();
    staged[0] = static_cast<int32_t>(at[hosts.size()]);
    ParallelFor(hosts.size(), [&](size_t i) {
      int32_t* out = staged + 1 + at[i];
      const int32_t eb = static_cast<int32_t>(edge_base[i]);
      if (retire[i]) {
        const JvHost& h = hosts[i];
        for (int32_t e = 0; e < h.m; ++e) {
          if (h.active[e]) *out++ = eb + e;
        }
      } else if (!retired[i] && commits[i] != nullptr) {
        for (const int32_t e : *commits[i]) *out++ = eb + e;
      }
    });
    for (size_t i = 0; i < hosts.size(); ++i) retired[i] = retired[i] || retire[i];
  }

  JvLaunch Args() {
    JvLaunch a{};
    a.edge_count = static_cast<int32_t>(slots);
    a.distance_type = 0;
    a.edge_uv = d_uv.get();
    a.edge_weight = d_weight.get();
    a.edge_active = csr.active();
    a.row_offsets = csr.row_offsets();
    a.neighbors = csr.neighbors();
    a.csr_edge_ids = csr.edge_ids();
    a.xy = d_xy.get();
    a.witnesses = d_witnesses.get();
    a.edge_mode = d_mode.get();
    a.one_per_wave = ctx->jv_edges_per_wave == 1 ? 1 : 0;
    return a;
  }

  // Launch and download (after a commit, which selects the CSR buffer).
  void EnqueueLaunch() {
    LaunchJv(Args(), ctx->jv_grid, ctx->stream);
    CheckHip(hipMemcpyAsync(h_witnesses.get(), d_witnesses.get(), std::max<size_t>(slots, 1) * 4,
                            hipMemcpyDeviceToHost, ctx->stream),
             "fleet download");
  }

  // The device side of one epoch; the first uploads the union, later ones
  // commit the staged list on the device.
  void EnqueueEpoch(bool first, bool use_graph, ee_jv_fleet_stats* st) {
    if (first) {
      const size_t adj = AssembleFull();
      csr.Upload(h_active.get(), h_row_offsets.get(), h_neighbors.get(), h_edge_ids.get(), adj, ctx->stream);
      CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
      EnqueueLaunch();
      return;
    }
    // the last snapshot's hash reads the flags this commit changes
    CheckHip(hipStreamWaitEvent(ctx->stream, ev_hashed, 0), "stream wait");
    csr.UploadCommits(ctx->stream);
    CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
    if (!use_graph) {
      csr.EnqueueCompact(ctx->stream);
      EnqueueLaunch();
      return;
    }
    const int p = csr.current();
    if (exec[p] == nullptr) {
      // the captured compaction reads the commit count from device memory,
      // so the graph serves every later commit of this buffer parity
      CheckHip(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      csr.EnqueueCompact(ctx->stream);
      EnqueueLaunch();
      CheckHip(hipStreamEndCapture(ctx->stream, &graph[p]), "end capture");
      CheckHip(hipGraphInstantiate(&exec[p], graph[p], nullptr, nullptr, 0), "graph instantiate");
    } else {
      csr.SetCurrent(p ^ 1);  // what the replayed compaction does
    }
    CheckHip(hipGraphLaunch(exec[p], ctx->stream), "graph launch");
    ++st->graph_launches;
  }

  void Run(uint32_t max_rounds, bool use_graph, ee_jv_fleet_stats* out) {
    if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
    ctx->Activate();
    const auto t_total = std::chrono::steady_clock::now();
    ee_jv_fleet_stats st{};
    st.instances = static_cast<uint32_t>(hosts.size());
    st.edge_slots = slots;
    retired.assign(hosts.size(), 0);
    // the graphs' host runs are independent: each phase runs over the host
    // cores; the previous run's hashes are settled first
    ParallelFor(hosts.size(), [&](size_t i) { hosts[i].StartRun(); });
    CheckHip(hipStreamSynchronize(hash_stream), "hipStreamSynchronize");
    hash_slots = 0;
    hashes_fetched = false;
    size_t max_m = 0;
    for (const JvHost& h : hosts) max_m = std::max(max_m, static_cast<size_t>(h.m));
    const size_t hash_capacity = (std::min<size_t>(max_rounds, max_m) + 1) * hosts.size();
    if (d_hashes.count() < hash_capacity) d_hashes.Allocate(hash_capacity, ctx->device);
    std::vector<const std::vector<int32_t>*> commits(hosts.size(), nullptr);
    auto t0 = std::chrono::steady_clock::now();
    EnqueueEpoch(true, use_graph, &st);
    EnqueueHash();
    st.assemble_ms += MsSince(t0);
    hash_pending = true;  // the uploaded snapshots are generation 0 of every graph
    ParallelFor(hosts.size(), [&](size_t i) { hosts[i].BeginEpoch(0); });
    hash_pending = false;
    for (uint32_t epoch = 0;; ++epoch) {
      CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
      CheckHip(hipEventSynchronize(ctx->ev_end), "event synchronize");
      float ms = 0.0f;
      CheckHip(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "event elapsed");
      st.device_ms += ms;
      const double propose_ms = MsSince(t0);
      st.propose_ms += propose_ms;
      ++st.epochs;
      auto t1 = std::chrono::steady_clock::now();
      ParallelFor(hosts.size(), [&](size_t i) {
        JvHost& h = hosts[i];
        commits[i] = nullptr;
        if (h.done) return;
        // witnesses are union node ids: back to the graph's own ids
        int32_t* w = h_witnesses.get() + edge_base[i];
        const int32_t nb = static_cast<int32_t>(node_base[i]);
        for (int32_t e = 0; e < h.m; ++e) {
          if (w[e] >= 0) w[e] -= nb;
        }
        commits[i] = &h.VerifyAndGate(w, propose_ms);
      });
      bool live = false;
      for (size_t i = 0; i < hosts.size(); ++i) live = live || (commits[i] != nullptr && !commits[i]->empty());
      const bool next = live && epoch + 1 < max_rounds;
      st.host_ms += MsSince(t1);
      if (next) {
        // the device commits and proposes while the hosts apply the commits
        t0 = std::chrono::steady_clock::now();
        StageCommits(commits);
        EnqueueEpoch(false, use_graph, &st);
        EnqueueHash();
        st.assemble_ms += MsSince(t0);
      }
      t1 = std::chrono::steady_clock::now();
      ParallelFor(hosts.size(), [&](size_t i) {
        if (commits[i] != nullptr) hosts[i].ApplyCommit();
      });
      if (!next) {
        st.host_ms += MsSince(t1);
        break;
      }
      hash_pending = true;  // a graph that committed takes the newest slot
      ParallelFor(hosts.size(), [&](size_t i) {
        if (!hosts[i].done) hosts[i].BeginEpoch(epoch + 1);
      });
      hash_pending = false;
      st.host_ms += MsSince(t1);
    }
    const double loop_ms = MsSince(t_total);
    ParallelFor(hosts.size(), [&](size_t i) { hosts[i].EndRun(loop_ms); });
    const double total = MsSince(t_total);  // including the wait for hashes
    for (const JvHost& h : hosts) {
      st.edges_scanned += h.stats.edges_scanned;
      st.proposed += h.stats.proposed;
      st.committed += h.stats.committed;
    }
    st.total_ms = total;
    if (out) *out = st;
  }
};

Fleet& AsFleet(const ee_jv_fleet* f) {
  if (f == nullptr) throw std::invalid_argument("fleet is NULL");
  return *reinterpret_cast<Fleet*>(const_cast<ee_jv_fleet*>(f));
}

const JvHost& Instance(const ee_jv_fleet* f, uint32_t index) {
  const Fleet& fl = AsFleet(f);
  if (index >= fl.hosts.size()) throw std::invalid_argument("fleet instance index out of range");
  return fl.hosts[index];
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

extern "C" {

ee_status ee_jv_fleet_create(ee_ctx* handle, const ee_graph_view* graphs, uint32_t count,
                             ee_jv_fleet** out) {
  return Guarded([&] {
    if (handle == nullptr || out == nullptr || (count > 0 && graphs == nullptr)) {
      throw std::invalid_argument("NULL argument");
    }
    *out = nullptr;
    if (count == 0) throw std::invalid_argument("a fleet needs at least one graph");
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    auto f = std::make_unique<ee::Fleet>(ctx);
    f->Load(graphs, count);
    *out = reinterpret_cast<ee_jv_fleet*>(f.release());
  });
}

void ee_jv_fleet_destroy(ee_jv_fleet* fleet) {
  if (fleet == nullptr) return;
  auto* f = reinterpret_cast<ee::Fleet*>(fleet);
  static_cast<void>(hipSetDevice(f->ctx->device));
  static_cast<void>(hipStreamSynchronize(f->ctx->stream));
  delete f;
}

ee_status ee_jv_fleet_run(ee_jv_fleet* fleet, uint32_t max_rounds, int use_graph,
                          ee_jv_fleet_stats* stats) {
  return Guarded([&] { ee::AsFleet(fleet).Run(max_rounds, use_graph != 0, stats); });
}

ee_status ee_jv_fleet_instance_stats(const ee_jv_fleet* fleet, uint32_t index, ee_jv_run_stats* stats) {
  return Guarded([&] {
    if (stats == nullptr) throw std::invalid_argument("stats is NULL");
    *stats = ee::Instance(fleet, index).stats;
  });
}

ee_status ee_jv_fleet_proof(const ee_jv_fleet* fleet, uint32_t index, char* buffer, size_t capacity,
                            size_t* length) {
  return Guarded([&] {
    const std::string& p = ee::Instance(fleet, index).proof;
    if (length) *length = p.size();
    if (buffer && capacity > 0) {
      const size_t len = std::min(p.size(), capacity - 1);
      std::memcpy(buffer, p.data(), len);
      buffer[len] = '\0';
    }
  });
}

ee_status ee_jv_fleet_active(const ee_jv_fleet* fleet, uint32_t index, uint8_t* active) {
  return Guarded([&] {
    if (active == nullptr) throw std::invalid_argument("active is NULL");
    const ee::JvHost& h = ee::Instance(fleet, index);
    std::copy(h.active.begin(), h.active.end(), active);
  });
}

}  // extern "C"
