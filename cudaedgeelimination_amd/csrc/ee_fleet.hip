// Many independent resident JV elimination runs driven in lockstep.
//
// Each graph keeps its own host run (ee_jvhost.h: re-verification, commit,
// proof) exactly as a standalone session does, so its proof equals the
// standalone one (src/cpu/jv_elimination.cpp:250-368 per graph).  The device
// side is shared: the graphs' snapshots sit back to back in one union (node
// ids, edge slots and adjacency offset per graph), and every epoch is one
// upload of the union's active flags and CSR, one JV launch over all edge
// slots and one download of the witnesses -- replayable as a captured HIP
// graph whose copies read fixed-capacity pinned staging.  A graph that has
// reached its fixed point has its slots marked inactive, so the kernel skips
// them, and its run is not advanced again.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {

void CheckGraphView(const ee_graph_view* g);

namespace {

struct Fleet {
  explicit Fleet(Context* c) : ctx(c) {}
  ~Fleet() { DropGraph(); }

  Context* ctx;
  std::vector<JvHost> hosts;
  std::vector<size_t> node_base, edge_base;  // per graph; entry count = totals
  size_t nodes = 0, slots = 0, adjacency = 0;  // adjacency: capacity, 2 x initially active edges
  // device: static union
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  DeviceArray<uint8_t> d_mode;
  // device: per epoch
  DeviceArray<int32_t> d_active, d_row_offsets, d_neighbors, d_edge_ids, d_witnesses;
  Pinned<int32_t> h_active, h_row_offsets, h_neighbors, h_edge_ids, h_witnesses;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void DropGraph() {
    if (exec != nullptr) static_cast<void>(hipGraphExecDestroy(exec));
    if (graph != nullptr) static_cast<void>(hipGraphDestroy(graph));
    exec = nullptr;
    graph = nullptr;
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
    if (nodes >= (size_t{1} << 31) || slots >= (size_t{1} << 31) || adjacency >= (size_t{1} << 31)) {
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
    d_active.Allocate(uv.size(), dev);
    d_witnesses.Allocate(uv.size(), dev);
    d_row_offsets.Allocate(nodes + 1, dev);
    d_neighbors.Allocate(adjacency, dev);
    d_edge_ids.Allocate(adjacency, dev);
    h_active.Reserve(uv.size());
    h_witnesses.Reserve(uv.size());
    h_row_offsets.Reserve(nodes + 1);
    h_neighbors.Reserve(adjacency);
    h_edge_ids.Reserve(adjacency);
    std::fill(h_neighbors.get(), h_neighbors.get() + adjacency, 0);
    std::fill(h_edge_ids.get(), h_edge_ids.get() + adjacency, 0);
    d_uv.Upload(uv.data(), uv.size(), ctx->stream);
    d_weight.Upload(weight.data(), weight.size(), ctx->stream);
    d_mode.Upload(mode.data(), mode.size(), ctx->stream);
    d_xy.Upload(xy.data(), xy.size(), ctx->stream);
    ctx->Sync();
  }

  // Writes the union's active flags and CSR into the pinned staging; graphs
  // whose run is done contribute inactive slots.
  void Assemble() {
    std::vector<size_t> adj_base(hosts.size() + 1, 0);
    for (size_t i = 0; i < hosts.size(); ++i) adj_base[i + 1] = adj_base[i] + hosts[i].neighbors.size();
    const size_t adj = adj_base[hosts.size()];
    if (adj > adjacency) throw std::logic_error("JV fleet adjacency grew");
    ParallelFor(hosts.size(), [&](size_t i) {
      const JvHost& h = hosts[i];
      const size_t eb = edge_base[i], nb = node_base[i], ab = adj_base[i];
      for (int32_t e = 0; e < h.m; ++e) h_active.get()[eb + e] = (!h.done && h.active[e]) ? 1 : 0;
      for (int32_t vtx = 0; vtx < h.n; ++vtx) {
        h_row_offsets.get()[nb + vtx] = static_cast<int32_t>(ab) + h.row_offsets[vtx];
      }
      for (size_t s = 0; s < h.neighbors.size(); ++s) {
        h_neighbors.get()[ab + s] = static_cast<int32_t>(nb) + h.neighbors[s];
        h_edge_ids.get()[ab + s] = static_cast<int32_t>(eb) + h.edge_ids[s];
      }
    });
    h_row_offsets.get()[nodes] = static_cast<int32_t>(adj);
  }

  JvLaunch Args() {
    JvLaunch a{};
    a.edge_count = static_cast<int32_t>(slots);
    a.distance_type = 0;
    a.edge_uv = d_uv.get();
    a.edge_weight = d_weight.get();
    a.edge_active = d_active.get();
    a.row_offsets = d_row_offsets.get();
    a.neighbors = d_neighbors.get();
    a.csr_edge_ids = d_edge_ids.get();
    a.xy = d_xy.get();
    a.witnesses = d_witnesses.get();
    a.edge_mode = d_mode.get();
    return a;
  }

  // Upload, launch, download: the whole device side of one epoch.
  void Enqueue() {
    const size_t m = std::max<size_t>(slots, 1);
    CheckHip(hipMemcpyAsync(d_active.get(), h_active.get(), m * 4, hipMemcpyHostToDevice, ctx->stream),
             "fleet upload");
    CheckHip(hipMemcpyAsync(d_row_offsets.get(), h_row_offsets.get(), (nodes + 1) * 4,
                            hipMemcpyHostToDevice, ctx->stream),
             "fleet upload");
    CheckHip(hipMemcpyAsync(d_neighbors.get(), h_neighbors.get(), adjacency * 4, hipMemcpyHostToDevice,
                            ctx->stream),
             "fleet upload");
    CheckHip(hipMemcpyAsync(d_edge_ids.get(), h_edge_ids.get(), adjacency * 4, hipMemcpyHostToDevice,
                            ctx->stream),
             "fleet upload");
    LaunchJv(Args(), ctx->jv_grid, ctx->stream);
    CheckHip(hipMemcpyAsync(h_witnesses.get(), d_witnesses.get(), m * 4, hipMemcpyDeviceToHost,
                            ctx->stream),
             "fleet download");
  }

  void Run(uint32_t max_rounds, bool use_graph, ee_jv_fleet_stats* out) {
    if (max_rounds == 0) throw std::invalid_argument("max_rounds must be > 0");
    ctx->Activate();
    const auto t_total = std::chrono::steady_clock::now();
    ee_jv_fleet_stats st{};
    st.instances = static_cast<uint32_t>(hosts.size());
    st.edge_slots = slots;
    // the graphs' host runs are independent: each phase runs over the host cores
    ParallelFor(hosts.size(), [&](size_t i) { hosts[i].StartRun(); });
    for (uint32_t epoch = 0; epoch < max_rounds; ++epoch) {
      bool live = false;
      for (const JvHost& h : hosts) live = live || !h.done;
      if (!live) break;
      ParallelFor(hosts.size(), [&](size_t i) {
        if (!hosts[i].done) hosts[i].BeginEpoch(epoch);
      });
      const auto t0 = std::chrono::steady_clock::now();
      Assemble();
      st.assemble_ms += MsSince(t0);
      if (slots > 0) {
        if (use_graph && exec == nullptr) {
          CheckHip(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
          Enqueue();
          CheckHip(hipStreamEndCapture(ctx->stream, &graph), "end capture");
          CheckHip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        }
        CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
        if (use_graph) {
          CheckHip(hipGraphLaunch(exec, ctx->stream), "graph launch");
          ++st.graph_launches;
        } else {
          Enqueue();
        }
        CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
        CheckHip(hipEventSynchronize(ctx->ev_end), "event synchronize");
        float ms = 0.0f;
        CheckHip(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "event elapsed");
        st.device_ms += ms;
      }
      const double propose_ms = MsSince(t0);
      st.propose_ms += propose_ms;
      const auto t1 = std::chrono::steady_clock::now();
      ParallelFor(hosts.size(), [&](size_t i) {
        JvHost& h = hosts[i];
        if (h.done) return;
        // witnesses are union node ids: back to the graph's own ids
        int32_t* w = h_witnesses.get() + edge_base[i];
        const int32_t nb = static_cast<int32_t>(node_base[i]);
        for (int32_t e = 0; e < h.m; ++e) {
          if (w[e] >= 0) w[e] -= nb;
        }
        h.FinishEpoch(w, propose_ms);
      });
      st.host_ms += MsSince(t1);
      ++st.epochs;
    }
    const double total = MsSince(t_total);
    ParallelFor(hosts.size(), [&](size_t i) { hosts[i].EndRun(total); });
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
