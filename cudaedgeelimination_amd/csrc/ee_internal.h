// Internal declarations shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "ee_hip.h"

namespace ee {

// --------------------------------------------------------------- errors
// Thrown types map onto ee_status codes at the C boundary.
struct Unavailable : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void CheckHip(hipError_t status, const char* what);

// ------------------------------------------------------------ k-opt plan
// Per-k reconnect plan derived from a template table: the distinct port pairs
// (p < q) the templates use, and for every template its k pair indices in
// ascending port order (the order the reference visits added edges,
// src/cuda/kopt_cost.cu:370-377).
constexpr int kMaxK = 5;
constexpr int kMaxPorts = 2 * kMaxK;
constexpr int kMaxPairs = kMaxPorts * (kMaxPorts - 1) / 2;  // 45
constexpr int kMaxTemplates = 945;                           // (2k-1)!! for k = 5

struct KoptPlan {
  uint32_t k = 0;
  uint32_t templates = 0;
  uint32_t pairs = 0;
  uint32_t reserved = 0;
  uint16_t pair_pq[kMaxPairs] = {};  // ports p | q << 8 of pair i, p < q
  uint64_t tpl[kMaxTemplates] = {};  // pair index of added edge j in byte j
};

// One rectangular block of the cost output: task rows [task_offset,
// task_offset + task_count) of instance `instance` evaluated against the plan
// for k, written row-major at cells [cost_offset, cost_offset + task_count*T).
struct KoptSegment {
  uint32_t instance;
  uint32_t k;
  uint64_t task_offset;
  uint64_t task_count;
  uint64_t cost_offset;
  uint64_t tile_begin;  // first global tile index of this segment
  uint32_t rows_per_tile;
  uint32_t reserved;
};

// One wavefront tile: `rows` task rows starting at task_index in the task
// pool, whose cells start at cost_index in the cost pool, on the instance
// whose coordinates start at coord_base in the coordinate pool.
struct KoptTile {
  uint64_t task_index;
  uint64_t cost_index;
  uint64_t coord_base;
  uint32_t rows;
  uint16_t distance_type;
  uint16_t near;
};

// Coordinates of one instance: xy[base .. base + n) in the coordinate pool.
// `near` = 1 when both coordinate extents are below 2^25 (NearExtent), so the
// kernels may use the FP64 exact distance.
struct KoptInstance {
  uint64_t coord_base;
  uint32_t n;
  uint16_t distance_type;
  uint16_t near;
};

// LDS layout of one k inside a launch: the block-shared plan (pair ports and
// packed template pair indices) and the per-wave tile arrays.  Cells are
// walked template-stationary: when T <= 64 the wave holds `group` = 64 / T
// copies of the templates, copy g walking rows g, g + group, g + 2 group, ...;
// otherwise `passes` blocks of 64 templates each walk all rows.  Pair records
// are stored pair-major with `pair_stride` records per pair.
struct KoptKLayout {
  uint32_t present;
  uint32_t P, T, rows;
  uint32_t group, passes, pair_stride;
  uint32_t pair_step_r, pair_step_p;  // 64 = pair_step_r * P + pair_step_p
  uint32_t pq_off, tpl_off;
  uint32_t xy_off, row_off, dcost_off, pair_off;
};

// A segment carried in the kernel arguments: for launches of at most
// kInlineSegments segments the kernel derives each tile from these instead of
// reading a tile descriptor, which takes one dependent memory access off the
// start of every tile.
struct KoptInlineSegment {
  uint64_t task_offset;
  uint64_t cost_offset;
  uint32_t tile_begin;  // global tile index of the segment's first tile
  uint32_t task_count;
  uint32_t coord_base;
  uint16_t distance_type;
  uint16_t near;
};
constexpr uint32_t kInlineSegments = 8;
// Cost pools start every segment on a multiple of this many int64 cells (128
// bytes, one L2 line): partial-line stores cost the memory system a
// read-modify-write.
constexpr uint64_t kSegmentAlignCells = 16;

// Tiles are numbered k by k: tiles [k_tile_begin[i], k_tile_begin[i] +
// k_tile_count[i]) belong to segments with k = 3 + i, and the k_block_count[i]
// workgroups from k_block_first[i] process exactly those tiles, so every
// workgroup runs code for one k only, in the order k = 3, 4, 5.
struct KoptLaunch {
  const longlong2* xy;         // coordinate pool
  const KoptTile* tiles;       // [tile_count]; unused when inline_count > 0
  uint64_t tile_count;
  uint64_t k_tile_begin[3];
  uint64_t k_tile_count[3];
  uint32_t k_block_first[3];   // filled by LaunchKopt
  uint32_t k_block_count[3];
  const ee_kopt_task* tasks;   // task pool
  const KoptPlan* plans;       // [3] for k = 3, 4, 5
  int64_t* costs;              // cost pool
  // non-null: instead of the cost cells, write per task row (global row r of
  // the task pool) the bit mask of its improving templates, bit j of word
  // masks[r * mask_words + j / 32] set iff cell (r, j) < the row's
  // deleted_cost (the cells TryReconnectFromCostRow rechecks)
  uint32_t* masks;
  uint32_t mask_words;
  KoptKLayout layout[3];
  uint32_t wave_base;   // LDS offset of the first wave's tile arrays
  uint32_t wave_bytes;  // LDS bytes per wave
  // inline_count > 0: the launch's segments, k by k in tile order;
  // inline_seg[k_inline_begin[i] .. k_inline_begin[i + 1]) have k = 3 + i
  uint32_t inline_count;
  uint32_t k_inline_begin[4];
  KoptInlineSegment inline_seg[kInlineSegments];
};

// Rows per wavefront tile for each k (at most 64); `grid` = 0 launches
// enough workgroups for one tile per wavefront.
struct KoptTuning {
  uint32_t rows_per_tile[3] = {32, 16, 8};
  uint32_t grid = 0;
  uint32_t block = 256;
  bool non_temporal = false;
};

// Number of tiles the segments split into under `tuning`.
uint64_t KoptCountTiles(const KoptSegment* segments, size_t count, const KoptTuning& tuning);
// Sets rows_per_tile and tile_begin of every segment, numbering tiles k by k,
// writes the descriptor of every tile to `tiles` (KoptCountTiles entries; the
// segments' instances index `instances`; may be null) and the per-k tile
// ranges to `a`.  Launches of at most kInlineSegments segments also get the
// inline segment table (a->inline_count > 0): their kernels ignore `tiles`,
// so callers need not upload the descriptors.
void KoptAssignTiles(KoptSegment* segments, size_t count, const KoptInstance* instances,
                     const KoptPlan plans[3], const KoptTuning& tuning, KoptTile* tiles,
                     KoptLaunch* a);
// Fills the layout of `a` for the plans present and returns the dynamic LDS
// bytes of one workgroup.  `a->masks` must already be set: mask output stages
// each row's deleted cost in LDS.
// near_only: every tile of the launch is on a near instance, whose port
// records take 16 bytes instead of 32.
size_t KoptPrepareLayout(KoptLaunch* a, const KoptPlan plans[3], const bool present[3],
                         const KoptTuning& tuning, bool near_only);
void LaunchKopt(const KoptLaunch& args, const KoptTuning& tuning, size_t shared_bytes,
                hipStream_t stream);
// Fills the workgroup ranges of every k (k_block_first, k_block_count) and
// returns the grid; 0 for an empty launch.
uint32_t KoptFinalize(KoptLaunch* args, const KoptTuning& tuning, size_t shared_bytes);
// Launches with the arguments read from `device_args`, a device copy of
// `finalized` (a KoptFinalize result) that stays unchanged until the launch
// has run.  stamps (measurement only, may be null): 2 x grid x waves per
// block entries, wavefront w writing the constant-rate device clock
// (s_memrealtime) at its start to stamps[2w] and at its end to stamps[2w+1].
void LaunchKoptResident(const KoptLaunch& finalized, const KoptLaunch* device_args, uint32_t grid,
                        const KoptTuning& tuning, size_t shared_bytes, hipStream_t stream,
                        unsigned long long* stamps = nullptr);

// ------------------------------------------------------------------- JV
// Edge slots edge_list[0..edge_count) (or 0..edge_count when edge_list is
// NULL) are tested; item i's witness goes to witnesses[i].
struct JvLaunch {
  int32_t edge_count;
  int32_t distance_type;
  int32_t near;                 // NearExtent of the coordinates: FP64 distances
  int32_t one_per_wave;         // 1: one edge slot per wavefront; 0: two when they fit
  const int2* edge_uv;          // [m]
  const int64_t* edge_weight;   // [m]
  const int32_t* edge_active;   // [m]
  const int32_t* row_offsets;   // [n+1]
  const int32_t* neighbors;     // [A]
  const int32_t* csr_edge_ids;  // [A]
  const longlong2* xy;          // [n]
  int32_t* witnesses;           // [m]
  // optional, for several graphs in one launch: per edge slot, bit 0 the
  // distance type and bit 1 the near flag of the slot's graph
  const uint8_t* edge_mode;     // [m] or NULL
  const int32_t* edge_list;     // [edge_count] or NULL
};

void LaunchJv(const JvLaunch& args, uint32_t grid, hipStream_t stream);

// ------------------------------------------------------- k-opt work queue
// A persistent kernel serves leaf-cost requests from a ring of entries in
// coherent host memory.  Workgroup 0 (one wavefront) polls the ring's tail,
// copies new entries to a device mirror and writes one ticket per tile,
// (entry + 1) << 24 | tile, into a device ticket ring.  Every other
// workgroup (one wavefront each) takes the next ticket index with one atomic
// add, polls that ticket's own slot until the poller fills it, clears it and
// runs the tile; so waves never retry a claim and idle waves poll distinct
// addresses.  After its cells each tile wave stores entry + 1 into the
// tile's mark in host memory, and the wave that ends an entry's last tile
// stores entry + 1 into the entry's completion flag there.  A wave's stores
// reach host memory in order, while the flag may arrive before another
// wave's cells do, so the host takes an entry's cells once its flag and all
// its tile marks are there.  Every wave ends on its own clock: the poller
// after `idle_ticks` without new entries, after `life_ticks`, or when the
// host sets `stop`; a tile wave once the poller has ended with its ticket
// unfilled.  The host launches the kernel again whenever it publishes into a
// queue whose kernel has ended; the new poller first rewinds the ticket
// index past the tickets the ended launch handed out but never filled.
//
// One entry (64 bytes).  A leaf-cost query (k = 3..5): task rows
// [task_offset, task_offset + task_count) of the task arena, cells from
// cost_offset of the cost arena.  A JV recheck (k = 0): task_count edge
// slots of the resident JV union as int32 from row task_offset of the task
// arena (the request holds `tiles` rows there, one mark each), witnesses as
// int32 from cell cost_offset of the cost arena; kQueueJvTileEdges slots per
// tile.
struct KoptQueueEntry {
  uint64_t seq;          // entry index; ring slot = seq % ring
  uint64_t task_offset;
  uint64_t cost_offset;
  uint64_t coord_base;
  uint64_t done_target;  // the slot's done counter once this entry's last tile ended
  uint32_t task_count;
  uint32_t tiles;
  uint32_t k;
  uint32_t node_count;
  uint16_t distance_type;
  uint16_t near;
  uint32_t reserved;
};
static_assert(sizeof(KoptQueueEntry) == 64, "a queue entry is 64 bytes");
// Flag value of an entry the kernel refused (bounds checks): entry + 1 with
// this bit; the queue is then out of service.
constexpr uint64_t kQueueRefused = uint64_t{1} << 63;

// Written by the host, read by the poller (coherent host memory).
struct KoptQueueHost {
  uint64_t tail;          // entries published
  uint32_t stop;          // nonzero: the poller ends at its next look
  uint32_t reserved;
  uint64_t ended_epoch;   // written by the poller: the launch it ended
};

// Device memory, persisting across launches.
struct KoptQueueDevice {
  uint64_t taken;         // ticket indices handed to tile waves
  uint64_t written;       // tickets the poller has written
  uint64_t published;     // entries copied to the mirror
  uint32_t exit_epoch;    // the launch whose poller has ended
  uint32_t open_epoch;    // the launch whose poller has rewound `taken`
  uint32_t refused;       // nonzero once an entry failed the bounds checks
  uint32_t reserved;
};
// Tiles of one entry and entries are numbered in a ticket's 24 and 40 bits.
constexpr uint64_t kQueueMaxTiles = (uint64_t{1} << 24) - 1;
// Edge slots per JV recheck tile: two pairs of edges on one wavefront.
constexpr uint32_t kQueueJvTileEdges = 4;

struct KoptQueueArgs {
  KoptQueueHost* host;             // device address of the host control words
  const KoptQueueEntry* ring;      // device address of the host ring
  uint64_t* flags;                 // device address of the host completion flags
  uint64_t* marks;                 // device address of the host tile marks [task_capacity]
  KoptQueueDevice* dev;
  KoptQueueEntry* mirror;          // [ring]
  uint64_t* done;                  // [ring] tiles ended, per slot, cumulative
  uint64_t* tickets;               // [ticket_capacity], 0 = empty
  uint64_t task_capacity;          // rows of the task arena
  uint64_t cost_capacity;          // cells of the cost arena
  uint64_t coord_count;            // coordinates in the pool
  uint64_t ticket_capacity;        // at least the task arena's rows: every tile has a row
  uint64_t idle_ticks, life_ticks; // s_memrealtime ticks
  uint32_t ring_size;
  uint32_t epoch;                  // this launch, counted from 1
  // JV rechecks: the resident union (edge_list, witnesses and edge_count are
  // set per tile) and a device staging row per workgroup (kQueueJvTileEdges
  // slots, then as many witnesses); jv_slot_count = 0 when the queue holds
  // no JV graph
  JvLaunch jv;
  uint64_t jv_slot_count;
  int32_t* jv_stage;
};

// LDS layout for the queue kernel: one wavefront per workgroup, the plans of
// all three k side by side.  Returns the dynamic LDS bytes.
size_t KoptQueueLayout(KoptLaunch* a, const KoptPlan plans[3], const KoptTuning& tuning);
// Launches 1 + workers one-wavefront workgroups.  `a` carries the layout,
// the coordinate pool, the plans, the task arena (tasks) and the cost arena
// (costs).
void LaunchKoptQueue(const KoptLaunch& a, const KoptQueueArgs& q, uint32_t workers, size_t shared_bytes,
                     hipStream_t stream);

}  // namespace ee
