/*
 * ee_hip.h - C ABI of the MI355X (gfx950) backend for the two candidate
 * generation hot paths of cudaEdgeElimination:
 *
 *   1. the Jonker-Volgenant fast-elimination edge test over a CSR graph, and
 *   2. the exact 3/4/5-opt leaf cost matrix (task rows x reconnect templates,
 *      int64 cells).
 *
 * The entry points below are what a reference-side binding calls in place of
 * the reference's CUDA translation units; each declaration names the
 * reference interface it replaces (file:line in the reference tree).  Plain
 * pointers and sizes only.  Every call is synchronous unless it says
 * otherwise; nothing here allocates memory that the caller must free except
 * through the matching *_destroy call.
 *
 * Error convention: every function returning ee_status returns EE_OK on
 * success; on failure the message of the failing call is available from
 * ee_last_error() on the same host thread.  The reference throws C++
 * exceptions instead (src/cuda/jv_cuda.cu:23-27, src/cuda/kopt_cost.cu:434-442);
 * the status codes map one-to-one onto the exception types it throws.
 *
 * Arithmetic: distances are the exact TSPLIB EUC_2D / CEIL_2D integers
 * computed from int64 coordinates (floor square root from an IEEE FP64 square
 * root plus an integer correction, bit-identical to the reference's integer
 * restoring method for every 64-bit radicand); every cost is int64.
 */
#ifndef EE_HIP_H
#define EE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EE_OK = 0,
  EE_ERR_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument */
  EE_ERR_OVERFLOW = 2,         /* reference: std::overflow_error */
  EE_ERR_RUNTIME = 3,          /* reference: std::runtime_error (device errors too) */
  EE_ERR_LOGIC = 4,            /* reference: std::logic_error */
  EE_ERR_UNAVAILABLE = 5       /* no usable device; reference: std::runtime_error */
} ee_status;

enum { EE_EUC_2D = 0, EE_CEIL_2D = 1 };

/* Message of the last failing call on this host thread ("" if none). */
const char* ee_last_error(void);
/* Library version string. */
const char* ee_version(void);

/* ------------------------------------------------------------------ device */

/* Number of visible HIP devices (0 when none or on error). */
int ee_device_count(void);

/* Replaces cudaee::CudaBackendAvailable (include/cuda_edge_elimination/
 * elimination.hpp:333) and cudaee::detail::KOptCostCudaAvailable
 * (include/cuda_edge_elimination/local_search.hpp:235).  device = -1 asks
 * whether any device is usable.  Returns EE_OK or EE_ERR_UNAVAILABLE with the
 * reason copied into `reason` (NUL-terminated, may be NULL). */
ee_status ee_backend_available(int device, char* reason, size_t reason_capacity);

/* A context owns the device-resident caches of one host thread on one device:
 * the reference keeps them in thread_local per-device caches
 * (src/cuda/jv_cuda.cu:113-139, src/cuda/kopt_cost.cu:111-134) and picks the
 * device with SelectDevice (src/cuda/jv_cuda.cu:290-347): forced ordinal,
 * else sticky, else the device with the most free memory.  Here the forced
 * ordinal is the `device` argument; device = -1 picks the device with the
 * most free memory.  A context must be used by one host thread at a time.
 * Batches, sessions and servers created on a context must be destroyed
 * before the context. */
typedef struct ee_ctx ee_ctx;
ee_status ee_ctx_create(int device, ee_ctx** out);
void ee_ctx_destroy(ee_ctx* ctx);
int ee_ctx_device(const ee_ctx* ctx);
/* Waits for all work the context enqueued. */
ee_status ee_ctx_synchronize(ee_ctx* ctx);
/* The snapshot content hash of ee_graph_content_hash (src/common/graph.cpp:
 * 323-346: FNV-1a 64 over the header, coordinates and edge slots) with the
 * edge part evaluated on the device: per chunk of slots the permutation of
 * the hash's low byte, walks giving every chunk's starting low byte, then
 * per chunk its weighted FNV sum.  Equal to the byte-serial hash; the JV
 * session hashes its snapshots this way.  No counterpart in the reference,
 * which hashes on the host. */
ee_status ee_graph_content_hash_hip(ee_ctx* ctx, int32_t n, uint8_t distance_type, const int64_t* x,
                                    const int64_t* y, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                                    const int64_t* edge_weight, const uint8_t* edge_active, uint64_t* hash);
/* JV kernel shape for the context's later JV launches (direct calls,
 * sessions, fleets): 2 (the default) puts two edge slots on one wavefront,
 * one per 32-lane half, when both rows of an edge fit the half, and runs the
 * others on a whole wavefront; 1 gives every edge slot a whole wavefront.
 * Both give the same witnesses.  No counterpart in the reference, whose
 * kernel (src/cuda/jv_cuda.cu:221-288) has one thread per edge slot. */
ee_status ee_ctx_set_jv_edges_per_wave(ee_ctx* ctx, uint32_t edges);

/* ---------------------------------------------------------------------- JV */

/* Host view of a graph snapshot (reference GraphSnapshot,
 * include/cuda_edge_elimination/graph.hpp:13-37).  Edges are in edge-id
 * order; the CSR rows are sorted by neighbour (as RebuildCsr produces them,
 * src/common/graph.cpp:230-282).  Coordinates must be integers whose bounding
 * box diagonal squared fits in uint64 (integer_distance_safe). */
typedef struct {
  int32_t dimension;           /* n */
  uint8_t distance_type;       /* EE_EUC_2D or EE_CEIL_2D */
  uint8_t integer_coordinates; /* must be 1 */
  uint8_t reserved[2];
  const int64_t* x;            /* [n] */
  const int64_t* y;            /* [n] */
  int32_t edge_count;          /* m, all edge slots, active or not */
  const int32_t* edge_u;       /* [m] */
  const int32_t* edge_v;       /* [m] */
  const int64_t* edge_weight;  /* [m], each in [0, 2^62] */
  const uint8_t* edge_active;  /* [m], 0 or 1 */
  const int32_t* row_offsets;  /* [n+1], row_offsets[0] = 0 */
  const int32_t* neighbors;    /* [row_offsets[n]] */
  const int32_t* csr_edge_ids; /* [row_offsets[n]] */
} ee_graph_view;

/* Reference JvCudaCacheUsage (include/cuda_edge_elimination/elimination.hpp:41-49). */
typedef struct {
  int32_t static_hit;
  int32_t workspace_hit;
  uint64_t resident_bytes; /* 24m + 20n + 4 + 8A on an exact first fit */
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
} ee_jv_usage;

/* Replaces cudaee::FindJvCandidatesCuda (src/cuda/jv_cuda.cu:488-566).
 * witnesses[e] (m entries) receives the first JV witness of edge slot e, or
 * -1 when the edge is inactive, has an endpoint of degree <= 2, or has no
 * witness among its ten best candidates.  The reference returns the same
 * information compacted into Candidate{edge_id, witness, JV} in edge-id
 * order.  usage may be NULL. */
ee_status ee_jv_find_candidates(ee_ctx* ctx, const ee_graph_view* graph, int32_t* witnesses,
                                ee_jv_usage* usage);

/* Replaces cudaee::ClearJvCudaCache (src/cuda/jv_cuda.cu:568-571). */
ee_status ee_jv_clear_cache(ee_ctx* ctx);

/* ------------------------------------------------------------------- k-opt */

/* Reference KOptCostTask (include/cuda_edge_elimination/local_search.hpp:54-57):
 * deleted edge e joins ports 2e and 2e+1; ports >= 2k are ignored. */
typedef struct {
  int32_t port_nodes[10];
  int64_t deleted_cost;
} ee_kopt_task;

/* Reference EndpointMatching (include/cuda_edge_elimination/path_system.hpp:275-280). */
typedef struct {
  uint8_t endpoint_count;
  uint8_t mate[14];
} ee_endpoint_matching;

/* Reference KOptCudaCacheUsage (include/cuda_edge_elimination/local_search.hpp:59-64)
 * plus wall-clock phase times. */
typedef struct {
  int32_t snapshot_hit;
  int32_t template_hit;
  int32_t workspace_hit;
  int32_t reserved;
  uint64_t resident_bytes;
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
} ee_kopt_usage;

/* Replaces cudaee::BuildKOptReconnectTable (src/cpu/kopt_search.cpp:1143-1193):
 * the canonical proper k-opt reconnect templates (4 / 25 / 208 for
 * k = 3 / 4 / 5) and their FNV-1a generator hash. */
ee_status ee_kopt_build_templates(uint32_t k, ee_endpoint_matching* templates, uint32_t capacity,
                                  uint32_t* count, uint64_t* generator_hash);

/* Replaces cudaee::detail::EvaluateKOptTemplateCostsCuda
 * (src/cuda/kopt_cost.cu:415-470).  costs receives task_count x
 * template_count int64 cells, row-major; a cell is INT64_MAX when its
 * template adds a self loop, re-adds a deleted edge, adds an edge twice, or
 * overflows.  Tasks are validated like ValidateKOptCostTask
 * (src/cpu/kopt_search.cpp:784-807).  usage may be NULL. */
ee_status ee_kopt_eval_costs(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                             uint8_t distance_type, uint32_t k,
                             const ee_endpoint_matching* templates, uint32_t template_count,
                             uint64_t generator_hash, const ee_kopt_task* tasks,
                             size_t task_count, int64_t* costs, ee_kopt_usage* usage);

/* Replaces cudaee::detail::ClearKOptCostCudaCache (src/cuda/kopt_cost.cu:472). */
ee_status ee_kopt_clear_cache(ee_ctx* ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ee_kopt_eval_costs with the same arguments and the same
 * cells, bit for bit, but no allocation and no cache.  Its device staging
 * (the reconnect plan, x and y, the interleaved coordinates, the task rows
 * and the cost cells; about 23 KB + 32 n + 48 task_count + 8 task_count x
 * template_count bytes, each part rounded up to 256 bytes) lives in one
 * static array of EE_NOALLOC_SCRATCH_BYTES per GPU, which each GPU the
 * library runs on holds until the process ends; a call that needs more fails
 * with EE_ERR_INVALID_ARGUMENT before touching the device.  Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU (from several
 * threads or contexts) race and must be serialised by the caller.  usage
 * reports resident_bytes = the scratch bytes the call used and no cache
 * hits. */
#define EE_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
ee_status ee_kopt_eval_costs_noalloc(ee_ctx* ctx, const int64_t* x, const int64_t* y,
                                     int32_t dimension, uint8_t distance_type, uint32_t k,
                                     const ee_endpoint_matching* templates,
                                     uint32_t template_count, uint64_t generator_hash,
                                     const ee_kopt_task* tasks, size_t task_count,
                                     int64_t* costs, ee_kopt_usage* usage);

/* -------------------------------------------------------- HT replies */
/* Reference HtNeighborPair, NodeEdge, HtEndReplyTask
 * (include/cuda_edge_elimination/hamilton_tutte.hpp:26-32, 76-81) and
 * HtReplyCudaCacheUsage (:515-519) plus timings. */
typedef struct {
  int32_t center;
  int32_t first;
  int32_t second;
} ee_ht_neighbor_pair;
typedef struct {
  int32_t u;
  int32_t v;
} ee_node_edge;
typedef struct {
  int32_t endpoint;
  int32_t internal_neighbor;
} ee_ht_end_task;
typedef struct {
  int32_t graph_hit;
  int32_t workspace_hit;
  uint64_t resident_bytes;
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
} ee_ht_usage;

/* Replaces cudaee::EvaluateHtHamiltonRepliesCuda (src/cuda/ht_replies.cu:585-663):
 * for each center in order, every pair of its CSR neighbours (first before
 * second) that is not the target edge, whose two center edges survive the
 * 2-opt test against the target, and whose target + center edges cost at
 * most the pair's edge + the center-to-target-end edges
 * (src/cpu/hamilton_tutte.cpp:230-262).  The target may be given in either
 * order.  offsets receives center_count + 1 entries (center i's replies are
 * [offsets[i], offsets[i+1])), replies the first min(*total, capacity)
 * replies (ee_ht_copy_replies copies the rest), *total their count.  A count
 * pass runs first, so the device holds exactly *total replies, never the
 * sum over centers of C(degree, 2).  The snapshot must pass the reference's HT graph checks
 * (ht_replies.cu:521-553): a symmetric CSR of exactly the active edges,
 * rows strictly ascending. */
ee_status ee_ht_hamilton_replies(ee_ctx* ctx, const ee_graph_view* graph, int32_t target_u,
                                 int32_t target_v, const int32_t* centers, size_t center_count,
                                 uint64_t* offsets, ee_ht_neighbor_pair* replies, uint64_t capacity,
                                 uint64_t* total, ee_ht_usage* usage);
/* Replaces cudaee::EvaluateHtEndRepliesCuda (src/cuda/ht_replies.cu:667-751):
 * per task every active edge of the endpoint except the one to its internal
 * neighbour, canonical (u < v), in CSR order.  Outputs as
 * ee_ht_hamilton_replies (*total = the sum of degree - 1). */
ee_status ee_ht_end_replies(ee_ctx* ctx, const ee_graph_view* graph, const ee_ht_end_task* tasks,
                            size_t task_count, uint64_t* offsets, ee_node_edge* replies,
                            uint64_t capacity, uint64_t* total, ee_ht_usage* usage);
/* Copies replies [first, first + count) of the context's last completed
 * ee_ht_hamilton_replies (12-byte ee_ht_neighbor_pair records) or
 * ee_ht_end_replies (8-byte ee_node_edge records) call into `replies`; they
 * stay on the device until the next reply call.  A caller that does not know
 * the reply count beforehand passes capacity 0 (or a guess) to the reply
 * call, sizes its buffer from *total and copies the rest here, as the
 * reference sizes its vector from the scanned offsets
 * (src/cuda/ht_replies.cu:626-647). */
ee_status ee_ht_copy_replies(ee_ctx* ctx, uint64_t first, uint64_t count, void* replies);
/* Replaces cudaee::ClearHtReplyCudaCache (src/cuda/ht_replies.cu:753). */
ee_status ee_ht_clear_cache(ee_ctx* ctx);

/* ------------------------------------------- device deletion-set producer */
/* The TourContext fields the reference's k-opt cost cursor reads
 * (src/cpu/kopt_search.cpp:125-134): the tour (distinct snapshot nodes),
 * the exact cost of the edge leaving each tour position (-1 where that
 * edge is not a path edge), and the selectable positions (the path-edge
 * positions, with the required edge's first, kopt_search.cpp:318-349). */
typedef struct {
  const int32_t* tour;        /* [tour_length] */
  const int64_t* edge_costs;  /* [tour_length] */
  const uint64_t* selectable; /* [selectable_count], at most 65536 */
  uint32_t tour_length;
  uint32_t selectable_count;
} ee_tour_view;

/* Number of deletion sets of size k the reference cursor enumerates
 * (KOptSearchCursor::PrepareNextK, kopt_search.cpp:2258-2271, with
 * AdvanceCombination, kopt_search.cpp:736-752, which never moves index 0):
 * C(selectable_count - 1, k - 1) for 3 <= k <= 5 and k < selectable_count,
 * else 0. */
ee_status ee_kopt_cursor_count(const ee_tour_view* tour, uint32_t k, uint64_t* count);

/* Replaces the host loop of KOptSearchCursor::PrepareNextBlock
 * (src/cpu/kopt_search.cpp:2123-2156) and the task upload of
 * EvaluateKOptTemplateCostsCuda (src/cuda/kopt_cost.cu:449): the task rows
 * of deletion sets [first, first + count) of size k, in the cursor's order,
 * are built on the device by BuildKOptCostTask's rule
 * (kopt_search.cpp:561-584) and evaluated there like ee_kopt_eval_costs.
 * count is clamped to the sets left; *produced (may be NULL) receives the
 * rows made.  Outputs (each may be NULL): tasks (the rows), positions (each
 * row's k sorted tour positions, the cursor's deleted_positions), and at
 * most one of costs (rows x template_count cells) and masks (rows x
 * ceil(template_count / 32) words, bit j set iff cell j < the row's
 * deleted_cost).  With neither costs nor masks only the rows are made, and
 * x, y, distance_type and the templates are not read.  Fails with
 * EE_ERR_RUNTIME where a deleted cost sum overflows int64 or a selectable
 * position's cost is negative or out of the tour, as BuildKOptCostTask
 * throws; the positions are checked up front, not as the sets reach them.
 * usage->h2d_ms covers building the rows on the device. */
ee_status ee_kopt_cursor_eval(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                              uint8_t distance_type, const ee_tour_view* tour, uint32_t k,
                              const ee_endpoint_matching* templates, uint32_t template_count,
                              uint64_t generator_hash, uint64_t first, uint32_t count,
                              ee_kopt_task* tasks, uint32_t* positions, int64_t* costs,
                              uint32_t* masks, uint32_t* produced, ee_kopt_usage* usage);

/* ------------------------------------------------ certified leaf pipeline */
/* Replaces the certification step of cudaee::EvaluateKOptTemplateCosts
 * (src/cpu/kopt_search.cpp:1455-1475) and the row scan of
 * TryReconnectFromCostRow (kopt_search.cpp:677-734): recomputes every cell of
 * `cells` (task_count x template_count, row-major, e.g. a device matrix) on
 * the host by the reference CPU scorer's rules (KOptCostTaskCpuScorer,
 * kopt_search.cpp:962-1042) and compares them.  When all are equal it returns
 * EE_OK and, when `improving` is not NULL, writes per row
 * ceil(template_count / 32) words whose bit j is set iff cell j is below the
 * row's deleted_cost (the templates the consumer rechecks, in template
 * order).  Otherwise it fails with EE_ERR_RUNTIME and *mismatch (may be NULL)
 * receives the first differing cell, as the reference throws with the cell
 * index.  Runs on every core the process may use; pure host code. */
ee_status ee_kopt_certify(const int64_t* x, const int64_t* y, int32_t dimension,
                          uint8_t distance_type, uint32_t k, const ee_endpoint_matching* templates,
                          uint32_t template_count, const ee_kopt_task* tasks, size_t task_count,
                          const int64_t* cells, uint32_t* improving, uint64_t* mismatch,
                          uint64_t* improving_count);

/* No reference counterpart: the certified leaf loop of
 * ProvePathSystemsByKOptImpl (src/cpu/kopt_search.cpp:2651-2764) as a
 * pipeline.  The reference evaluates one block of task rows per k on the
 * device, waits for it, downloads every cell, recomputes the whole matrix on
 * the host to certify it and only then builds the next block, so device and
 * host time add up.  Here the caller submits up to `depth` blocks ahead
 * (rows uploaded from the host, or built on the device from the cost
 * cursor's deletion-set ranks); each runs its cost launch and downloads its
 * rows and cells into pinned host memory on its own, while ee_leaf_next
 * certifies the oldest block on the host threads (ee_kopt_certify's rule,
 * plus, for device-built rows, a host rebuild of every row from the cursor
 * order) and hands the consumer its certified cells and improving masks.
 * Blocks come back in submission order.  A block that fails certification,
 * or whose rows overflow their deleted cost, fails ee_leaf_next with
 * EE_ERR_RUNTIME and its cells are never returned.  Templates are the
 * canonical ones (ee_kopt_build_templates). */
typedef struct ee_leaf_pipeline ee_leaf_pipeline;

typedef struct {
  uint64_t tag;                 /* the submitter's tag */
  uint64_t first;               /* cursor blocks: rank of the first deletion set; else 0 */
  uint32_t k;
  uint32_t rows;
  uint32_t template_count;
  uint32_t mask_words;          /* ceil(template_count / 32) */
  const ee_kopt_task* tasks;    /* [rows] */
  const uint32_t* positions;    /* cursor blocks: [rows x k] sorted tour positions; else NULL */
  const int64_t* costs;         /* [rows x template_count], certified */
  const uint32_t* improving;    /* [rows x mask_words] */
  uint64_t improving_cells;     /* set bits of `improving` */
} ee_leaf_block;

typedef struct {
  uint64_t blocks;
  uint64_t rows;
  uint64_t cells;               /* certified */
  uint64_t improving_cells;
  double kernel_ms;             /* device: row build + cost launch, summed over blocks */
  double copy_ms;               /* device: download of rows and cells, summed */
  double wait_ms;               /* host time ee_leaf_next spent waiting for the device */
  double certify_ms;            /* host time certifying (rows and cells) */
  uint32_t certify_threads;
  uint32_t reserved;
} ee_leaf_stats;

/* depth in [1, 4]: blocks outstanding (submitted, not yet returned by
 * ee_leaf_next); max_rows in [1, 65536]: rows per block.  The coordinates are
 * copied (device and host) and must be distance safe. */
ee_status ee_leaf_create(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                         uint8_t distance_type, uint32_t depth, uint32_t max_rows,
                         ee_leaf_pipeline** out);
void ee_leaf_destroy(ee_leaf_pipeline* pipeline);
/* The tour context of later cursor blocks (as ee_kopt_cursor_eval); waits
 * for the blocks in flight first. */
ee_status ee_leaf_set_tour(ee_leaf_pipeline* pipeline, const ee_tour_view* tour);
/* Enqueues one block of host task rows (validated like ee_kopt_eval_costs). */
ee_status ee_leaf_submit_tasks(ee_leaf_pipeline* pipeline, uint32_t k, const ee_kopt_task* tasks,
                               uint32_t rows, uint64_t tag);
/* Enqueues the deletion sets [first, first + count) of size k of the tour
 * context, clamped to the sets left; *rows receives the rows (0: nothing was
 * enqueued). */
ee_status ee_leaf_submit_cursor(ee_leaf_pipeline* pipeline, uint32_t k, uint64_t first,
                                uint32_t count, uint64_t tag, uint32_t* rows);
/* Waits for the oldest outstanding block and certifies it; *block views stay
 * valid until the next ee_leaf_next or ee_leaf_destroy.  EE_ERR_LOGIC when
 * nothing is outstanding. */
ee_status ee_leaf_next(ee_leaf_pipeline* pipeline, ee_leaf_block* block);
ee_status ee_leaf_stats_get(const ee_leaf_pipeline* pipeline, ee_leaf_stats* stats);
/* Fault injection for tests of the fail-closed path (no counterpart in the
 * reference, whose certification is tested by construction): the next block
 * handed out by ee_leaf_next has cell `cell` of its host copy set to `value`
 * before certification, which must then refuse the block with
 * EE_ERR_RUNTIME naming the cell (unless the value is the true one). */
ee_status ee_leaf_inject_cell(ee_leaf_pipeline* pipeline, uint64_t cell, int64_t value);

/* --------------------------------------------------- resident k-opt batch */
/* No reference counterpart.  The reference evaluates one (graph, k) matrix
 * per synchronous call with a full upload and download
 * (src/cpu/kopt_search.cpp:2619-2692, src/cuda/kopt_cost.cu:437-462).  A
 * batch keeps the coordinates of many instances, all task rows and the whole
 * cost output resident on the device and evaluates every (instance, k)
 * segment in one kernel launch.  Cells are bit-identical to
 * ee_kopt_eval_costs with the canonical templates. */
typedef struct ee_kopt_batch ee_kopt_batch;

typedef struct {
  uint32_t instance;   /* id from ee_kopt_batch_add_instance */
  uint32_t k;          /* 3, 4 or 5 */
  uint64_t task_count; /* rows of this segment, consecutive in the task array */
} ee_kopt_segment_desc;

typedef struct {
  uint32_t rows_per_tile[3]; /* task rows per wavefront tile for k = 3, 4, 5, in [1, 64] */
  uint32_t grid;             /* 0: one tile per wavefront; else a persistent grid of
                                this many workgroups, shared among k by work */
  uint32_t block;            /* threads per workgroup: 64, 128, 192 or 256 */
  uint32_t non_temporal;     /* 1: cost stores bypass the caches */
} ee_kopt_tuning;

ee_status ee_kopt_batch_create(ee_ctx* ctx, ee_kopt_batch** out);
void ee_kopt_batch_destroy(ee_kopt_batch* batch);
/* Uploads one instance's coordinates; *instance receives its id. */
ee_status ee_kopt_batch_add_instance(ee_kopt_batch* batch, const int64_t* x, const int64_t* y,
                                     int32_t dimension, uint8_t distance_type,
                                     uint32_t* instance);
/* Validates and uploads the task rows of every segment (segment order) and
 * sizes the resident cost output.  cell_offsets (nullable, count + 1
 * entries) receives the first cell of every segment and the size of the cost
 * pool; segment s holds task_count * T(k) cells from cell_offsets[s], and
 * every segment starts on a multiple of 16 cells (128 bytes). */
ee_status ee_kopt_batch_set_work(ee_kopt_batch* batch, const ee_kopt_segment_desc* segments,
                                 uint32_t segment_count, const ee_kopt_task* tasks,
                                 uint64_t* cell_offsets);
ee_status ee_kopt_batch_set_tuning(ee_kopt_batch* batch, const ee_kopt_tuning* tuning);
ee_status ee_kopt_batch_get_tuning(const ee_kopt_batch* batch, ee_kopt_tuning* tuning);
/* Times `candidates_launches` launches of each tuning candidate (coordinate
 * descent over rows per tile, grid, block and store policy), keeps the
 * fastest, and, when cache_path is non-NULL, reads the choice from / writes
 * it to that file under a key made of the device name and the batch shape. */
ee_status ee_kopt_batch_autotune(ee_kopt_batch* batch, uint32_t launches_per_candidate,
                                 const char* cache_path, ee_kopt_tuning* chosen);
/* The same, timing each candidate in the run mode of `run_flags`: 0 (plain
 * launches, as ee_kopt_batch_autotune) or EE_KOPT_RUN_PIPELINED (at the
 * batch's pipeline depth, which is part of the cache key).  No reference
 * counterpart. */
ee_status ee_kopt_batch_autotune_runs(ee_kopt_batch* batch, uint32_t launches_per_candidate, int run_flags,
                                      const char* cache_path, ee_kopt_tuning* chosen);
/* Enqueues `launches` back-to-back evaluations of the whole batch and waits;
 * *device_ms receives the mean device time of one launch, measured with HIP
 * events from the first launch's start to the last one's end.  flags:
 * EE_KOPT_RUN_GRAPH replays a captured HIP graph of one launch instead of
 * launching directly; EE_KOPT_RUN_PIPELINED deals the launches round-robin
 * over the context's stream and further streams, each writing its own cost
 * pool of the same size (cost output only; ee_kopt_batch_set_pipeline_depth),
 * so that a launch starts while the previous one drains; every launch
 * evaluates the whole batch, and the cost pool that ee_kopt_batch_download
 * reads holds the full result afterwards.  0 = plain launches on the
 * context's stream.  The flags do not combine. */
#define EE_KOPT_RUN_GRAPH 1
#define EE_KOPT_RUN_PIPELINED 2
ee_status ee_kopt_batch_run(ee_kopt_batch* batch, uint32_t launches, int flags, double* device_ms);
/* ee_kopt_batch_run of `launches` (at most 1024) plain (flags 0) or
 * pipelined (EE_KOPT_RUN_PIPELINED) launches in which every wavefront
 * records the device's constant-rate clock (s_memrealtime) when it starts
 * and ends: start_us[i] and end_us[i] receive launch i's first wavefront
 * start and last wavefront end, in microseconds from the run's first start.
 * Launches whose intervals intersect ran at the same time.  No reference
 * counterpart (a measurement aid). */
ee_status ee_kopt_batch_run_stamped(ee_kopt_batch* batch, uint32_t launches, int flags, double* start_us,
                                    double* end_us);
/* Streams (and cost pools) of pipelined runs, 2 to 4 (default 2); each pool
 * beyond the first is allocated on the first pipelined run that needs it.
 * No reference counterpart. */
ee_status ee_kopt_batch_set_pipeline_depth(ee_kopt_batch* batch, uint32_t depth);
ee_status ee_kopt_batch_download(ee_kopt_batch* batch, uint64_t first_cell, uint64_t cell_count,
                                 int64_t* costs);
/* Cells of cost pool `pool` of pipelined runs (0 is the pool that
 * ee_kopt_batch_download reads; 1 .. depth - 1 the further streams' pools).
 * Fails with EE_ERR_LOGIC for a pool that no launch of the current work has
 * written.  No reference counterpart: it lets a caller check every pool a
 * pipelined run filled. */
ee_status ee_kopt_batch_download_pool(ee_kopt_batch* batch, uint32_t pool, uint64_t first_cell,
                                      uint64_t cell_count, int64_t* costs);
/* Output mode of later runs: 0 = the int64 cost cells (default); 1 = per
 * task row (rows numbered in segment order) the bit mask of its improving
 * templates, bit j of word j / 32 set iff cell (row, j) < the row's
 * deleted_cost, i.e. exactly the templates TryReconnectFromCostRow rechecks
 * (src/cpu/kopt_search.cpp:693-697), in template order.  No reference
 * counterpart: it replaces the 8-byte-per-cell download by T / 8 bytes per
 * row for callers that only consume improving cells. */
ee_status ee_kopt_batch_set_output(ee_kopt_batch* batch, uint32_t mode);
/* 32-bit words per row of the mask output (ceil(T / 32) of the widest k). */
ee_status ee_kopt_batch_mask_words(const ee_kopt_batch* batch, uint32_t* words);
ee_status ee_kopt_batch_download_masks(ee_kopt_batch* batch, uint64_t first_row, uint64_t row_count,
                                       uint32_t* masks);
/* Device bytes held by the batch. */
uint64_t ee_kopt_batch_resident_bytes(const ee_kopt_batch* batch);

/* ------------------------------------------------------- resident JV epochs */
/* Replaces the propose step of cudaee::RunJvElimination with backend cuda
 * (src/cpu/jv_elimination.cpp:250-322) and the CSR refresh of its commit
 * (src/cpu/elimination_commit.cpp:64-72): the snapshot stays resident on the
 * device across epochs; each epoch proposes on the device, re-verifies every
 * candidate on the host (an epoch with a rejected candidate fails closed,
 * nothing committed) and runs the deterministic commit rule there; the
 * committed edge slots go to the device, which clears their active flags and
 * compacts its CSR while the host applies the same commit to its own copy
 * (kept for the re-verification) and hashes snapshots on a helper thread;
 * the proof is recorded.  The proof is the reference's V1 text
 * (src/cpu/jv_elimination.cpp:324-368) with backend name "hip". */
typedef struct ee_jv_session ee_jv_session;

typedef struct {
  uint32_t epochs;
  uint32_t reserved;
  uint64_t edges_scanned;   /* sum over epochs of active edges before the epoch */
  uint64_t proposed;
  uint64_t committed;
  uint64_t initial_hash;
  uint64_t final_hash;
  double device_ms;         /* JV kernel time, HIP events */
  double h2d_ms;            /* the first epoch's upload of active flags and CSR */
  double d2h_ms;
  double propose_ms;        /* host wall time from enqueueing an epoch's device work to its witnesses */
  double verify_ms;         /* host re-verification */
  double commit_ms;         /* host commit rule plus applying it to the host snapshot */
  double total_ms;
  double compact_ms;        /* device commit: deactivate and CSR compaction, HIP events */
  double hash_wait_ms;      /* time the end of the run waited for snapshot hashes */
} ee_jv_run_stats;

ee_status ee_jv_session_create(ee_ctx* ctx, const ee_graph_view* graph, ee_jv_session** out);
void ee_jv_session_destroy(ee_jv_session* session);
ee_status ee_jv_session_run(ee_jv_session* session, uint32_t max_rounds, ee_jv_run_stats* stats);
/* V1 proof text of the last run (NUL-terminated when capacity > 0);
 * *length receives the full length. */
/* Fault injection for tests of the fail-closed path (no counterpart in the
 * reference): in the next run's first epoch, the host copy of the device's
 * witness for `edge` is replaced by `witness` before re-verification, which
 * must stop the run with EE_ERR_RUNTIME and commit nothing, as the
 * reference stops on a rejected candidate (src/cpu/jv_elimination.cpp:
 * 301-303), unless the replacement is itself a valid witness. */
ee_status ee_jv_session_inject_witness(ee_jv_session* session, int32_t edge, int32_t witness);
ee_status ee_jv_session_proof(const ee_jv_session* session, char* buffer, size_t capacity,
                              size_t* length);
/* Active flags after the last run (edge_count entries). */
ee_status ee_jv_session_active(const ee_jv_session* session, uint8_t* active);
/* No reference counterpart: the session's device-resident snapshot, the one
 * the last epoch of the last run proposed on (the final snapshot when the
 * run reached its fixed point; one commit behind the host's when max_rounds
 * cut the run after a committing epoch, which the next run's upload
 * resolves), for checks against the host snapshot.  active: edge_count flags; row_offsets:
 * dimension + 1 entries; neighbors and edge_ids: row_offsets[dimension]
 * entries each, which must not exceed `capacity`.  Before any run the CSR
 * buffers hold no snapshot. */
ee_status ee_jv_session_device_snapshot(const ee_jv_session* session, uint8_t* active,
                                        int32_t* row_offsets, int32_t* neighbors, int32_t* edge_ids,
                                        size_t capacity);

/* ------------------------------------------------------- many JV graphs */
/* No single reference counterpart: runs RunJvElimination
 * (src/cpu/jv_elimination.cpp:250-368) for many independent graphs at once.
 * Every graph keeps its own host run (re-verification, commit, V1 proof
 * identical to a standalone ee_jv_session run of that graph); per epoch the
 * device side is one upload of all graphs' active flags and CSR, one JV
 * launch over all their edge slots and one download, optionally replayed as a
 * captured HIP graph (use_graph != 0).  Graphs that reached their fixed point
 * are skipped in later epochs.  Graphs may mix EUC_2D and CEIL_2D. */
typedef struct ee_jv_fleet ee_jv_fleet;

typedef struct {
  uint32_t instances;
  uint32_t epochs;          /* lockstep epochs until every graph stopped */
  uint64_t edge_slots;      /* edge slots of the union */
  uint64_t edges_scanned;   /* summed over graphs and epochs */
  uint64_t proposed;
  uint64_t committed;
  double device_ms;         /* per epoch: device commit + JV launch + download, or the replayed graph */
  double assemble_ms;       /* host packing of the union's arrays (first epoch) and commit lists */
  double propose_ms;        /* assemble + device round trip */
  double host_ms;           /* re-verification, commit rule and host CSR rebuild of all graphs */
  double total_ms;
  uint32_t graph_launches;  /* epochs replayed as a graph: every epoch after the first */
  uint32_t reserved;
} ee_jv_fleet_stats;

ee_status ee_jv_fleet_create(ee_ctx* ctx, const ee_graph_view* graphs, uint32_t count,
                             ee_jv_fleet** out);
void ee_jv_fleet_destroy(ee_jv_fleet* fleet);
ee_status ee_jv_fleet_run(ee_jv_fleet* fleet, uint32_t max_rounds, int use_graph,
                          ee_jv_fleet_stats* stats);
/* The last run of graph `index`: its stats (device fields zero), V1 proof
 * text and final active flags. */
ee_status ee_jv_fleet_instance_stats(const ee_jv_fleet* fleet, uint32_t index,
                                     ee_jv_run_stats* stats);
ee_status ee_jv_fleet_proof(const ee_jv_fleet* fleet, uint32_t index, char* buffer,
                            size_t capacity, size_t* length);
ee_status ee_jv_fleet_active(const ee_jv_fleet* fleet, uint32_t index, uint8_t* active);

/* ------------------------------------------------------- serving k-opt */
/* No reference counterpart.  A request loop over device-resident graph
 * snapshots: callers submit leaf-cost queries (one graph, one k, task rows)
 * and JV rechecks (ee_server_submit_jv) from any thread; a worker thread packs whatever has arrived into the next
 * batched launch (continuous batching) as long as the batch stays within the
 * cell and HBM budgets, runs up to `pipeline_depth` batches concurrently on
 * separate streams (upload, kernel and download of one batch overlap with the
 * next), copies each request's cells to the caller's buffer and marks it
 * complete.  Cells are bit-identical to ee_kopt_eval_costs with the canonical
 * templates.  With max_batch_requests = 1 and pipeline_depth = 1 the server
 * is the plain per-query path: one synchronous upload, launch and download
 * per request, in arrival order. */
typedef struct ee_server ee_server;

typedef struct {
  uint64_t hbm_budget_bytes;   /* device bytes for all in-flight batch buffers */
  uint64_t max_batch_cells;    /* cells per launch (a larger request runs alone) */
  uint32_t max_batch_requests; /* requests per launch */
  uint32_t pipeline_depth;     /* batches in flight: 1 or 2 */
} ee_server_config;

typedef struct {
  uint64_t request_id;
  int32_t status;       /* ee_status of the request */
  uint32_t batch_size;  /* requests in the launch that served it */
  uint64_t batch_id;
  double queue_ms;      /* submit -> its batch was issued */
  double service_ms;    /* issue -> cells in the caller's buffer */
  double latency_ms;    /* submit -> complete */
} ee_request_result;

typedef struct {
  uint64_t requests;
  uint64_t batches;
  uint64_t cells;       /* leaf-cost cells served */
  uint64_t jv_edges;    /* edge slots rechecked */
  uint64_t peak_inflight_bytes;
  double device_ms;     /* sum of batch kernel times (HIP events) */
} ee_server_stats;

ee_status ee_server_create(ee_ctx* ctx, const ee_server_config* config, ee_server** out);
/* Stops the worker (finishing queued requests) and frees everything. */
void ee_server_destroy(ee_server* server);
/* Uploads a graph snapshot's coordinates (before or while serving). */
ee_status ee_server_add_graph(ee_server* server, const int64_t* x, const int64_t* y,
                              int32_t dimension, uint8_t distance_type, uint32_t* graph);
/* Queues a request; validates it first (EE_ERR_INVALID_ARGUMENT when a task
 * is malformed or the request alone exceeds the HBM budget).  costs (count x
 * templates(k) cells) must stay valid until the request completes. */
ee_status ee_server_submit(ee_server* server, uint32_t graph, uint32_t k,
                           const ee_kopt_task* tasks, uint64_t count, int64_t* costs,
                           uint64_t* request_id);
/* Makes a JV graph snapshot (coordinates, edges, active flags) resident for
 * recheck requests (before or while serving); *jv_graph receives its id. */
ee_status ee_server_add_jv_graph(ee_server* server, const ee_graph_view* graph, uint32_t* jv_graph);
/* Queues a JV recheck of edge slots edges[0..count) of a resident JV graph:
 * witnesses[i] receives the JV witness of slot edges[i] on that snapshot
 * (the FindJvCandidatesCuda result, src/cuda/jv_cuda.cu:221-288), or -1.
 * Rechecks share launches with leaf-cost requests; witnesses must stay valid
 * until the request completes. */
ee_status ee_server_submit_jv(ee_server* server, uint32_t jv_graph, const int32_t* edges,
                              uint64_t count, int32_t* witnesses, uint64_t* request_id);
/* Blocks until the request completes (timeout_ms < 0: no limit). */
ee_status ee_server_wait(ee_server* server, uint64_t request_id, double timeout_ms,
                         ee_request_result* result);

/* One request of ee_server_submit_many. */
typedef struct {
  uint32_t kind;     /* 0: leaf-cost query; 1: JV recheck */
  uint32_t graph;    /* graph id (kind 0) or JV graph id (kind 1) */
  uint32_t k;        /* 3, 4 or 5 (kind 0) */
  uint32_t reserved;
  const void* input; /* ee_kopt_task rows (kind 0) or int32 edge slots (kind 1) */
  uint64_t count;    /* rows or edge slots */
  void* output;      /* count x T(k) int64 cells (kind 0) or count int32 witnesses (kind 1) */
} ee_server_request;
/* Queues `count` requests at once, in order, as a front end receiving many
 * queries in one message would: every request is validated first (in
 * parallel; on an error none is queued), then all are queued under one lock.
 * Unlike ee_server_submit the task rows are not copied: inputs and outputs
 * must stay valid until each request completes.  request_ids[i] receives the
 * id of requests[i]. */
ee_status ee_server_submit_many(ee_server* server, uint32_t count, const ee_server_request* requests,
                                uint64_t* request_ids);
/* ee_server_wait for each id in turn within one overall timeout; results
 * (nullable) has count entries.  Returns the first failure after waiting for
 * every request. */
ee_status ee_server_wait_many(ee_server* server, uint32_t count, const uint64_t* request_ids,
                              double timeout_ms, ee_request_result* results);
ee_status ee_server_stats_get(const ee_server* server, ee_server_stats* stats);

/* ------------------------------------------------- device work queue */
/* No reference counterpart.  Leaf-cost queries and JV rechecks served by a
 * persistent kernel: requests go into a ring in coherent host memory, the
 * kernel (polling that ring) writes each request's cells or witnesses into a
 * host arena and raises its completion flag there; a completing thread
 * copies them to the caller's buffer.  No launch or copy command is issued
 * per request.  The kernel ends after idle_exit_ms without requests and is
 * launched again by the next submit.  Cells are bit-identical to
 * ee_kopt_eval_costs with the canonical templates, witnesses to
 * ee_jv_find_candidates on the same snapshot.  While the kernel runs, calls
 * that wait for the whole device (hipDeviceSynchronize, hipFree) wait for it
 * to end. */
typedef struct ee_queue ee_queue;

typedef struct {
  uint64_t task_rows;     /* task arena (host memory): rows in flight */
  uint64_t cost_cells;    /* cost arena (host memory): cells in flight */
  uint64_t max_nodes;     /* device coordinate pool: nodes of all graphs */
  uint32_t ring_entries;  /* requests in flight */
  uint32_t workers;       /* tile wavefronts of the kernel */
  double idle_exit_ms;    /* the kernel ends after this long without a request (0, 1000] */
  uint64_t max_jv_nodes;      /* JV snapshots: nodes, edge slots and adjacency entries */
  uint64_t max_jv_edges;      /* of all JV graphs (all 0: no JV rechecks) */
  uint64_t max_jv_adjacency;
} ee_queue_config;

typedef struct {
  uint64_t requests;
  uint64_t cells;
  uint64_t launches;            /* kernel launches (one per busy period) */
  uint64_t peak_inflight_cells; /* cost-arena cells held by requests in flight */
  uint64_t jv_edges;            /* edge slots rechecked */
} ee_queue_stats;

ee_status ee_queue_create(ee_ctx* ctx, const ee_queue_config* config, ee_queue** out);
/* Serves every queued request, stops the kernel and frees everything. */
void ee_queue_destroy(ee_queue* queue);
/* Uploads a graph's coordinates into the pool (before or while serving; a
 * running kernel first serves what is queued and ends). */
ee_status ee_queue_add_graph(ee_queue* queue, const int64_t* x, const int64_t* y, int32_t dimension,
                             uint8_t distance_type, uint32_t* graph);
/* Makes a JV snapshot resident for rechecks, like ee_server_add_jv_graph. */
ee_status ee_queue_add_jv_graph(ee_queue* queue, const ee_graph_view* graph, uint32_t* jv_graph);
/* Queues a JV recheck of edge slots edges[0..count) of a resident JV graph;
 * witnesses[i] receives slot edges[i]'s witness (graph node id) or -1. */
ee_status ee_queue_submit_jv(ee_queue* queue, uint32_t jv_graph, const int32_t* edges, uint64_t count,
                             int32_t* witnesses, uint64_t* request_id);
/* Validates and queues a request (the rows are copied); waits for room when
 * the arenas or the ring are full.  costs (count x templates(k) cells) must
 * stay valid until the request is collected. */
ee_status ee_queue_submit(ee_queue* queue, uint32_t graph, uint32_t k, const ee_kopt_task* tasks,
                          uint64_t count, int64_t* costs, uint64_t* request_id);
/* ee_queue_submit / ee_queue_submit_jv for each request in order; every
 * request is validated first (on an error none is queued). */
ee_status ee_queue_submit_many(ee_queue* queue, uint32_t count, const ee_server_request* requests,
                               uint64_t* request_ids);
/* Blocks until the request's cells are in the caller's buffer (timeout_ms <
 * 0: no limit).  batch_size is 1, batch_id the ring entry, queue_ms 0. */
ee_status ee_queue_wait(ee_queue* queue, uint64_t request_id, double timeout_ms, ee_request_result* result);
ee_status ee_queue_wait_many(ee_queue* queue, uint32_t count, const uint64_t* request_ids, double timeout_ms,
                             ee_request_result* results);
ee_status ee_queue_stats_get(const ee_queue* queue, ee_queue_stats* stats);
/* No reference counterpart; a test hook for the fail-closed path.  Queues
 * one k = 3 leaf-cost request (one valid task row) whose ring entry points
 * past the task arena.  The kernel's bounds check refuses the entry before
 * any tile of it runs; waiting for the request then reports EE_ERR_RUNTIME,
 * and the queue refuses every later submit. */
ee_status ee_queue_inject_refusal(ee_queue* queue, uint32_t graph, const ee_kopt_task* task, int64_t* costs,
                                  uint64_t* request_id);

#ifdef __cplusplus
}
#endif

#endif /* EE_HIP_H */
