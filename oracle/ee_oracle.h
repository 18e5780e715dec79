/*
 * CPU oracle for the JV fast-elimination edge test and the 3/4/5-opt exact
 * leaf cost matrix.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of the
 * reference CPU path (LinHuanli/cudaEdgeElimination).  It exists so that the
 * test suite, the smoke check and the benchmark's CPU-baseline leg can check
 * and time the HIP product path against an independent implementation.  The
 * product library never links, loads or calls it.
 *
 * Every entry point cites the reference function it restates.  All functions
 * return 0 on success and a negative value on failure unless stated
 * otherwise; the failure text is available from orc_last_error().
 */
#ifndef EE_ORACLE_H
#define EE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph orc_graph;

/* Same 48-byte layout as the reference KOptCostTask
 * (include/cuda_edge_elimination/local_search.hpp:54-57). */
typedef struct {
  int32_t port_nodes[10];
  int64_t deleted_cost;
} orc_kopt_task;

const char* orc_last_error(void);

/* src/common/distance.cpp:9-27 (restoring square root). */
uint64_t orc_isqrt_floor(uint64_t value);
/* src/common/distance.cpp:29-68.  type: 0 = EUC_2D, 1 = CEIL_2D. */
int orc_exact_distance(int64_t ax, int64_t ay, int64_t bx, int64_t by, int type, int64_t* out);

/* src/common/graph.cpp:98-228 (TSPLIB .tsp + Concorde .edg load). */
orc_graph* orc_graph_load(const char* tsp_path, const char* edge_path);
/* A hand-built snapshot with integer coordinates: edges are taken in the
 * given order (no sort, no weight check), then the CSR is rebuilt
 * (src/common/graph.cpp:230-282), as the reference unit tests build graphs
 * (tests/unit/test_core.cpp:86-98). */
orc_graph* orc_graph_from_arrays(int32_t dimension, int distance_type, const int64_t* x,
                                 const int64_t* y, int32_t edge_count, const int32_t* edge_u,
                                 const int32_t* edge_v, const int64_t* edge_weight,
                                 const uint8_t* edge_active);
orc_graph* orc_graph_clone(const orc_graph* graph);
void orc_graph_free(orc_graph* graph);

int32_t orc_graph_dimension(const orc_graph* graph);
int32_t orc_graph_edge_count(const orc_graph* graph);
int64_t orc_graph_adjacency_count(const orc_graph* graph);
int32_t orc_graph_active_edge_count(const orc_graph* graph);
/* distance_type, integer_coordinates, integer_distance_safe. */
void orc_graph_flags(const orc_graph* graph, int* distance_type, int* integer_coordinates,
                     int* integer_distance_safe);
/* Copies out any non-NULL array.  Sizes: x,y [n]; edge arrays [m];
 * row_offsets [n+1]; neighbors, csr_edge_ids, csr_weights [A]. */
void orc_graph_export(const orc_graph* graph, int64_t* x, int64_t* y, int32_t* edge_u,
                      int32_t* edge_v, int64_t* edge_weight, uint8_t* edge_active,
                      int32_t* row_offsets, int32_t* neighbors, int32_t* csr_edge_ids,
                      int64_t* csr_weights);
int orc_graph_set_active(orc_graph* graph, int32_t edge_id, int active);
/* src/common/graph.cpp:230-282. */
void orc_graph_rebuild_csr(orc_graph* graph);
/* src/common/graph.cpp:323-346 (FNV-1a 64 over the snapshot's member bytes). */
uint64_t orc_graph_hash(const orc_graph* graph);
/* src/common/graph.cpp:356-363. */
int orc_graph_has_active_edge(const orc_graph* graph, int32_t u, int32_t v);
/* src/common/graph.cpp:300-316. */
int orc_graph_distance(const orc_graph* graph, int32_t u, int32_t v, int64_t* out);

/* src/cpu/jv_elimination.cpp:209-224 with :45-117.  witnesses[m] receives the
 * first JV witness of each edge slot or -1.  Returns the candidate count.
 * distance_evaluations (nullable) receives the exact number of distance
 * evaluations performed, a deterministic work count. */
int64_t orc_jv_find_candidates(const orc_graph* graph, int32_t* witnesses,
                               uint64_t* distance_evaluations);
/* src/cpu/jv_elimination.cpp:226-248.  1 = verified, 0 = rejected. */
int orc_jv_verify(const orc_graph* graph, int32_t edge_id, int32_t witness);
/* src/cpu/elimination_commit.cpp:13-75.  Commits the candidates on `graph`
 * in place.  committed_edge_ids / committed_witnesses receive the committed
 * candidates in commit order; returns how many were committed. */
int64_t orc_jv_commit(orc_graph* graph, int64_t count, const int32_t* edge_ids,
                      const int32_t* witnesses, uint64_t expected_hash,
                      int32_t* committed_edge_ids, int32_t* committed_witnesses);
/* src/cpu/jv_elimination.cpp:250-322 (CPU backend) and :324-368 (V1 proof
 * text).  The propose/verify wall times of each metrics line are written as
 * measured (seconds of steady clock, in milliseconds).  Returns the number of
 * proof bytes (the text is truncated to capacity-1 and NUL-terminated when
 * proof_out is non-NULL); negative on error.  The graph is updated in place. */
int64_t orc_jv_run(orc_graph* graph, uint32_t max_rounds, const char* backend_name,
                   char* proof_out, size_t capacity);
/* src/cpu/jv_elimination.cpp:370-523 (V1 subset) and :525-605.  Replays the
 * proof on a copy of `graph`; on success the graph takes the final state. */
int orc_jv_replay(orc_graph* graph, const char* proof_text);

/* src/cpu/kopt_search.cpp:1143-1193 with src/cpu/path_system.cpp:40-71,
 * 377-421.  mates receives count x 14 bytes (unused entries 0xff). */
int orc_kopt_templates(uint32_t k, uint8_t* mates, uint32_t capacity, uint32_t* count,
                       uint64_t* generator_hash);
/* src/cpu/kopt_search.cpp:784-807, 865-1090, 1279-1302: the CPU exact cost
 * matrix, row-major [task][template].  max_threads caps the OpenMP team the
 * way kKOptCpuMaxThreads does (kopt_search.cpp:40, 1070-1071); pass 8 for the
 * reference policy or 0 for all host cores.  threads_used and
 * distance_cache_nodes (nullable) report what the reference reports in
 * KOptCostBatchResult. */
int orc_kopt_costs(const int64_t* x, const int64_t* y, int32_t dimension, int distance_type,
                   uint32_t k, const orc_kopt_task* tasks, size_t task_count, int64_t* costs,
                   int max_threads, uint32_t* threads_used, uint32_t* distance_cache_nodes);

/* The deletion sets of KOptSearchCursor (src/cpu/kopt_search.cpp:2123-2156,
 * 2258-2271): combinations of k selectable positions, advanced by
 * AdvanceCombination (kopt_search.cpp:736-752), which never moves the first
 * index, so every set holds selectable[0]; each set's positions sorted and
 * turned into a task by BuildKOptCostTask (kopt_search.cpp:561-584).
 * Writes the tasks of sets [first, first + count) and, when positions is not
 * NULL, their sorted tour positions (count x k).  Returns the number of sets
 * written (fewer than count where the enumeration ends; 0 when k < 3, k > 5
 * or k >= selectable_count, as PrepareNextK), negative on failure. */
int64_t orc_kopt_cursor_tasks(const int32_t* tour, const int64_t* edge_costs, uint32_t tour_length,
                              const uint64_t* selectable, uint32_t selectable_count, uint32_t k,
                              uint64_t first, uint64_t count, orc_kopt_task* tasks,
                              uint64_t* positions);

/* HT Hamilton replies of the reference (src/cpu/hamilton_tutte.cpp:230-262,
 * 79-87, 386-440 without the center deduplication, which the device entry
 * EvaluateHtHamiltonRepliesCuda does not do either): for each center in
 * order, every pair of its CSR neighbours (first before second) that is not
 * the target edge, whose two center edges each survive the 2-opt test
 * against the target, and with target + center edges <= the pair's edge +
 * the center-to-target-ends edges.  offsets receives count + 1 entries;
 * replies (triples center, first, second) receives min(total, capacity)
 * entries.  Returns the total, negative on failure (invalid target or
 * center, as ValidateTarget / the center checks). */
int64_t orc_ht_hamilton_replies(const orc_graph* graph, int32_t target_u, int32_t target_v,
                                const int32_t* centers, size_t count, uint64_t* offsets,
                                int32_t* replies, uint64_t capacity);
/* HT end replies (src/cpu/hamilton_tutte.cpp:264-277, 497-560 without the
 * task deduplication): per task (endpoint, internal neighbour) every CSR
 * neighbour of the endpoint except the internal one, as canonical (u < v)
 * edges.  tasks: count pairs; replies: pairs.  Returns the total, negative
 * on failure (a task that is not an active edge). */
int64_t orc_ht_end_replies(const orc_graph* graph, const int32_t* tasks, size_t count,
                           uint64_t* offsets, int32_t* replies, uint64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* EE_ORACLE_H */
