/*
 * ee_gen.h - seeded synthetic workloads and host graph helpers of the gfx950
 * backend.  These have no counterpart among the reference's CUDA entry
 * points: the reference reads TSPLIB files (src/common/graph.cpp:98-228) and
 * produces k-opt task rows inside its HT search.  They exist so that
 * benchmarks, tests and the serving loop can build the same inputs anywhere
 * from a seed.
 *
 * Randomness: xoshiro256** seeded through SplitMix64, unbiased bounded
 * integers by Lemire's multiply-shift rejection method, Gaussians by
 * Box-Muller in FP64 rounded with llround.  Identical seeds give identical
 * outputs on every host with the same C library.
 */
#ifndef EE_GEN_H
#define EE_GEN_H

#include "ee_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* n distinct points with integer coordinates uniform in [lo, hi]^2. */
ee_status ee_gen_points_uniform(uint64_t seed, int32_t n, int64_t lo, int64_t hi, int64_t* x,
                                int64_t* y);

/* A side_x x side_y lattice with the given spacing; exactly
 * round(drop_fraction * sites) sites are dropped at random, every kept site
 * is jittered by an integer uniform in [-jitter, jitter] per axis.  Kept
 * sites are numbered in row-major order.  x, y need side_x * side_y entries;
 * *n_out receives the number of points. */
ee_status ee_gen_points_lattice(uint64_t seed, int32_t side_x, int32_t side_y, int64_t spacing,
                                double drop_fraction, int64_t jitter, int32_t* n_out, int64_t* x,
                                int64_t* y);

/* n distinct points drawn from `clusters` isotropic Gaussians (standard
 * deviation sigma) whose centres are uniform in [lo, hi]^2; coordinates are
 * rounded and clamped to [lo, hi]. */
ee_status ee_gen_points_clustered(uint64_t seed, int32_t n, int32_t clusters, double sigma,
                                  int64_t lo, int64_t hi, int64_t* x, int64_t* y);

/* Symmetric union of every node's k nearest neighbours (ties broken by
 * squared distance, then node id), canonical (u < v), sorted by (u, v), with
 * exact distance weights.  With edge_u, edge_v and edge_weight all NULL the
 * call only counts the edges (EE_OK); otherwise it returns
 * EE_ERR_INVALID_ARGUMENT when `capacity` is too small.  *edge_count
 * receives the number of edges either way. */
ee_status ee_gen_knn_graph(const int64_t* x, const int64_t* y, int32_t n, uint8_t distance_type,
                           int32_t k, int32_t* edge_u, int32_t* edge_v, int64_t* edge_weight,
                           int64_t capacity, int64_t* edge_count);

/* CSR of the active edges exactly as the reference RebuildCsr lays it out
 * (src/common/graph.cpp:230-282): rows filled in edge-id order from both
 * endpoints, then sorted by (neighbour, edge id) unless the edge list is
 * canonically sorted.  row_offsets [n+1]; neighbors, csr_edge_ids
 * [2 x active edges]. */
ee_status ee_graph_build_csr(int32_t n, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                             const uint8_t* edge_active, int32_t* row_offsets, int32_t* neighbors,
                             int32_t* csr_edge_ids);

/* FNV-1a 64 snapshot hash with the reference's byte layout
 * (src/common/graph.cpp:323-346) for integer-coordinate graphs. */
uint64_t ee_graph_content_hash(int32_t n, uint8_t distance_type, const int64_t* x,
                               const int64_t* y, int32_t m, const int32_t* edge_u,
                               const int32_t* edge_v, const int64_t* edge_weight,
                               const uint8_t* edge_active);

/* `count` k-opt task rows whose deleted edges are k distinct edges, in walk
 * order, of a random simple walk of k..2k edges on the CSR graph; port
 * 2e / 2e+1 are the walk nodes of the e-th chosen edge, deleted_cost is the
 * sum of their exact distances, unused ports are 0.  ks (may be NULL) gives a
 * per-row k in [3, 5]; otherwise every row uses `k`. */
ee_status ee_gen_kopt_tasks(uint64_t seed, const int64_t* x, const int64_t* y, int32_t n,
                            uint8_t distance_type, const int32_t* row_offsets,
                            const int32_t* neighbors, uint32_t k, const uint8_t* ks, uint64_t count,
                            ee_kopt_task* tasks);

/* floor(sqrt(s)) for each s with the same FP64-root-plus-correction code the
 * kernels use, evaluated on the host (ee_isqrt_host) or on the context's
 * device (ee_isqrt_device, synchronous, count <= 2^28). */
void ee_isqrt_host(const uint64_t* s, uint64_t* root, size_t count);
ee_status ee_isqrt_device(ee_ctx* ctx, const uint64_t* s, uint64_t* root, size_t count);

/* Exact distances of node pairs (a[i], b[i]) on the context's device
 * (synchronous); coordinates must be distance safe.  When both coordinate
 * extents are below 2^25 this runs the FP64 distance the k-opt kernel uses
 * for such instances, otherwise the 64-bit integer one. */
ee_status ee_distances_device(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t n,
                              uint8_t distance_type, const int32_t* a, const int32_t* b,
                              size_t count, int64_t* out);

/* A random arrangement of quota[0] 3s, quota[1] 4s and quota[2] 5s. */
ee_status ee_gen_shuffled_ks(uint64_t seed, const uint64_t quota[3], uint8_t* ks);

#ifdef __cplusplus
}
#endif

#endif /* EE_GEN_H */
