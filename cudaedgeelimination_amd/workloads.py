"""Seeded synthetic workloads for the five benchmark configurations.

1. plain JV: n = 300 uniform in [0, 5000]^2, EUC_2D, 12-NN union graph;
2. multi-epoch JV: 60 x 60 lattice, spacing 25, 17% of sites dropped, +-3
   jitter, EUC_2D, 10-NN union graph;
3. k-opt matrix: n = 15,000 uniform in [0, 10^6]^2, 8-NN; 65,536 rows with
   k = 3 / 4 / 5 quotas 32,768 / 19,661 / 13,107, shuffled; every row's
   deleted edges are k distinct edges of a random simple walk, in walk order;
4. many instances: 256 graphs, n log-uniform in [200, 4000], uniform or
   8-cluster Gaussian coordinates (sigma = 20,000 in a 10^6 box), CEIL_2D for
   odd instance indices, 10-NN; 4,096 rows each (2,048 / 1,229 / 819);
5. served queries: resident graphs with n = 1k, 5k, 15k, 30k, uniform, 10-NN.

Every input is a pure function of the seed (native generator, ee_gen.h).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib
from .graph import CEIL_2D, EUC_2D, GraphSnapshot, exact_distances

SEEDS = {1: 0xEE000001, 2: 0xEE000002, 3: 0xEE000003, 4: 0xEE000004, 5: 0xEE000005}
KOPT_QUOTA_FRACTIONS = (0.5, 0.3, 0.2)


def _points_uniform(seed: int, n: int, lo: int, hi: int):
    x = np.empty(n, dtype=np.int64)
    y = np.empty(n, dtype=np.int64)
    _lib.check(_lib.lib().ee_gen_points_uniform(seed, n, lo, hi, _lib.ptr(x), _lib.ptr(y)))
    return x, y


def _points_clustered(seed: int, n: int, clusters: int, sigma: float, lo: int, hi: int):
    x = np.empty(n, dtype=np.int64)
    y = np.empty(n, dtype=np.int64)
    _lib.check(_lib.lib().ee_gen_points_clustered(seed, n, clusters, sigma, lo, hi,
                                                  _lib.ptr(x), _lib.ptr(y)))
    return x, y


def _points_lattice(seed: int, side: int, spacing: int, drop: float, jitter: int):
    x = np.empty(side * side, dtype=np.int64)
    y = np.empty(side * side, dtype=np.int64)
    n = C.c_int32()
    _lib.check(_lib.lib().ee_gen_points_lattice(seed, side, side, spacing, drop, jitter,
                                                C.byref(n), _lib.ptr(x), _lib.ptr(y)))
    return x[:n.value].copy(), y[:n.value].copy()


def knn_graph(x, y, distance_type: int, k: int) -> GraphSnapshot:
    x = np.ascontiguousarray(x, dtype=np.int64)
    y = np.ascontiguousarray(y, dtype=np.int64)
    count = C.c_int64()
    lib = _lib.lib()
    _lib.check(lib.ee_gen_knn_graph(_lib.ptr(x), _lib.ptr(y), len(x), distance_type, k,
                                    None, None, None, 0, C.byref(count)))
    m = count.value
    u = np.empty(m, dtype=np.int32)
    v = np.empty(m, dtype=np.int32)
    w = np.empty(m, dtype=np.int64)
    _lib.check(lib.ee_gen_knn_graph(_lib.ptr(x), _lib.ptr(y), len(x), distance_type, k,
                                    _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), m, C.byref(count)))
    return GraphSnapshot.from_edges(x, y, distance_type, u, v, w)


def kopt_quota(rows: int) -> tuple[int, int, int]:
    """Rows per k in {3, 4, 5} for a 50/30/20 mix (k = 3 takes the remainder)."""
    k4 = int(round(rows * KOPT_QUOTA_FRACTIONS[1]))
    k5 = int(round(rows * KOPT_QUOTA_FRACTIONS[2]))
    return rows - k4 - k5, k4, k5


def shuffled_ks(seed: int, quota) -> np.ndarray:
    q = np.asarray(quota, dtype=np.uint64)
    ks = np.empty(int(q.sum()), dtype=np.uint8)
    _lib.check(_lib.lib().ee_gen_shuffled_ks(seed, _lib.ptr(q), _lib.ptr(ks)))
    return ks


def kopt_tasks(seed: int, graph: GraphSnapshot, ks) -> np.ndarray:
    ks = np.ascontiguousarray(ks, dtype=np.uint8)
    tasks = np.zeros(len(ks), dtype=_lib.KOPT_TASK_DTYPE)
    _lib.check(_lib.lib().ee_gen_kopt_tasks(
        seed, _lib.ptr(graph.x), _lib.ptr(graph.y), graph.dimension, graph.distance_type,
        _lib.ptr(graph.row_offsets), _lib.ptr(graph.neighbors), 0, _lib.ptr(ks), len(ks),
        _lib.ptr(tasks)))
    return tasks


def tour_context(seed: int, graph: GraphSnapshot, length: int, selectable: int,
                 required: bool = True):
    """A tour context of the shape BuildTourContext makes
    (src/cpu/kopt_search.cpp:270-350): a cycle over `length` distinct nodes
    of the snapshot whose `selectable` path-edge positions carry their exact
    cost (-1 on the other positions, the outside-matching edges), listed in
    tour order and, with `required`, rotated so that a random one (the
    required edge) comes first.  numpy's PCG64 stream of `seed` picks the
    nodes and positions."""
    from .backend import TourContext
    if not 3 <= length <= graph.dimension or not 0 <= selectable <= length:
        raise ValueError("tour_context: 3 <= length <= n and selectable <= length")
    rng = np.random.default_rng(seed)
    tour = rng.permutation(graph.dimension)[:length].astype(np.int32)
    picked = np.sort(rng.permutation(length)[:selectable]).astype(np.uint64)
    if required and selectable:
        picked = np.roll(picked, -int(rng.integers(selectable)))
    nxt = np.roll(tour, -1)
    costs = np.full(length, -1, dtype=np.int64)
    at = picked.astype(np.int64)
    costs[at] = exact_distances(graph.x[tour[at]], graph.y[tour[at]], graph.x[nxt[at]],
                                graph.y[nxt[at]], graph.distance_type)
    return TourContext(tour, costs, picked)


@dataclasses.dataclass
class KoptWorkload:
    graph: GraphSnapshot
    ks: np.ndarray      # k of every row, in production order
    tasks: np.ndarray   # KOPT_TASK_DTYPE rows, in production order

    def bucket(self, k: int) -> np.ndarray:
        """Rows of one k in production order (the reference evaluates one
        matrix per k, src/cpu/kopt_search.cpp:2619-2651)."""
        return self.tasks[self.ks == k]

    def cells(self) -> int:
        return sum(int(np.count_nonzero(self.ks == k)) * t for k, t in ((3, 4), (4, 25), (5, 208)))


def config1(seed: int = SEEDS[1], n: int = 300) -> GraphSnapshot:
    x, y = _points_uniform(seed, n, 0, 5000)
    return knn_graph(x, y, EUC_2D, 12)


def config2(seed: int = SEEDS[2], side: int = 60) -> GraphSnapshot:
    x, y = _points_lattice(seed, side, 25, 0.17, 3)
    return knn_graph(x, y, EUC_2D, 10)


def config3(seed: int = SEEDS[3], n: int = 15000, rows: int = 65536) -> KoptWorkload:
    x, y = _points_uniform(seed, n, 0, 1_000_000)
    graph = knn_graph(x, y, EUC_2D, 8)
    ks = shuffled_ks(seed ^ 0x5151, kopt_quota(rows))
    return KoptWorkload(graph, ks, kopt_tasks(seed ^ 0xA5A5, graph, ks))


def config4_instance(index: int, seed: int = SEEDS[4], rows: int = 4096,
                     n_range: tuple[int, int] = (200, 4000)):
    """One instance of the many-instance batch: (graph, KoptWorkload).  n is
    log-uniform in n_range; instances alternate uniform / 8-cluster Gaussian
    coordinates in pairs, and odd indices use CEIL_2D."""
    rng = np.random.Generator(np.random.PCG64(seed + index))
    lo, hi = n_range
    n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    sub = seed * 1000003 + index
    if (index // 2) % 2 == 0:
        x, y = _points_uniform(sub, n, 0, 1_000_000)
    else:
        x, y = _points_clustered(sub, n, 8, 20_000.0, 0, 1_000_000)
    distance_type = CEIL_2D if index % 2 == 1 else EUC_2D
    graph = knn_graph(x, y, distance_type, 10)
    ks = shuffled_ks(sub ^ 0x5151, (2048 * rows // 4096, 1229 * rows // 4096,
                                    rows - 2048 * rows // 4096 - 1229 * rows // 4096))
    return graph, KoptWorkload(graph, ks, kopt_tasks(sub ^ 0xA5A5, graph, ks))


def config5_graphs(seed: int = SEEDS[5], sizes=(1000, 5000, 15000, 30000)):
    graphs = []
    for i, n in enumerate(sizes):
        x, y = _points_uniform(seed + i, n, 0, 1_000_000)
        graphs.append(knn_graph(x, y, EUC_2D, 10))
    return graphs


SERVED_TEMPLATES = {3: 4, 4: 25, 5: 208}


def config5_row_pools(graphs, seed: int = SEEDS[5], rows: int = 2048):
    """Per served graph and k, `rows` task rows on random walks of the graph
    (the rows that leaf-cost queries take contiguous slices of)."""
    pools = []
    for i, g in enumerate(graphs):
        pools.append({k: kopt_tasks(seed + 17 * i + k, g, np.full(rows, k, np.uint8))
                      for k in (3, 4, 5)})
    return pools


def config5_requests(graphs, pools, count: int, rng, jv_share: float = 0.2):
    """A stream of served requests: (graph index, k, rows) leaf-cost queries
    whose size is lognormal in cells (median 2,896, sigma 1.04) clipped to
    [128, 65,536], k = 3 / 4 / 5 at 50/30/20%, rows a contiguous slice of the
    graph's pool; and, with probability `jv_share`, (graph index, 0, edge
    slots) JV rechecks of lognormally many (median 256, sigma 1.0, clipped to
    [16, 4,096]) distinct random edge slots."""
    cells = np.clip(np.exp(rng.normal(np.log(2896.0), 1.04, count)), 128, 65536)
    kinds = rng.choice([3, 4, 5], size=count, p=[0.5, 0.3, 0.2])
    which = rng.integers(0, len(graphs), count)
    is_jv = rng.random(count) < jv_share
    jv_sizes = np.clip(np.exp(rng.normal(np.log(256.0), 1.0, count)), 16, 4096).astype(int)
    requests = []
    for i in range(count):
        g = int(which[i])
        if is_jv[i]:
            m = graphs[g].edge_count
            requests.append((g, 0, rng.choice(m, size=min(int(jv_sizes[i]), m),
                                              replace=False).astype(np.int32)))
            continue
        k = int(kinds[i])
        rows = max(1, int(np.ceil(cells[i] / SERVED_TEMPLATES[k])))
        pool = pools[g][k]
        rows = min(rows, len(pool))
        start = int(rng.integers(0, len(pool) - rows + 1))
        requests.append((g, k, pool[start:start + rows]))
    return requests, is_jv

