"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when it was
built, to the reference's own CPU path (oracle/_ref/libcudaee_ref.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import re

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"
REF_PATH = ROOT / "oracle" / "_ref" / "libcudaee_ref.so"

TASK_DTYPE = np.dtype([("port_nodes", np.int32, (10,)), ("deleted_cost", np.int64)], align=True)
P, I32, U32, I64, U64, SZ = C.c_void_p, C.c_int32, C.c_uint32, C.c_int64, C.c_uint64, C.c_size_t

_ORACLE = {
    "orc_last_error": (C.c_char_p, []),
    "orc_isqrt_floor": (U64, [U64]),
    "orc_exact_distance": (C.c_int, [I64, I64, I64, I64, C.c_int, C.POINTER(I64)]),
    "orc_graph_load": (P, [C.c_char_p, C.c_char_p]),
    "orc_graph_from_arrays": (P, [I32, C.c_int, P, P, I32, P, P, P, P]),
    "orc_graph_clone": (P, [P]),
    "orc_graph_free": (None, [P]),
    "orc_graph_dimension": (I32, [P]),
    "orc_graph_edge_count": (I32, [P]),
    "orc_graph_adjacency_count": (I64, [P]),
    "orc_graph_active_edge_count": (I32, [P]),
    "orc_graph_flags": (None, [P, P, P, P]),
    "orc_graph_export": (None, [P, P, P, P, P, P, P, P, P, P, P]),
    "orc_graph_set_active": (C.c_int, [P, I32, C.c_int]),
    "orc_graph_rebuild_csr": (None, [P]),
    "orc_graph_hash": (U64, [P]),
    "orc_graph_has_active_edge": (C.c_int, [P, I32, I32]),
    "orc_graph_distance": (C.c_int, [P, I32, I32, C.POINTER(I64)]),
    "orc_jv_find_candidates": (I64, [P, P, C.POINTER(U64)]),
    "orc_jv_verify": (C.c_int, [P, I32, I32]),
    "orc_jv_commit": (I64, [P, I64, P, P, U64, P, P]),
    "orc_jv_run": (I64, [P, U32, C.c_char_p, C.c_char_p, SZ]),
    "orc_jv_replay": (C.c_int, [P, C.c_char_p]),
    "orc_kopt_templates": (C.c_int, [U32, P, U32, C.POINTER(U32), C.POINTER(U64)]),
    "orc_kopt_costs": (C.c_int, [P, P, I32, C.c_int, U32, P, SZ, P, C.c_int, C.POINTER(U32),
                                 C.POINTER(U32)]),
    "orc_kopt_cursor_tasks": (I64, [P, P, U32, P, U32, U32, U64, U64, P, P]),
    "orc_ht_hamilton_replies": (I64, [P, I32, I32, P, SZ, P, P, U64]),
    "orc_ht_end_replies": (I64, [P, P, SZ, P, P, U64]),
}
_REF = {
    "ref_last_error": (C.c_char_p, []),
    "ref_isqrt_floor": (U64, [U64]),
    "ref_exact_distance": (C.c_int, [I64, I64, I64, I64, C.c_int, C.POINTER(I64)]),
    "ref_graph_hash": (U64, [I32, C.c_int, P, P, I32, P, P, P, P]),
    "ref_jv_candidates_files": (I64, [C.c_char_p, C.c_char_p, P, I32]),
    "ref_jv_candidates": (I64, [I32, C.c_int, P, P, I32, P, P, P, P, P]),
    "ref_jv_run_files": (I64, [C.c_char_p, C.c_char_p, U32, C.c_char_p, SZ]),
    "ref_jv_replay_files": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "ref_kopt_templates": (C.c_int, [U32, P, U32, C.POINTER(U32), C.POINTER(U64)]),
    "ref_kopt_costs": (C.c_int, [P, P, I32, C.c_int, U32, P, SZ, P, C.POINTER(U32), C.POINTER(U32)]),
    "ref_ht_hamilton_replies": (I64, [I32, C.c_int, P, P, I32, P, P, P, P, I32, I32, P, SZ, P, P,
                                      U64]),
    "ref_ht_end_replies": (I64, [I32, C.c_int, P, P, I32, P, P, P, P, P, SZ, P, P, U64]),
}


def _load(path, protos):
    lib = C.CDLL(str(path))
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = _load(ORACLE_PATH, _ORACLE)
    return _oracle


def ref_available() -> bool:
    return REF_PATH.exists()


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_PATH, _REF)
    return _ref


def p(a):
    return None if a is None else a.ctypes.data


class OracleError(RuntimeError):
    pass


def _raise(lib_err):
    raise OracleError(lib_err().decode())


# ------------------------------------------------------------------ graphs
class OGraph:
    """Owning handle of an oracle snapshot."""

    def __init__(self, handle):
        if not handle:
            _raise(oracle().orc_last_error)
        self.h = handle

    @classmethod
    def load(cls, tsp, edg):
        return cls(oracle().orc_graph_load(str(tsp).encode(), str(edg).encode()))

    @classmethod
    def from_snapshot(cls, g):
        """From a cudaedgeelimination_amd.GraphSnapshot (edges in its order)."""
        return cls(oracle().orc_graph_from_arrays(
            g.dimension, g.distance_type, p(g.x), p(g.y), g.edge_count, p(g.edge_u), p(g.edge_v),
            p(g.edge_weight), p(g.edge_active)))

    @classmethod
    def from_arrays(cls, x, y, dtype, u, v, w, active=None):
        x, y = np.ascontiguousarray(x, np.int64), np.ascontiguousarray(y, np.int64)
        u, v = np.ascontiguousarray(u, np.int32), np.ascontiguousarray(v, np.int32)
        w = np.ascontiguousarray(w, np.int64)
        a = None if active is None else np.ascontiguousarray(active, np.uint8)
        return cls(oracle().orc_graph_from_arrays(len(x), dtype, p(x), p(y), len(u), p(u), p(v),
                                                  p(w), p(a)))

    def __del__(self):
        if getattr(self, "h", None):
            oracle().orc_graph_free(self.h)
            self.h = None

    def clone(self):
        return OGraph(oracle().orc_graph_clone(self.h))

    @property
    def n(self):
        return oracle().orc_graph_dimension(self.h)

    @property
    def m(self):
        return oracle().orc_graph_edge_count(self.h)

    def arrays(self):
        lib = oracle()
        n, m, a = self.n, self.m, lib.orc_graph_adjacency_count(self.h)
        out = dict(x=np.empty(n, np.int64), y=np.empty(n, np.int64), edge_u=np.empty(m, np.int32),
                   edge_v=np.empty(m, np.int32), edge_weight=np.empty(m, np.int64),
                   edge_active=np.empty(m, np.uint8), row_offsets=np.empty(n + 1, np.int32),
                   neighbors=np.empty(a, np.int32), csr_edge_ids=np.empty(a, np.int32),
                   csr_weights=np.empty(a, np.int64))
        lib.orc_graph_export(self.h, *[p(out[k]) for k in (
            "x", "y", "edge_u", "edge_v", "edge_weight", "edge_active", "row_offsets",
            "neighbors", "csr_edge_ids", "csr_weights")])
        dt = C.c_int()
        lib.orc_graph_flags(self.h, C.byref(dt), None, None)
        out["distance_type"] = dt.value
        return out

    def hash(self):
        return int(oracle().orc_graph_hash(self.h))

    def distance(self, u, v):
        out = C.c_int64()
        if oracle().orc_graph_distance(self.h, u, v, C.byref(out)) != 0:
            _raise(oracle().orc_last_error)
        return out.value

    def witnesses(self):
        w = np.empty(self.m, np.int32)
        evals = C.c_uint64()
        if oracle().orc_jv_find_candidates(self.h, p(w), C.byref(evals)) < 0:
            _raise(oracle().orc_last_error)
        return w

    def verify(self, edge_id, witness):
        return oracle().orc_jv_verify(self.h, edge_id, witness) == 1

    def set_active(self, edge_id, active):
        oracle().orc_graph_set_active(self.h, edge_id, 1 if active else 0)

    def rebuild_csr(self):
        oracle().orc_graph_rebuild_csr(self.h)


def jv_run(graph: OGraph, max_rounds=100, backend="cpu") -> str:
    """RunJvElimination(kCpu) + WriteProof V1 on a clone; returns the text."""
    lib = oracle()
    capacity = 1 << 20
    while True:
        g = graph.clone()
        buf = C.create_string_buffer(capacity)
        size = lib.orc_jv_run(g.h, max_rounds, backend.encode(), buf, len(buf))
        if size < 0:
            _raise(lib.orc_last_error)
        if size < len(buf):
            return buf.value.decode()
        capacity = size + 1


def jv_replay(graph: OGraph, proof: str) -> None:
    g = graph.clone()
    if oracle().orc_jv_replay(g.h, proof.encode()) != 0:
        _raise(oracle().orc_last_error)


def canonical_proof(text: str) -> str:
    """Proof text without the run-specific fields: the backend name and the
    two wall-time columns of each metrics line
    (src/cpu/jv_elimination.cpp:331-347)."""
    out = []
    for line in text.splitlines():
        if line.startswith("backend "):
            out.append("backend *")
        elif line.startswith("metrics "):
            out.append(" ".join(line.split()[:7]))
        else:
            out.append(line)
    return "\n".join(out) + "\n"


# ------------------------------------------------------------------- k-opt
def templates(k, lib_name="orc"):
    lib = oracle() if lib_name == "orc" else ref()
    mates = np.zeros((945, 14), np.uint8)
    count, h = C.c_uint32(), C.c_uint64()
    fn = lib.orc_kopt_templates if lib_name == "orc" else lib.ref_kopt_templates
    if fn(k, p(mates), 945, C.byref(count), C.byref(h)) != 0:
        raise OracleError("templates")
    return mates[:count.value].copy(), int(h.value)


def kopt_costs(x, y, dtype, k, tasks, max_threads=8):
    x, y = np.ascontiguousarray(x, np.int64), np.ascontiguousarray(y, np.int64)
    tasks = np.ascontiguousarray(tasks, TASK_DTYPE)
    T = {3: 4, 4: 25, 5: 208}[k]
    out = np.empty(len(tasks) * T, np.int64)
    threads, nodes = C.c_uint32(), C.c_uint32()
    lib = oracle()
    if lib.orc_kopt_costs(p(x), p(y), len(x), dtype, k, p(tasks) if len(tasks) else None,
                          len(tasks), p(out) if len(out) else None, max_threads,
                          C.byref(threads), C.byref(nodes)) != 0:
        _raise(lib.orc_last_error)
    return out, threads.value, nodes.value


def cursor_tasks(tour, edge_costs, selectable, k, first, count):
    """Oracle deletion-set cursor: (tasks, positions[rows, k])."""
    tour = np.ascontiguousarray(tour, np.int32)
    edge_costs = np.ascontiguousarray(edge_costs, np.int64)
    selectable = np.ascontiguousarray(selectable, np.uint64)
    tasks = np.zeros(max(count, 1), TASK_DTYPE)
    positions = np.zeros((max(count, 1), k), np.uint64)
    lib = oracle()
    n = lib.orc_kopt_cursor_tasks(p(tour), p(edge_costs), len(tour),
                                  p(selectable) if len(selectable) else None, len(selectable), k,
                                  first, count, p(tasks), p(positions))
    if n < 0:
        _raise(lib.orc_last_error)
    return tasks[:n], positions[:n]


def _ht_call(fn, err, lead, items, width, bound):
    items = np.ascontiguousarray(items, np.int32)
    count = len(items) if items.ndim == 1 else len(items)
    offsets = np.zeros(count + 1, np.uint64)
    replies = np.zeros((max(bound, 1), width), np.int32)
    total = fn(*lead, p(items) if count else None, count, p(offsets), p(replies), bound)
    if total < 0:
        _raise(err)
    assert total <= bound
    return offsets, replies[:total].copy()


def ht_hamilton_replies(graph: OGraph, target, centers, bound):
    lib = oracle()
    return _ht_call(lib.orc_ht_hamilton_replies, lib.orc_last_error,
                    (graph.h, int(target[0]), int(target[1])), centers, 3, bound)


def ht_end_replies(graph: OGraph, tasks, bound):
    lib = oracle()
    return _ht_call(lib.orc_ht_end_replies, lib.orc_last_error, (graph.h,),
                    np.asarray(tasks, np.int32).reshape(-1, 2), 2, bound)


def _ref_graph_args(x, y, dtype, u, v, w, active):
    arrs = [np.ascontiguousarray(a, t) for a, t in ((x, np.int64), (y, np.int64), (u, np.int32),
                                                     (v, np.int32), (w, np.int64),
                                                     (active, np.uint8))]
    return arrs, (len(arrs[0]), dtype, p(arrs[0]), p(arrs[1]), len(arrs[2]), p(arrs[2]),
                  p(arrs[3]), p(arrs[4]), p(arrs[5]))


def ref_ht_hamilton_replies(x, y, dtype, u, v, w, active, target, centers, bound):
    lib = ref()
    keep, lead = _ref_graph_args(x, y, dtype, u, v, w, active)
    return _ht_call(lib.ref_ht_hamilton_replies, lib.ref_last_error,
                    lead + (int(target[0]), int(target[1])), centers, 3, bound)


def ref_ht_end_replies(x, y, dtype, u, v, w, active, tasks, bound):
    lib = ref()
    keep, lead = _ref_graph_args(x, y, dtype, u, v, w, active)
    return _ht_call(lib.ref_ht_end_replies, lib.ref_last_error, lead,
                    np.asarray(tasks, np.int32).reshape(-1, 2), 2, bound)


def ref_kopt_costs(x, y, dtype, k, tasks):
    x, y = np.ascontiguousarray(x, np.int64), np.ascontiguousarray(y, np.int64)
    tasks = np.ascontiguousarray(tasks, TASK_DTYPE)
    T = {3: 4, 4: 25, 5: 208}[k]
    out = np.empty(len(tasks) * T, np.int64)
    threads, nodes = C.c_uint32(), C.c_uint32()
    lib = ref()
    if lib.ref_kopt_costs(p(x), p(y), len(x), dtype, k, p(tasks), len(tasks), p(out),
                          C.byref(threads), C.byref(nodes)) != 0:
        _raise(lib.ref_last_error)
    return out, threads.value, nodes.value


def edge_list_cost(dist, k, ports, mate):
    """Independent per-cell oracle with plain edge lists (the reference's test
    oracle ReferenceKOptTemplateCost, tests/unit/test_kopt.cpp:48-74)."""
    canon = lambda a, b: (a, b) if a < b else (b, a)  # noqa: E731
    deleted = [canon(ports[2 * e], ports[2 * e + 1]) for e in range(k)]
    added, total = [], 0
    for port in range(2 * k):
        partner = mate[port]
        if port >= partner:
            continue
        edge = canon(ports[port], ports[partner])
        if edge[0] == edge[1] or edge in deleted or edge in added:
            return (1 << 63) - 1
        added.append(edge)
        total += dist(edge[0], edge[1])
    if len(added) != k or total > (1 << 63) - 1:
        return (1 << 63) - 1
    return total


def header_functions(path) -> list[str]:
    """Function names declared in a C header."""
    text = re.sub(r"/\*.*?\*/", "", pathlib.Path(path).read_text(), flags=re.S)
    names = re.findall(r"\b(ee_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))
