"""Host container of a graph snapshot with the reference's field layout
(``GraphSnapshot``, include/cuda_edge_elimination/graph.hpp:13-37).

Only integer-coordinate snapshots are represented: the HIP paths, like the
reference's CUDA paths (src/cuda/jv_cuda.cu:493-495), refuse anything else.
CSR construction and the content hash run in the native library.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import pathlib

import numpy as np

from . import _lib

EUC_2D = 0
CEIL_2D = 1
DISTANCE_NAMES = {"EUC_2D": EUC_2D, "CEIL_2D": CEIL_2D}


def exact_distances(ax, ay, bx, by, distance_type: int) -> np.ndarray:
    """Vectorised exact TSPLIB distance (src/common/distance.cpp:29-68).

    floor(sqrt(S)) from the FP64 root with an integer correction; exact while
    S < 2^63, which every distance-safe graph in this package satisfies.
    """
    dx = np.abs(np.asarray(ax, dtype=np.int64) - np.asarray(bx, dtype=np.int64)).astype(np.uint64)
    dy = np.abs(np.asarray(ay, dtype=np.int64) - np.asarray(by, dtype=np.int64)).astype(np.uint64)
    s = dx * dx + dy * dy
    r = np.floor(np.sqrt(s.astype(np.float64))).astype(np.uint64)
    r = np.where(r * r > s, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= s, r + 1, r)
    if distance_type == EUC_2D:
        up = (s - r * r) > r
    else:
        up = (r * r) != s
    return (r + up.astype(np.uint64)).astype(np.int64)


@dataclasses.dataclass
class GraphSnapshot:
    dimension: int
    distance_type: int
    x: np.ndarray
    y: np.ndarray
    edge_u: np.ndarray
    edge_v: np.ndarray
    edge_weight: np.ndarray
    edge_active: np.ndarray
    row_offsets: np.ndarray = dataclasses.field(default=None)
    neighbors: np.ndarray = dataclasses.field(default=None)
    csr_edge_ids: np.ndarray = dataclasses.field(default=None)
    integer_coordinates: bool = True

    # ------------------------------------------------------------ builders
    @classmethod
    def from_edges(cls, x, y, distance_type, edge_u, edge_v, edge_weight=None, edge_active=None):
        x = np.ascontiguousarray(x, dtype=np.int64)
        y = np.ascontiguousarray(y, dtype=np.int64)
        u = np.ascontiguousarray(edge_u, dtype=np.int32)
        v = np.ascontiguousarray(edge_v, dtype=np.int32)
        if edge_weight is None:
            edge_weight = exact_distances(x[u], y[u], x[v], y[v], distance_type)
        w = np.ascontiguousarray(edge_weight, dtype=np.int64)
        a = (np.ones(len(u), dtype=np.uint8) if edge_active is None
             else np.ascontiguousarray(edge_active, dtype=np.uint8))
        g = cls(int(len(x)), int(distance_type), x, y, u, v, w, a)
        g.rebuild_csr()
        return g

    @classmethod
    def load(cls, tsp_path, edge_path):
        """TSPLIB .tsp + Concorde .edg (src/common/graph.cpp:98-228), integer
        coordinates only: edges normalised to u < v, duplicates rejected,
        weights must equal the exact distance, sorted by (u, v)."""
        dimension = 0
        distance_type = EUC_2D
        coords = None
        in_coords = False
        for raw in pathlib.Path(tsp_path).read_text().splitlines():
            line = raw.strip()
            if not line:
                continue
            if not in_coords:
                if line == "NODE_COORD_SECTION":
                    if dimension <= 0:
                        raise ValueError("DIMENSION missing before coordinates")
                    coords = [None] * dimension
                    in_coords = True
                    continue
                if ":" not in line:
                    continue
                key, value = (part.strip() for part in line.split(":", 1))
                if key == "DIMENSION":
                    dimension = int(value)
                elif key == "EDGE_WEIGHT_TYPE":
                    if value not in DISTANCE_NAMES:
                        raise ValueError(f"unsupported EDGE_WEIGHT_TYPE {value}")
                    distance_type = DISTANCE_NAMES[value]
                continue
            if line == "EOF":
                break
            ident, sx, sy = line.split()[:3]
            fx, fy = float(sx), float(sy)
            if not (fx.is_integer() and fy.is_integer()):
                raise ValueError("only integer coordinates are supported")
            index = int(ident) - 1
            if not 0 <= index < dimension or coords[index] is not None:
                raise ValueError("node id out of range or repeated")
            coords[index] = (int(fx), int(fy))
        if coords is None or any(c is None for c in coords):
            raise ValueError("incomplete coordinates")
        x = np.array([c[0] for c in coords], dtype=np.int64)
        y = np.array([c[1] for c in coords], dtype=np.int64)
        tokens = pathlib.Path(edge_path).read_text().split()
        if int(tokens[0]) != dimension:
            raise ValueError("edge file dimension mismatch")
        count = int(tokens[1])
        body = np.array(tokens[2:2 + 3 * count], dtype=np.int64).reshape(count, 3)
        if len(tokens) != 2 + 3 * count:
            raise ValueError("edge file has extra or missing data")
        u = np.minimum(body[:, 0], body[:, 1])
        v = np.maximum(body[:, 0], body[:, 1])
        if (u < 0).any() or (v >= dimension).any() or (u == v).any() or (body[:, 2] < 0).any():
            raise ValueError("edge out of range, self loop or negative weight")
        keys = u * dimension + v
        if len(np.unique(keys)) != count:
            raise ValueError("duplicate edge")
        if not np.array_equal(body[:, 2], exact_distances(x[u], y[u], x[v], y[v], distance_type)):
            raise ValueError("edge weight differs from the exact distance")
        order = np.lexsort((v, u))
        return cls.from_edges(x, y, distance_type, u[order], v[order], body[order, 2])

    def write(self, tsp_path, edge_path, name="ee-instance"):
        """Writes the active edges in the reference's input formats."""
        kind = "EUC_2D" if self.distance_type == EUC_2D else "CEIL_2D"
        lines = [f"NAME : {name}", "TYPE : TSP", f"DIMENSION : {self.dimension}",
                 f"EDGE_WEIGHT_TYPE : {kind}", "NODE_COORD_SECTION"]
        lines += [f"{i + 1} {int(a)} {int(b)}" for i, (a, b) in enumerate(zip(self.x, self.y))]
        lines.append("EOF")
        pathlib.Path(tsp_path).write_text("\n".join(lines) + "\n")
        keep = self.edge_active != 0
        rows = [f"{a} {b} {w}" for a, b, w in
                zip(self.edge_u[keep], self.edge_v[keep], self.edge_weight[keep])]
        pathlib.Path(edge_path).write_text(f"{self.dimension} {len(rows)}\n" + "\n".join(rows) + "\n")

    # ------------------------------------------------------------- queries
    @property
    def edge_count(self) -> int:
        return int(len(self.edge_u))

    @property
    def adjacency_count(self) -> int:
        return int(self.row_offsets[-1])

    def active_edge_count(self) -> int:
        return int(np.count_nonzero(self.edge_active))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets)

    def rebuild_csr(self) -> None:
        """src/common/graph.cpp:230-282, in the native library."""
        n, m = self.dimension, self.edge_count
        active = int(np.count_nonzero(self.edge_active))
        self.row_offsets = np.zeros(n + 1, dtype=np.int32)
        self.neighbors = np.zeros(2 * active, dtype=np.int32)
        self.csr_edge_ids = np.zeros(2 * active, dtype=np.int32)
        _lib.check(_lib.lib().ee_graph_build_csr(
            n, m, _lib.ptr(self.edge_u), _lib.ptr(self.edge_v), _lib.ptr(self.edge_active),
            _lib.ptr(self.row_offsets), _lib.ptr(self.neighbors), _lib.ptr(self.csr_edge_ids)))

    def content_hash(self) -> int:
        """src/common/graph.cpp:323-346."""
        return int(_lib.lib().ee_graph_content_hash(
            self.dimension, self.distance_type, _lib.ptr(self.x), _lib.ptr(self.y),
            self.edge_count, _lib.ptr(self.edge_u), _lib.ptr(self.edge_v),
            _lib.ptr(self.edge_weight), _lib.ptr(self.edge_active)))

    def view(self) -> _lib.GraphView:
        v = _lib.GraphView()
        v.dimension = self.dimension
        v.distance_type = self.distance_type
        v.integer_coordinates = 1 if self.integer_coordinates else 0
        v.x, v.y = _lib.ptr(self.x), _lib.ptr(self.y)
        v.edge_count = self.edge_count
        v.edge_u, v.edge_v = _lib.ptr(self.edge_u), _lib.ptr(self.edge_v)
        v.edge_weight, v.edge_active = _lib.ptr(self.edge_weight), _lib.ptr(self.edge_active)
        v.row_offsets = _lib.ptr(self.row_offsets)
        v.neighbors = _lib.ptr(self.neighbors) if len(self.neighbors) else None
        v.csr_edge_ids = _lib.ptr(self.csr_edge_ids) if len(self.csr_edge_ids) else None
        return v

    def copy(self) -> "GraphSnapshot":
        return dataclasses.replace(self, **{f.name: (getattr(self, f.name).copy()
                                                     if isinstance(getattr(self, f.name), np.ndarray)
                                                     else getattr(self, f.name))
                                            for f in dataclasses.fields(self)})
