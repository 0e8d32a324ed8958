"""MI355X (gfx950) VExpand / multi-way intersection, Python side.

Thin numpy-facing wrappers over the C ABI of ``libvertexsurge.so``
(include/vertexsurge.h).  All computation happens in the library's HIP
kernels; this module only moves arrays and handles.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (DIR_BACKWARD, DIR_BOTH, DIR_FORWARD, LAYOUT_DVSM, LAYOUT_NATIVE,
                   LAYOUT_ROWMAJOR, PATH_ANY, PATH_FIND_SHORTEST, PATH_SHORTEST,
                   SCHED_HEURISTIC, SCHED_LEARNED, SCHED_PULL, SCHED_PUSH, VSError)

__all__ = [
    "HostCSR", "Graph", "GraphStore", "store_write", "VExpandResult", "vexpand", "find_shortest", "mintersect",
    "ColMat", "mintersect_colmat", "mintersect_varrow", "pack_colmain", "unpack_colmain",
    "sample_sources", "sample_paths3", "Scheduler", "vexpand_batch",
    "vexpand_batch_results", "autotune_batch", "BatchQuery", "device_count",
    "PATH_ANY", "PATH_SHORTEST", "PATH_FIND_SHORTEST", "DIR_FORWARD", "DIR_BACKWARD",
    "DIR_BOTH", "SCHED_HEURISTIC", "SCHED_PUSH", "SCHED_PULL", "SCHED_LEARNED",
    "LAYOUT_NATIVE", "LAYOUT_DVSM", "LAYOUT_ROWMAJOR", "VSError",
]

_PATHS = {"any": PATH_ANY, "shortest": PATH_SHORTEST}
_DIRS = {"forward": DIR_FORWARD, "backward": DIR_BACKWARD, "both": DIR_BOTH}
_SCHED = {"heuristic": SCHED_HEURISTIC, "push": SCHED_PUSH, "pull": SCHED_PULL,
          "learned": SCHED_LEARNED}


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    n = C.c_int(0)
    _lib.call("vs_device_count", C.byref(n))
    return n.value


class HostCSR:
    """Host CSR produced by the library's generators (offsets uint64[V+1], adj uint32[E])."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        V = C.c_uint32()
        E = C.c_uint64()
        off = _lib.u64p()
        adj = _lib.u32p()
        _lib.call("vs_csr_info", handle, C.byref(V), C.byref(E), C.byref(off), C.byref(adj))
        self.V = V.value
        self.E = E.value
        self.offsets = np.ctypeslib.as_array(off, shape=(self.V + 1,))
        self.adj = (np.ctypeslib.as_array(adj, shape=(self.E,)) if self.E
                    else np.zeros(0, dtype=np.uint32))

    @classmethod
    def rmat(cls, scale: int, edge_factor: int, a=0.57, b=0.19, c=0.19, seed=1) -> "HostCSR":
        h = C.c_void_p()
        _lib.call("vs_rmat_csr", scale, edge_factor, a, b, c, seed, C.byref(h))
        return cls(h)

    @classmethod
    def erdos_renyi(cls, num_vertices: int, num_pairs: int, seed=1) -> "HostCSR":
        h = C.c_void_p()
        _lib.call("vs_er_csr", num_vertices, num_pairs, seed, C.byref(h))
        return cls(h)

    @classmethod
    def from_edges(cls, num_vertices: int, src, dst, symmetrize=True,
                   drop_self_loops=True) -> "HostCSR":
        s, d = _u32(src), _u32(dst)
        h = C.c_void_p()
        _lib.call("vs_csr_from_edges", num_vertices, _ptr(s, C.c_uint32), _ptr(d, C.c_uint32),
                  len(s), int(symmetrize), int(drop_self_loops), C.byref(h))
        return cls(h)

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)

    def close(self):
        if self._h:
            _lib.call("vs_csr_free", self._h)
            self._h = None
            self.offsets = self.adj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample_sources(num_vertices: int, S: int, seed: int) -> np.ndarray:
    out = np.zeros(S, dtype=np.uint32)
    _lib.call("vs_sample_sources", num_vertices, S, seed, _ptr(out, C.c_uint32))
    return out


def sample_paths3(offsets: np.ndarray, adj: np.ndarray, T: int, seed: int) -> np.ndarray:
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    a = _u32(adj)
    out = np.zeros((T, 3), dtype=np.uint32)
    _lib.call("vs_sample_paths3", _ptr(off, C.c_uint64), _ptr(a, C.c_uint32), len(off) - 1,
              T, seed, _ptr(out, C.c_uint32))
    return out


class Graph:
    """CSR resident in HBM (replaces the reference's global Graph* store)."""

    def __init__(self, offsets: np.ndarray, adj: np.ndarray, symmetric: bool = True):
        self._off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self._adj = _u32(adj)
        self.V = len(self._off) - 1
        self.E = int(self._off[-1])
        self._h = C.c_void_p()
        _lib.call("vs_graph_create", _ptr(self._off, C.c_uint64), _ptr(self._adj, C.c_uint32),
                  self.V, self.E, _lib.GRAPH_SYMMETRIC if symmetric else 0, C.byref(self._h))
        self._off = self._adj = None  # the library keeps its own copy

    @classmethod
    def from_csr(cls, csr: HostCSR, symmetric: bool = True) -> "Graph":
        return cls(csr.offsets, csr.adj, symmetric)

    @classmethod
    def _wrap(cls, handle: C.c_void_p) -> "Graph":
        g = cls.__new__(cls)
        g._h = handle
        V, E, dev = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _lib.call("vs_graph_info", handle, C.byref(V), C.byref(E), C.byref(dev))
        g.V, g.E = V.value, E.value
        g._off = g._adj = None
        return g

    @classmethod
    def load(cls, directory, edge: str | None = None) -> "Graph":
        """Device graph from a directory in the reference's storage format
        (schema.json + DiskArray files; Graph::load, graph_store.cpp:47-73)."""
        with GraphStore(directory, edge) as st:
            return st.graph()

    @property
    def handle(self):
        return self._h

    def set_tuning(self, tile_words=0, push_alpha=16, pull_hubs=0, sparse_pct=0, wpush_pml=0):
        t = _lib.Tuning(tile_words, push_alpha, sparse_pct, pull_hubs, wpush_pml)
        _lib.call("vs_graph_set_tuning", self._h, C.byref(t))

    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.call("vs_graph_get_tuning", self._h, C.byref(t))
        return {"tile_words": t.tile_words, "push_alpha": t.push_alpha,
                "pull_hubs": t.pull_hubs, "sparse_pct": t.sparse_pct,
                "wpush_pml": t.wpush_pml}

    def set_scheduler(self, sched: "Scheduler | None"):
        # the library keeps its own reference to an attached scheduler
        _lib.call("vs_graph_set_scheduler", self._h, sched.handle if sched else None)

    def trim(self):
        _lib.call("vs_graph_trim", self._h)

    def close(self):
        if self._h:
            _lib.call("vs_graph_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraphStore:
    """One edge relation of a graph directory in the reference's storage
    format, opened on the host (vs_store_open): schema.json + the relation's
    source-side CSR DiskArray files, with lists normalised to ascending sets."""

    def __init__(self, directory, edge: str | None = None):
        self._h = C.c_void_p()
        _lib.call("vs_store_open", str(directory).encode(),
                  edge.encode() if edge else None, C.byref(self._h))
        st = _lib.StoreInfo()
        _lib.call("vs_store_get_info", self._h, C.byref(st))
        self.info = {k: getattr(st, k) for k, _ in _lib.StoreInfo._fields_}
        self.V = st.num_vertices
        self.E = st.num_edges

    def csr(self):
        """(offsets uint64[V+1], adj uint32[E]) copies of the normalised CSR."""
        off, adj = _lib.u64p(), _lib.u32p()
        _lib.call("vs_store_csr", self._h, C.byref(off), C.byref(adj))
        o = np.ctypeslib.as_array(off, shape=(self.V + 1,)).copy()
        a = (np.ctypeslib.as_array(adj, shape=(self.E,)).copy() if self.E
             else np.zeros(0, dtype=np.uint32))
        return o, a

    def graph(self) -> "Graph":
        h = C.c_void_p()
        _lib.call("vs_store_graph", self._h, C.byref(h))
        return Graph._wrap(h)

    def close(self):
        if self._h:
            _lib.call("vs_store_close", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def store_write(directory, offsets, adj, vertex_type: str = "Vertex", edge_type: str = "edge",
                directed: bool = False, graph_name: str = "",
                flat_partitions: int = 0) -> None:
    """Writes one relation of vertex_type to itself as a store the reference's
    Graph::load reads (vs_store_write).  flat_partitions > 0 also writes the
    src_dst / hilbert flat edge files, whole and in that many destination
    partitions.  directory must exist."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    a = _u32(adj)
    V = len(off) - 1
    _lib.call("vs_store_write", str(directory).encode(), graph_name.encode(),
              vertex_type.encode(), edge_type.encode(), int(directed), _ptr(off, C.c_uint64),
              _ptr(a, C.c_uint32), V, int(off[-1]) if V >= 0 else 0, int(flat_partitions))


def stats_dict(st: _lib.VexpandStats) -> dict:
    k = st.k_actual_max
    return {
        "k_actual_max": k, "tile_words": st.tile_words, "num_tiles": st.num_tiles,
        "te": st.te, "bytes_alg": st.bytes_alg, "visit_count": st.visit_count,
        "ms_total": st.ms_total, "ms_hops": st.ms_hops,
        "ms_hop": list(st.ms_hop[:k]), "mode_hop": list(st.mode_hop[:k]),
        "te_hop": list(st.te_hop[:k]), "bytes_hop": list(st.bytes_hop[:k]),
        "frontier_bits": list(st.frontier_bits[:k]), "new_bits": list(st.new_bits[:k]),
        "ms_pull": st.ms_pull, "ms_push": st.ms_push, "n_pull": st.n_pull,
        "n_push": st.n_push, "bytes_pull": st.bytes_pull, "bytes_push": st.bytes_push,
        "frontier_act": list(st.frontier_act[:k]), "frontier_deg": list(st.frontier_deg[:k]),
        "vis_frac": list(st.vis_frac[:k]),
        "src_degsum": st.src_degsum, "src_vol2": st.src_vol2,
        "tile_words_learned": bool(st.tile_words_learned),
        "wpush_words": list(st.wpush_words[:k]),
    }


class VExpandResult:
    """VarRowMatrixTable (one merged matrix) or MultipleMatrixTable (one per hop)."""

    def __init__(self, handle: C.c_void_p, sources: np.ndarray):
        self._h = handle
        self.row_map = sources  # the reference sets row_map = start vertices
        rows, cols, n, k = (C.c_uint32() for _ in range(4))
        _lib.call("vs_result_info", handle, C.byref(rows), C.byref(cols), C.byref(n), C.byref(k))
        self.rows, self.cols = rows.value, cols.value
        self.num_matrices, self.k_actual_max = n.value, k.value

    def distance(self, i: int = 0):
        d = C.c_uint32()
        _lib.call("vs_result_distance", self._h, i, C.byref(d))
        return None if d.value == 0xFFFFFFFF else d.value

    def tuple_count(self, i: int = 0) -> int:
        c = C.c_uint64()
        _lib.call("vs_result_count", self._h, i, C.byref(c))
        return c.value

    def row_counts(self, i: int = 0) -> np.ndarray:
        out = np.zeros(self.rows, dtype=np.uint64)
        if self.rows:
            _lib.call("vs_result_row_counts", self._h, i, _ptr(out, C.c_uint64))
        return out

    def row_digests(self, i: int = 0) -> np.ndarray:
        out = np.zeros(self.rows, dtype=np.uint64)
        if self.rows:
            _lib.call("vs_result_row_digests", self._h, i, _ptr(out, C.c_uint64))
        return out

    def export(self, i: int = 0, layout: int = LAYOUT_ROWMAJOR) -> np.ndarray:
        n = C.c_size_t()
        _lib.call("vs_result_export_size", self._h, i, layout, C.byref(n))
        out = np.zeros(n.value // 8, dtype=np.uint64)
        if n.value:
            _lib.call("vs_result_export", self._h, i, layout, out.ctypes.data_as(C.c_void_p),
                      n.value)
        if layout == LAYOUT_ROWMAJOR:
            return out.reshape(self.rows, (self.cols + 63) // 64)
        return out

    def get(self, row: int, col: int, i: int = 0) -> bool:
        b = C.c_int()
        _lib.call("vs_result_get", self._h, i, row, col, C.byref(b))
        return bool(b.value)

    def to_flat(self, with_distance: bool = False):
        """(start, target[, distance]) arrays, the CSR strategy's flat table
        (merge_to_Flat, vexpand.cpp:380-446), ordered by row then target."""
        n = C.c_uint64()
        _lib.call("vs_result_flat", self._h, None, None, None, 0, C.byref(n))
        s = np.zeros(max(n.value, 1), dtype=np.uint32)
        t = np.zeros(max(n.value, 1), dtype=np.uint32)
        d = np.zeros(max(n.value, 1), dtype=np.int32) if with_distance else None
        if n.value:
            _lib.call("vs_result_flat", self._h, _ptr(s, C.c_uint32), _ptr(t, C.c_uint32),
                      _ptr(d, C.c_int32) if with_distance else None, n.value, C.byref(n))
        s, t = s[:n.value], t[:n.value]
        return (s, t, d[:n.value]) if with_distance else (s, t)

    def col_map(self):
        """Vertex ids of the columns of a column select; None for VExpand output."""
        if not getattr(self, "_selected", False):
            return None
        out = np.zeros(self.cols, dtype=np.uint32)
        if self.cols:
            _lib.call("vs_result_col_map", self._h, _ptr(out, C.c_uint32))
        return out

    def to_colmain(self, i: int = 0, transpose: bool = False) -> "ColMat":
        """TableConvert VarRow -> ColMain (maps travel; transpose swaps them)."""
        h = C.c_void_p()
        _lib.call("vs_table_convert", self._h, i, int(transpose), C.byref(h))
        return ColMat(h)

    def select_cols(self, col_ids, i: int = 0) -> "VExpandResult":
        """MatrixTableSelect: a new result with columns col_ids (vertex ids)."""
        ids = _u32(col_ids)
        h = C.c_void_p()
        _lib.call("vs_result_select_cols", self._h, i, _ptr(ids, C.c_uint32), len(ids),
                  C.byref(h))
        r = VExpandResult(h, self.row_map)
        r._selected = True
        return r

    def col_counts(self, i: int = 0, order: str | None = None):
        """MatrixColumnCount: (ids, counts) per column; order None/'asc'/'desc'."""
        return _col_counts("vs_result_col_counts", (self._h, i), self.cols, order)

    def row_or(self, i: int = 0) -> np.ndarray:
        """MatrixRowOrSum: ids of the columns with any set bit."""
        return _row_or("vs_result_row_or", (self._h, i))

    def close(self):
        if self._h:
            _lib.call("vs_result_free", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vexpand(graph: Graph, sources, k_min: int, k_max: int, path: str = "shortest",
            direction: str = "forward", with_distance: bool = False,
            sched: str = "heuristic", tile_words: int = 0, timing: bool = False,
            keep_order: bool = False):
    """Multi-source VExpand; returns (VExpandResult, stats dict).  Rows are
    expanded packed by 2-hop volume and returned in the caller's order unless
    keep_order is set (then they are also expanded in that order)."""
    src = _u32(sources)
    flags = (_lib.VX_TIMING if timing else 0) | (_lib.VX_KEEP_ORDER if keep_order else 0)
    p = _lib.VexpandParams(k_min, k_max, _PATHS[path], _DIRS[direction], int(with_distance),
                           _SCHED[sched], tile_words, flags)
    st = _lib.VexpandStats()
    h = C.c_void_p()
    _lib.call("vs_vexpand", graph.handle, _ptr(src, C.c_uint32), len(src), C.byref(p),
              C.byref(h), C.byref(st))
    return VExpandResult(h, src), stats_dict(st)


def find_shortest(graph: Graph, sources, targets, k_max: int = 50,
                  direction: str = "forward") -> np.ndarray:
    src, tgt = _u32(sources), _u32(targets)
    if len(src) != len(tgt):
        raise ValueError("sources and targets differ in length")
    dist = np.zeros(len(src), dtype=np.int32)
    st = _lib.VexpandStats()
    _lib.call("vs_find_shortest", graph.handle, _ptr(src, C.c_uint32), _ptr(tgt, C.c_uint32),
              len(src), _DIRS[direction], k_max, _ptr(dist, C.c_int32), C.byref(st))
    return dist


def mintersect(deps: list, tuples, with_lists: bool = True):
    """Multi-way intersection; returns (counts, offsets, lists, stats)."""
    tup = np.ascontiguousarray(tuples, dtype=np.uint32)
    if tup.ndim == 1:
        tup = tup.reshape(-1, len(deps))
    T, arity = tup.shape
    if arity != len(deps):
        raise ValueError("tuple arity differs from the number of dependencies")
    arr = (C.c_void_p * arity)(*[d.handle for d in deps])
    counts = np.zeros(T, dtype=np.uint32)
    offsets = np.zeros(T + 1, dtype=np.uint64)
    total = C.c_uint64()
    st = _lib.MintersectStats()
    lists = None
    if not with_lists:
        _lib.call("vs_mintersect", arr, arity, _ptr(tup, C.c_uint32), T,
                  _ptr(counts, C.c_uint32), _ptr(offsets, C.c_uint64), None, 0,
                  C.byref(total), C.byref(st))
    else:
        # one intersection pass when the lists fit a first guess of one entry per
        # tuple; a capacity error still returns counts, offsets and the total,
        # and the call runs again with room for exactly that
        cap = max(T, 1)
        while True:
            lists = np.zeros(cap, dtype=np.uint32)
            try:
                _lib.call("vs_mintersect", arr, arity, _ptr(tup, C.c_uint32), T,
                          _ptr(counts, C.c_uint32), _ptr(offsets, C.c_uint64),
                          _ptr(lists, C.c_uint32), cap, C.byref(total), C.byref(st))
                break
            except VSError as e:
                if e.code != _lib.VS_ERR_CAPACITY or total.value <= cap:
                    raise
                cap = total.value
        lists = lists[:total.value]
    return counts, offsets, lists, {"ms": st.ms, "total": st.total, "bytes_alg": st.bytes_alg}


_ORDERS = {None: _lib.ORDER_NONE, "asc": _lib.ORDER_ASC, "desc": _lib.ORDER_DESC}


def _col_counts(fn, head, n, order):
    ids = np.zeros(n, dtype=np.uint32)
    counts = np.zeros(n, dtype=np.int32)
    _lib.call(fn, *head, _ORDERS[order], _ptr(ids, C.c_uint32), _ptr(counts, C.c_int32))
    return ids, counts


def _row_or(fn, head):
    n = C.c_uint32()
    _lib.call(fn, *head, None, 0, C.byref(n))
    ids = np.zeros(max(n.value, 1), dtype=np.uint32)
    _lib.call(fn, *head, _ptr(ids, C.c_uint32), len(ids), C.byref(n))
    return ids[:n.value]


class ColMat:
    """Device DenseColMainMatrix + row/column maps (ColMainMatrixTable)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        rows, cols, ncm = C.c_uint32(), C.c_uint32(), C.c_uint32()
        pitch = C.c_uint64()
        hr = C.c_int()
        _lib.call("vs_colmat_info", handle, C.byref(rows), C.byref(cols), C.byref(pitch),
                  C.byref(hr), C.byref(ncm))
        self.rows, self.cols, self.pitch = rows.value, cols.value, pitch.value
        self.has_row_map, self.n_col_map = bool(hr.value), ncm.value

    @classmethod
    def create(cls, graph: Graph, rows: int, cols: int, words=None, row_map=None,
               col_map=None) -> "ColMat":
        """words: cols x pitch uint64 (pitch = align8(ceil(rows/64))) or None."""
        w = None
        if words is not None:
            w = np.ascontiguousarray(words, dtype=np.uint64)
            pitch = (((rows + 63) // 64) + 7) // 8 * 8
            if w.size != cols * pitch:
                raise ValueError(f"words: {w.size} words, expected cols x pitch = {cols * pitch}")
        h = C.c_void_p()
        _lib.call("vs_colmat_create", graph.handle, rows, cols,
                  _ptr(w, C.c_uint64) if w is not None else None, C.byref(h))
        m = cls(h)
        if row_map is not None or col_map is not None:
            m.set_maps(row_map, col_map)
        return m

    @classmethod
    def from_bool(cls, graph: Graph, M, row_map=None, col_map=None) -> "ColMat":
        """From a rows x cols boolean array."""
        M = np.asarray(M, dtype=bool)
        return cls.create(graph, M.shape[0], M.shape[1], pack_colmain(M), row_map, col_map)

    @property
    def handle(self):
        return self._h

    def set_maps(self, row_map=None, col_map=None):
        rm = _u32(row_map) if row_map is not None else None
        cm = _u32(col_map) if col_map is not None else None
        if rm is not None and len(rm) != self.rows:
            raise ValueError("row map length differs from the row count")
        if cm is not None and len(cm) > self.cols:
            raise ValueError("column map longer than the column count")
        _lib.call("vs_colmat_set_maps", self._h,
                  _ptr(rm, C.c_uint32) if rm is not None else None,
                  _ptr(cm, C.c_uint32) if cm is not None else None,
                  len(cm) if cm is not None else 0)
        self.has_row_map = rm is not None
        self.n_col_map = len(cm) if cm is not None else 0

    @property
    def has_col_map(self) -> bool:
        return self.n_col_map > 0

    def maps(self):
        rm = np.zeros(self.rows, dtype=np.uint32) if self.has_row_map else None
        cm = np.zeros(self.n_col_map, dtype=np.uint32) if self.has_col_map else None
        _lib.call("vs_colmat_get_maps", self._h,
                  _ptr(rm, C.c_uint32) if rm is not None else None,
                  _ptr(cm, C.c_uint32) if cm is not None else None)
        return rm, cm

    def export(self) -> np.ndarray:
        """cols x pitch uint64 words, the reference DenseColMainMatrix layout."""
        out = np.zeros((self.cols, self.pitch), dtype=np.uint64)
        if out.size:
            _lib.call("vs_colmat_export", self._h, _ptr(out, C.c_uint64), out.nbytes)
        return out

    def to_bool(self) -> np.ndarray:
        return unpack_colmain(self.export(), self.rows)

    def get(self, row: int, col: int) -> bool:
        b = C.c_int()
        _lib.call("vs_colmat_get", self._h, row, col, C.byref(b))
        return bool(b.value)

    def col_counts(self, order: str | None = None):
        return _col_counts("vs_colmat_col_counts", (self._h,), self.cols, order)

    def row_or(self) -> np.ndarray:
        return _row_or("vs_colmat_row_or", (self._h,))

    def close(self):
        if self._h:
            _lib.call("vs_colmat_free", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_colmain(M) -> np.ndarray:
    """rows x cols bool -> cols x align8(ceil(rows/64)) uint64 words."""
    M = np.asarray(M, dtype=bool)
    R, Cn = M.shape
    nw = (R + 63) // 64
    pitch = (nw + 7) // 8 * 8
    bits = np.zeros((Cn, pitch * 64), dtype=np.uint8)
    bits[:, :R] = M.T
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(Cn, pitch)


def unpack_colmain(words, rows: int) -> np.ndarray:
    w = np.ascontiguousarray(words, dtype=np.uint64)
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :rows].T.astype(bool)


def mintersect_colmat(deps: list, tuples, with_lists: bool = True):
    """Mintersect over ColMain dependencies; returns (counts, offsets, lists, stats)."""
    tup = np.ascontiguousarray(tuples, dtype=np.uint32)
    if tup.ndim == 1:
        tup = tup.reshape(-1, len(deps))
    T, nd = tup.shape
    if nd != len(deps):
        raise ValueError("tuple width differs from the number of dependencies")
    arr = (C.c_void_p * nd)(*[d.handle for d in deps])
    counts = np.zeros(T, dtype=np.uint32)
    offsets = np.zeros(T + 1, dtype=np.uint64)
    total = C.c_uint64()
    st = _lib.MintersectStats()
    _lib.call("vs_mintersect_colmat", arr, nd, _ptr(tup, C.c_uint32), T,
              _ptr(counts, C.c_uint32), _ptr(offsets, C.c_uint64), None, 0, C.byref(total),
              C.byref(st))
    lists = None
    if with_lists:
        lists = np.zeros(max(total.value, 1), dtype=np.uint32)
        _lib.call("vs_mintersect_colmat", arr, nd, _ptr(tup, C.c_uint32), T,
                  _ptr(counts, C.c_uint32), _ptr(offsets, C.c_uint64),
                  _ptr(lists, C.c_uint32), len(lists), C.byref(total), C.byref(st))
        lists = lists[:total.value]
    return counts, offsets, lists, {"ms": st.ms, "total": st.total, "bytes_alg": st.bytes_alg}


def mintersect_varrow(base: "VExpandResult", deps: list, var_idx, i: int = 0,
                      with_lists: bool = True):
    """Mintersect over a VExpand result as the base table (the reference's
    VarRow-base branch): dependency d intersects the column bound to the start
    (var_idx[d] == 0) or target (1) vertex of every set bit.  deps: all ColMat
    or all Graph.  Returns (start, target, new) arrays ordered by base tuple
    and new vertex, or the total when with_lists is False."""
    nd = len(deps)
    vi = _u32(var_idx)
    if len(vi) != nd:
        raise ValueError("one base variable per dependency")
    arr = (C.c_void_p * nd)(*[d.handle for d in deps])
    mats = all(isinstance(d, ColMat) for d in deps)
    if not mats and not all(isinstance(d, Graph) for d in deps):
        raise ValueError("dependencies must be all ColMat or all Graph")
    n = C.c_uint64()
    st = _lib.MintersectStats()
    args = (arr, None) if mats else (None, arr)
    _lib.call("vs_mintersect_varrow", base._h, i, *args, nd, _ptr(vi, C.c_uint32), None, None,
              None, 0, C.byref(n), C.byref(st))
    if not with_lists:
        return n.value
    out = [np.zeros(max(1, n.value), dtype=np.uint32) for _ in range(3)]
    if n.value:
        _lib.call("vs_mintersect_varrow", base._h, i, *args, nd, _ptr(vi, C.c_uint32),
                  *[_ptr(o, C.c_uint32) for o in out], n.value, C.byref(n), C.byref(st))
    return tuple(o[:n.value] for o in out)


class Scheduler:
    """The learned push/pull scheduler (FP32 MLP with HIP forward/backward)."""

    def __init__(self, seed: int = 1):
        self._h = C.c_void_p()
        _lib.call("vs_sched_create", seed, C.byref(self._h))
        n = C.c_size_t()
        _lib.call("vs_sched_num_params", C.byref(n))
        self.num_params = n.value
        self.domain = None  # (rows, radius) once set

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, dtype=np.float32)
        _lib.call("vs_sched_get_params", self._h, _ptr(p, C.c_float), len(p))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.call("vs_sched_set_params", self._h, _ptr(p, C.c_float), len(p))

    def loss_grad(self, X, Y, M):
        X, Y, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, M))
        loss = C.c_float()
        grad = np.zeros(self.num_params, dtype=np.float32)
        _lib.call("vs_sched_loss_grad", self._h, _ptr(X, C.c_float), _ptr(Y, C.c_float),
                  _ptr(M, C.c_float), len(X), C.byref(loss), _ptr(grad, C.c_float))
        return loss.value, grad

    def train(self, X, Y, M, epochs: int, lr: float = 1e-3) -> np.ndarray:
        X, Y, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, M))
        curve = np.zeros(epochs, dtype=np.float32)
        _lib.call("vs_sched_train", self._h, _ptr(X, C.c_float), _ptr(Y, C.c_float),
                  _ptr(M, C.c_float), len(X), epochs, lr, _ptr(curve, C.c_float))
        return curve

    def set_domain(self, rows, radius: float):
        """Feature rows (normally the training samples) and a radius: the MLP
        overrules the heuristic only for decisions within radius of a row."""
        rows = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, _lib.MLP_IN)
        _lib.call("vs_sched_set_domain", self._h, _ptr(rows, C.c_float), len(rows),
                  float(radius))
        self.domain = (rows.copy(), float(radius))

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), _lib.MLP_OUT), dtype=np.float32)
        _lib.call("vs_sched_predict", self._h, _ptr(X, C.c_float), len(X), _ptr(Y, C.c_float))
        return Y

    def close(self):
        if self._h:
            _lib.call("vs_sched_free", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BatchQuery:
    sources: np.ndarray
    k_min: int
    k_max: int
    path: str = "shortest"


def _batch_args(queries, tile_words, sched, use_graph, replays, mem_budget, direction, timing,
                keep_order):
    keep = [_u32(q.sources) for q in queries]
    arr = (_lib.Query * len(queries))()
    for i, q in enumerate(queries):
        arr[i].sources = _ptr(keep[i], C.c_uint32)
        arr[i].S = len(keep[i])
        arr[i].k_min = q.k_min
        arr[i].k_max = q.k_max
        arr[i].path_type = _PATHS[q.path]
    p = _lib.BatchParams(tile_words, _SCHED[sched], int(use_graph), replays, mem_budget,
                         _DIRS[direction], (_lib.VX_TIMING if timing else 0) |
                         (_lib.BATCH_KEEP_ORDER if keep_order else 0))
    return keep, arr, p


def _batch_stats(st) -> dict:
    return {"te": st.te, "ms_total": st.ms_total, "ms_capture": st.ms_capture,
            "num_tiles": st.num_tiles, "num_waves": st.num_waves, "launches": st.launches,
            "push_decisions": st.push_decisions, "pull_decisions": st.pull_decisions,
            "ms_expand": st.ms_expand, "n_expand": st.n_expand, "bytes_alg": st.bytes_alg,
            "ms_wall": st.ms_wall, "ms_expand_sparse": st.ms_expand_sparse,
            "n_expand_sparse": st.n_expand_sparse, "bytes_sparse": st.bytes_sparse,
            "wpush_words": st.wpush_words}


def autotune_batch(graph: Graph, queries: list, tile_words: int = 0, sched: str = "heuristic",
                   direction: str = "forward"):
    """Tunes the batched engine on these queries (vs_autotune_batch); the
    best tuning is stored in the graph.  Returns (tuning dict, best ms, start ms)."""
    keep, arr, p = _batch_args(queries, tile_words, sched, True, 1, 0, direction, False, False)
    t = _lib.Tuning()
    best, start = C.c_double(), C.c_double()
    _lib.call("vs_autotune_batch", graph.handle, arr, len(queries), C.byref(p), C.byref(t),
              C.byref(best), C.byref(start))
    return ({"tile_words": t.tile_words, "push_alpha": t.push_alpha,
             "sparse_pct": t.sparse_pct, "pull_hubs": t.pull_hubs,
             "wpush_pml": t.wpush_pml}, best.value, start.value)


def vexpand_batch_results(graph: Graph, queries: list, tile_words: int = 0,
                          sched: str = "heuristic", use_graph: bool = True, replays: int = 1,
                          mem_budget: int = 0, direction: str = "forward",
                          timing: bool = False, keep_order: bool = False):
    """Runs many VExpand queries packed into shared tiles and returns
    ([VExpandResult per query], stats): each result is what vexpand() returns
    for that query (rows = its sources in its order)."""
    keep, arr, p = _batch_args(queries, tile_words, sched, use_graph, replays, mem_budget,
                               direction, timing, keep_order)
    hs = (C.c_void_p * max(1, len(queries)))()
    st = _lib.BatchStats()
    _lib.call("vs_vexpand_batch_results", graph.handle, arr, len(queries), C.byref(p), hs,
              C.byref(st))
    return [VExpandResult(C.c_void_p(hs[i]), keep[i]) for i in range(len(queries))], \
        _batch_stats(st)


def vexpand_batch(graph: Graph, queries: list, tile_words: int = 0, sched: str = "heuristic",
                  use_graph: bool = True, replays: int = 1, mem_budget: int = 0,
                  direction: str = "forward", want_digests: bool = True,
                  timing: bool = False, keep_order: bool = False):
    """Runs many VExpand queries packed into shared tiles.  Returns
    (row_counts, row_digests, stats) concatenated over queries, in query order.
    Rows are packed by source degree unless keep_order is set."""
    keep = [_u32(q.sources) for q in queries]
    arr = (_lib.Query * len(queries))()
    for i, q in enumerate(queries):
        arr[i].sources = _ptr(keep[i], C.c_uint32)
        arr[i].S = len(keep[i])
        arr[i].k_min = q.k_min
        arr[i].k_max = q.k_max
        arr[i].path_type = _PATHS[q.path]
    total = sum(len(k) for k in keep)
    counts = np.zeros(total, dtype=np.uint64)
    digests = np.zeros(total, dtype=np.uint64) if want_digests else None
    p = _lib.BatchParams(tile_words, _SCHED[sched], int(use_graph), replays, mem_budget,
                         _DIRS[direction], (_lib.VX_TIMING if timing else 0) |
                         (_lib.BATCH_KEEP_ORDER if keep_order else 0))
    st = _lib.BatchStats()
    _lib.call("vs_vexpand_batch", graph.handle, arr, len(queries), C.byref(p),
              _ptr(counts, C.c_uint64),
              _ptr(digests, C.c_uint64) if digests is not None else None, C.byref(st))
    return counts, digests, {
        "te": st.te, "ms_total": st.ms_total, "ms_capture": st.ms_capture,
        "num_tiles": st.num_tiles, "num_waves": st.num_waves, "launches": st.launches,
        "push_decisions": st.push_decisions, "pull_decisions": st.pull_decisions,
        "ms_expand": st.ms_expand, "n_expand": st.n_expand, "bytes_alg": st.bytes_alg,
        "ms_wall": st.ms_wall, "ms_expand_sparse": st.ms_expand_sparse,
        "n_expand_sparse": st.n_expand_sparse, "bytes_sparse": st.bytes_sparse,
            "wpush_words": st.wpush_words}
