"""ctypes binding of libvertexsurge.so (the C ABI in include/vertexsurge.h).

The library is loaded from this package directory only; there is no fallback:
if it has not been built, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvertexsurge.so"

VS_OK = 0
VS_ERR_INVALID = 1
VS_ERR_NOMEM = 2
VS_ERR_HIP = 3
VS_ERR_CAPACITY = 4
VS_ERR_UNSUPPORTED = 5
VS_ERR_NODEVICE = 6
VS_ERR_IO = 7
VS_ERR_INTERNAL = 8

PATH_ANY, PATH_SHORTEST, PATH_FIND_SHORTEST = 0, 1, 2
DIR_FORWARD, DIR_BACKWARD, DIR_BOTH = 0, 1, 2
SCHED_HEURISTIC, SCHED_PUSH, SCHED_PULL, SCHED_LEARNED = 0, 1, 2, 3
LAYOUT_NATIVE, LAYOUT_DVSM, LAYOUT_ROWMAJOR = 0, 1, 2
VX_TIMING = 0x1
VX_KEEP_ORDER = 0x2
BATCH_KEEP_ORDER = 0x2
ORDER_NONE, ORDER_ASC, ORDER_DESC = 0, 1, 2
GRAPH_SYMMETRIC = 0x1
MAX_HOPS = 64
MLP_IN, MLP_HIDDEN, MLP_OUT = 8, 64, 8
NOALLOC_MAX_TUPLES = 4718592  # VS_NOALLOC_MAX_TUPLES
SCHED_MAX_DOMAIN = 1024  # VS_SCHED_MAX_DOMAIN

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class VexpandParams(C.Structure):
    _fields_ = [("k_min", C.c_uint32), ("k_max", C.c_uint32), ("path_type", C.c_int32),
                ("direction", C.c_int32), ("with_distance", C.c_int32), ("sched", C.c_int32),
                ("tile_words", C.c_uint32), ("flags", C.c_uint32)]


class VexpandStats(C.Structure):
    _fields_ = [("k_actual_max", C.c_uint32), ("tile_words", C.c_uint32),
                ("num_tiles", C.c_uint32), ("reserved", C.c_uint32),
                ("te", C.c_uint64), ("bytes_alg", C.c_uint64), ("visit_count", C.c_uint64),
                ("ms_total", C.c_double), ("ms_hops", C.c_double),
                ("ms_hop", C.c_double * MAX_HOPS), ("mode_hop", C.c_int32 * MAX_HOPS),
                ("te_hop", C.c_uint64 * MAX_HOPS), ("bytes_hop", C.c_uint64 * MAX_HOPS),
                ("frontier_bits", C.c_uint64 * MAX_HOPS), ("new_bits", C.c_uint64 * MAX_HOPS),
                ("ms_pull", C.c_double), ("ms_push", C.c_double),
                ("n_pull", C.c_uint64), ("n_push", C.c_uint64),
                ("bytes_pull", C.c_uint64), ("bytes_push", C.c_uint64),
                ("frontier_act", C.c_uint64 * MAX_HOPS), ("frontier_deg", C.c_uint64 * MAX_HOPS),
                ("vis_frac", C.c_double * MAX_HOPS),
                ("src_degsum", C.c_uint64), ("src_vol2", C.c_uint64),
                ("tile_words_learned", C.c_uint32), ("reserved2", C.c_uint32),
                ("wpush_words", C.c_uint32 * MAX_HOPS)]


class Tuning(C.Structure):
    _fields_ = [("tile_words", C.c_uint32), ("push_alpha", C.c_uint32),
                ("sparse_pct", C.c_uint32), ("pull_hubs", C.c_uint32),
                ("wpush_pml", C.c_uint32)]


class Query(C.Structure):
    _fields_ = [("sources", u32p), ("S", C.c_uint32), ("k_min", C.c_uint32),
                ("k_max", C.c_uint32), ("path_type", C.c_int32)]


class BatchParams(C.Structure):
    _fields_ = [("tile_words", C.c_uint32), ("sched", C.c_int32), ("use_graph", C.c_uint32),
                ("replays", C.c_uint32), ("mem_budget", C.c_size_t), ("direction", C.c_int32),
                ("flags", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("te", C.c_uint64), ("ms_total", C.c_double), ("ms_capture", C.c_double),
                ("num_tiles", C.c_uint32), ("num_waves", C.c_uint32), ("launches", C.c_uint32),
                ("reserved", C.c_uint32), ("push_decisions", C.c_uint64),
                ("pull_decisions", C.c_uint64), ("ms_expand", C.c_double),
                ("n_expand", C.c_uint64), ("bytes_alg", C.c_uint64), ("ms_wall", C.c_double),
                ("ms_expand_sparse", C.c_double), ("n_expand_sparse", C.c_uint64),
                ("bytes_sparse", C.c_uint64), ("wpush_words", C.c_uint64)]


class StoreInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("stored_entries", C.c_uint64),
                ("num_edges", C.c_uint64), ("duplicates", C.c_uint64),
                ("unsorted_lists", C.c_uint64), ("self_loops", C.c_uint64),
                ("directed", C.c_int32), ("allow_multiple_edge", C.c_int32)]


class MintersectStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("total", C.c_uint64), ("bytes_alg", C.c_uint64)]


# name -> (restype, argtypes)
SIGNATURES = {
    "vs_last_error": (C.c_char_p, []),
    "vs_version": (C.c_char_p, []),
    "vs_rmat_csr": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                              C.c_uint64, C.POINTER(vp)]),
    "vs_er_csr": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(vp)]),
    "vs_csr_from_edges": (C.c_int, [C.c_uint32, u32p, u32p, C.c_uint64, C.c_int, C.c_int,
                                    C.POINTER(vp)]),
    "vs_csr_info": (C.c_int, [vp, u32p, u64p, C.POINTER(u64p), C.POINTER(u32p)]),
    "vs_csr_free": (C.c_int, [vp]),
    "vs_sample_sources": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, u32p]),
    "vs_sample_paths3": (C.c_int, [u64p, u32p, C.c_uint32, C.c_uint64, C.c_uint64, u32p]),
    "vs_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "vs_set_device": (C.c_int, [C.c_int]),
    "vs_device_synchronize": (C.c_int, []),
    "vs_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
    "vs_dev_free": (C.c_int, [vp]),
    "vs_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
    "vs_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
    "vs_device_copy": (C.c_int, [vp, vp, C.c_size_t]),
    "vs_graph_create": (C.c_int, [u64p, u32p, C.c_uint32, C.c_uint64, C.c_uint32,
                                  C.POINTER(vp)]),
    "vs_graph_destroy": (C.c_int, [vp]),
    "vs_graph_info": (C.c_int, [vp, u32p, u64p, u64p]),
    "vs_graph_trim": (C.c_int, [vp]),
    "vs_graph_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
    "vs_graph_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
    "vs_result_flat": (C.c_int, [vp, u32p, u32p, i32p, C.c_uint64, u64p]),
    "vs_store_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
    "vs_store_get_info": (C.c_int, [vp, C.POINTER(StoreInfo)]),
    "vs_store_csr": (C.c_int, [vp, C.POINTER(u64p), C.POINTER(u32p)]),
    "vs_store_graph": (C.c_int, [vp, C.POINTER(vp)]),
    "vs_store_close": (C.c_int, [vp]),
    "vs_store_write": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                 u64p, u32p, C.c_uint32, C.c_uint64, C.c_uint32]),
    "vs_vexpand": (C.c_int, [vp, u32p, C.c_uint32, C.POINTER(VexpandParams), C.POINTER(vp),
                             C.POINTER(VexpandStats)]),
    "vs_find_shortest": (C.c_int, [vp, u32p, u32p, C.c_uint32, C.c_int32, C.c_uint32, i32p,
                                   C.POINTER(VexpandStats)]),
    "vs_result_info": (C.c_int, [vp, u32p, u32p, u32p, u32p]),
    "vs_result_distance": (C.c_int, [vp, C.c_uint32, u32p]),
    "vs_result_count": (C.c_int, [vp, C.c_uint32, u64p]),
    "vs_result_row_counts": (C.c_int, [vp, C.c_uint32, u64p]),
    "vs_result_row_digests": (C.c_int, [vp, C.c_uint32, u64p]),
    "vs_result_export_size": (C.c_int, [vp, C.c_uint32, C.c_int32, C.POINTER(C.c_size_t)]),
    "vs_result_export": (C.c_int, [vp, C.c_uint32, C.c_int32, vp, C.c_size_t]),
    "vs_result_get": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]),
    "vs_result_free": (C.c_int, [vp]),
    "vs_vexpand_batch": (C.c_int, [vp, C.POINTER(Query), C.c_uint32, C.POINTER(BatchParams),
                                   u64p, u64p, C.POINTER(BatchStats)]),
    "vs_vexpand_batch_results": (C.c_int, [vp, C.POINTER(Query), C.c_uint32,
                                           C.POINTER(BatchParams), C.POINTER(vp),
                                           C.POINTER(BatchStats)]),
    "vs_mintersect": (C.c_int, [C.POINTER(vp), C.c_uint32, u32p, C.c_uint64, u32p, u64p, u32p,
                                C.c_uint64, u64p, C.POINTER(MintersectStats)]),
    "vs_mintersect_device": (C.c_int, [C.POINTER(vp), C.c_uint32, vp, C.c_uint64, vp, vp, vp,
                                       C.c_uint64, u64p, C.POINTER(MintersectStats)]),
    "vs_mintersect_device_noalloc": (C.c_int, [C.POINTER(vp), C.c_uint32, vp, C.c_uint64, vp,
                                               vp, vp, C.c_uint64, u64p]),
    "vs_colmat_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, u64p, C.POINTER(vp)]),
    "vs_colmat_set_maps": (C.c_int, [vp, u32p, u32p, C.c_uint32]),
    "vs_colmat_info": (C.c_int, [vp, u32p, u32p, u64p, C.POINTER(C.c_int), u32p]),
    "vs_colmat_get_maps": (C.c_int, [vp, u32p, u32p]),
    "vs_colmat_export": (C.c_int, [vp, u64p, C.c_size_t]),
    "vs_colmat_get": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]),
    "vs_colmat_free": (C.c_int, [vp]),
    "vs_table_convert": (C.c_int, [vp, C.c_uint32, C.c_int, C.POINTER(vp)]),
    "vs_result_select_cols": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint32, C.POINTER(vp)]),
    "vs_result_col_map": (C.c_int, [vp, u32p]),
    "vs_result_col_counts": (C.c_int, [vp, C.c_uint32, C.c_int32, u32p, i32p]),
    "vs_colmat_col_counts": (C.c_int, [vp, C.c_int32, u32p, i32p]),
    "vs_result_row_or": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint32, u32p]),
    "vs_colmat_row_or": (C.c_int, [vp, u32p, C.c_uint32, u32p]),
    "vs_mintersect_varrow": (C.c_int, [vp, C.c_uint32, C.POINTER(vp), C.POINTER(vp), C.c_uint32,
                                       u32p, u32p, u32p, u32p, C.c_uint64, u64p,
                                       C.POINTER(MintersectStats)]),
    "vs_mintersect_colmat": (C.c_int, [C.POINTER(vp), C.c_uint32, u32p, C.c_uint64, u32p, u64p,
                                       u32p, C.c_uint64, u64p, C.POINTER(MintersectStats)]),
    "vs_sched_create": (C.c_int, [C.c_uint64, C.POINTER(vp)]),
    "vs_sched_free": (C.c_int, [vp]),
    "vs_sched_num_params": (C.c_int, [C.POINTER(C.c_size_t)]),
    "vs_sched_get_params": (C.c_int, [vp, f32p, C.c_size_t]),
    "vs_sched_set_params": (C.c_int, [vp, f32p, C.c_size_t]),
    "vs_sched_loss_grad": (C.c_int, [vp, f32p, f32p, f32p, C.c_uint32, f32p, f32p]),
    "vs_sched_train": (C.c_int, [vp, f32p, f32p, f32p, C.c_uint32, C.c_uint32, C.c_float,
                                 f32p]),
    "vs_sched_predict": (C.c_int, [vp, f32p, C.c_uint32, f32p]),
    "vs_sched_set_domain": (C.c_int, [vp, f32p, C.c_uint32, C.c_float]),
    "vs_graph_set_scheduler": (C.c_int, [vp, vp]),
    "vs_autotune_batch": (C.c_int, [vp, C.POINTER(Query), C.c_uint32, C.POINTER(BatchParams),
                                    C.POINTER(Tuning), f64p, f64p]),
    "vs_autotune_vexpand": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, C.c_int32,
                                      C.POINTER(Tuning), f64p]),
}

_LIB: C.CDLL | None = None


class VSError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code
        self.fn = fn


def load() -> C.CDLL:
    """Loads the in-tree library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def call(name: str, *args) -> None:
    """Calls an entry point returning vs_status and raises VSError on failure."""
    lib = load()
    code = getattr(lib, name)(*args)
    if code != VS_OK:
        msg = lib.vs_last_error().decode(errors="replace")
        raise VSError(name, code, msg)


def declared_symbols() -> list[str]:
    return sorted(SIGNATURES)
