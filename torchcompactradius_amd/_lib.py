"""ctypes binding of libtcr_hip.so (the C ABI declared in include/tcr.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded, every search raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtcr_hip.so")

MODES = {"symmetric": 0, "gather": 1, "scatter": 2, "superSymmetric": 3}
TCR_OK, TCR_ERR_ARG, TCR_ERR_HIP, TCR_ERR_DOMAIN, TCR_ERR_SPACE, TCR_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5


class TcrCloud(ctypes.Structure):
    """Mirror of ``tcr_cloud`` in include/tcr.h."""
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("periodic_mask", ctypes.c_uint32),
        ("search_range", ctypes.c_int32),
        ("domain_min", ctypes.c_float * 3),
        ("domain_max", ctypes.c_float * 3),
        ("hash_map_length", ctypes.c_uint32),
        ("self_search", ctypes.c_int32),
        ("q_off", ctypes.c_int64),
        ("nq", ctypes.c_int64),
        ("r_off", ctypes.c_int64),
        ("nr", ctypes.c_int64),
        ("q_foff", ctypes.c_int64),
        ("r_foff", ctypes.c_int64),
    ]


class TcrTuning(ctypes.Structure):
    """Mirror of ``tcr_tuning`` in include/tcr.h."""
    _fields_ = [("cell_workgroups", ctypes.c_int32), ("stencil_workgroups", ctypes.c_int32),
                ("cell_threads", ctypes.c_int32), ("cells_per_wave", ctypes.c_int32)]


class TcrMlmrefGrid(ctypes.Structure):
    """Mirror of ``tcr_mlmref_grid`` in include/tcr.h."""
    _fields_ = [("dim", ctypes.c_int32), ("levels", ctypes.c_int32), ("hcell", ctypes.c_float),
                ("hfine", ctypes.c_float), ("hmin", ctypes.c_float), ("hmax", ctypes.c_float),
                ("res", (ctypes.c_int32 * 3) * 17)]


class TcrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libtcr_hip error {code}: {msg}")
        self.code = code


class Unsupported(TcrError):
    """Inputs the multi-level search does not take (the single-level search
    gives the same output for them)."""


_lib: Optional[ctypes.CDLL] = None

# Every symbol include/tcr.h declares (checked by the ABI tests).
EXPORTED = (
    "tcr_search_workspace_size", "tcr_search_count", "tcr_search_fill", "tcr_search_launch",
    "tcr_search_launch_csr",
    "tcr_search_finish", "tcr_search_export", "tcr_timing_enable", "tcr_timing_collect",
    "tcr_graph_capture", "tcr_graph_launch", "tcr_graph_destroy", "tcr_set_tuning", "tcr_get_tuning",
    "tcr_mlm_workspace_size", "tcr_mlm_count", "tcr_mlm_fill", "tcr_csr_arrays",
    "tcr_compute_hash_indices", "tcr_count_neighbors", "tcr_build_neighbor_list",
    "tcr_naive_workspace_size", "tcr_naive_count", "tcr_naive_fill", "tcr_small_count", "tcr_small_fill",
    "tcr_last_error", "tcr_version", "tcr_search_noalloc",
    "tcr_mlmref_plan", "tcr_mlmref_count", "tcr_mlmref_fill", "tcr_mlmref_transpose_workspace_size",
    "tcr_mlmref_transpose",
)


def load() -> ctypes.CDLL:
    """Load libtcr_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python torchcompactradius_amd/_build.py`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, u8p = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p
    cl = ctypes.POINTER(TcrCloud)
    L.tcr_last_error.restype = ctypes.c_char_p
    L.tcr_last_error.argtypes = []
    L.tcr_version.restype = ctypes.c_char_p
    L.tcr_version.argtypes = []
    L.tcr_search_workspace_size.restype = ctypes.c_size_t
    L.tcr_search_workspace_size.argtypes = [cl, i32]
    L.tcr_search_count.restype = ctypes.c_int
    L.tcr_search_count.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp,
                                   ctypes.POINTER(ctypes.c_int64), vp]
    L.tcr_search_fill.restype = ctypes.c_int
    L.tcr_search_fill.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, i32, vp]
    L.tcr_search_launch.restype = ctypes.c_int
    L.tcr_search_launch.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, i32, i64, vp]
    L.tcr_search_launch_csr.restype = ctypes.c_int
    L.tcr_search_launch_csr.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp, i32, i64, vp, vp, vp]
    L.tcr_graph_capture.restype = ctypes.c_int
    L.tcr_graph_capture.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, i32, i64,
                                    ctypes.POINTER(ctypes.c_void_p), vp]
    L.tcr_graph_launch.restype = ctypes.c_int
    L.tcr_graph_launch.argtypes = [vp, vp]
    L.tcr_graph_destroy.restype = ctypes.c_int
    L.tcr_graph_destroy.argtypes = [vp]
    L.tcr_set_tuning.restype = ctypes.c_int
    L.tcr_set_tuning.argtypes = [ctypes.POINTER(TcrTuning)]
    L.tcr_get_tuning.restype = ctypes.c_int
    L.tcr_get_tuning.argtypes = [ctypes.POINTER(TcrTuning)]
    L.tcr_mlm_workspace_size.restype = ctypes.c_size_t
    L.tcr_mlm_workspace_size.argtypes = [cl]
    L.tcr_mlm_count.restype = ctypes.c_int
    L.tcr_mlm_count.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int32), vp, i64, ctypes.POINTER(ctypes.c_int32), vp]
    L.tcr_mlm_fill.restype = ctypes.c_int
    L.tcr_mlm_fill.argtypes = [cl, i32, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, vp, vp, i32, vp, vp, i64, vp]
    L.tcr_csr_arrays.restype = ctypes.c_int
    L.tcr_csr_arrays.argtypes = [vp, i64, vp, vp, vp]
    L.tcr_timing_enable.restype = ctypes.c_int
    L.tcr_timing_enable.argtypes = [i32]
    L.tcr_timing_collect.restype = ctypes.c_int
    L.tcr_timing_collect.argtypes = [vp, vp, vp, i32, ctypes.POINTER(ctypes.c_int32)]
    L.tcr_search_finish.restype = ctypes.c_int
    L.tcr_search_finish.argtypes = [cl, i32, vp, ctypes.c_size_t, vp, ctypes.POINTER(ctypes.c_int64), vp]
    L.tcr_search_export.restype = ctypes.c_int
    L.tcr_search_export.argtypes = [cl, i32, i32, vp, ctypes.c_size_t, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.tcr_compute_hash_indices.restype = ctypes.c_int
    L.tcr_compute_hash_indices.argtypes = [vp, i64, i32, i64, vp, vp]
    L.tcr_count_neighbors.restype = ctypes.c_int
    L.tcr_count_neighbors.argtypes = [vp, vp, i64, i32, vp, vp, i64, vp, i64, vp, vp, i64, vp,
                                      ctypes.c_double, vp, vp, u8p, i32, i32, vp, vp]
    L.tcr_build_neighbor_list.restype = ctypes.c_int
    L.tcr_build_neighbor_list.argtypes = [vp, vp, i64, vp, vp, i64, i32, vp, vp, i64, vp, i64, vp, vp, i64, vp,
                                          ctypes.c_double, vp, vp, u8p, i32, i32, vp, vp, vp]
    L.tcr_search_noalloc.restype = ctypes.c_int
    L.tcr_search_noalloc.argtypes = [cl, i32, vp, vp, vp, vp, vp, vp, vp, i32, i64, ctypes.POINTER(ctypes.c_int64),
                                     vp]
    gp = ctypes.POINTER(TcrMlmrefGrid)
    L.tcr_mlmref_plan.restype = ctypes.c_int
    L.tcr_mlmref_plan.argtypes = [vp, i64, i64, i32, vp, vp, vp, gp, ctypes.POINTER(ctypes.c_size_t), vp]
    L.tcr_mlmref_count.restype = ctypes.c_int
    L.tcr_mlmref_count.argtypes = [gp, vp, vp, i64, vp, vp, i64, vp, vp, ctypes.c_uint32, i32, i64, vp,
                                   ctypes.c_size_t, vp, vp, ctypes.POINTER(ctypes.c_int64), vp]
    L.tcr_mlmref_fill.restype = ctypes.c_int
    L.tcr_mlmref_fill.argtypes = [gp, vp, vp, i64, vp, vp, i64, vp, vp, ctypes.c_uint32, i32, i64, vp,
                                  ctypes.c_size_t, vp, i64, vp, vp, vp]
    L.tcr_mlmref_transpose_workspace_size.restype = ctypes.c_size_t
    L.tcr_mlmref_transpose_workspace_size.argtypes = [i64, i64]
    L.tcr_mlmref_transpose.restype = ctypes.c_int
    L.tcr_mlmref_transpose.argtypes = [vp, vp, i64, i64, vp, ctypes.c_size_t, vp, vp, vp, vp, vp]
    dp = ctypes.POINTER(ctypes.c_double)
    L.tcr_naive_workspace_size.restype = ctypes.c_size_t
    L.tcr_naive_workspace_size.argtypes = [i64, i64, i32]
    for name in ("tcr_naive_count", "tcr_small_count"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [vp, vp, i64, vp, vp, i64, i32, vp, vp, ctypes.c_uint32, i32, i32, dp, vp,
                                     ctypes.c_size_t, vp, vp]
    for name in ("tcr_naive_fill", "tcr_small_fill"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [vp, vp, i64, vp, vp, i64, i32, vp, vp, ctypes.c_uint32, i32, i32, dp, vp,
                                     ctypes.c_size_t, vp, vp, vp, i32, vp]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != TCR_OK:
        msg = load().tcr_last_error().decode(errors="replace")
        if rc == TCR_ERR_ARG:
            raise ValueError(msg)
        if rc == TCR_ERR_UNSUPPORTED:
            raise Unsupported(rc, msg)
        raise TcrError(rc, msg)


def ptr(t) -> Optional[int]:
    """Device pointer of a tensor (None stays None)."""
    if t is None:
        return None
    return t.data_ptr()
