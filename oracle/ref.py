"""ctypes binding of oracle/_ref/libtcr_ref.so: the reference's own CPU ops
(torchCompactRadius 0.5.5 ``countNeighbors``, ``buildNeighborList``,
``computeHashIndices``; cpp/neighborhoodDynamic_cpu.cpp,
cpp/hashing_cpu.cpp), compiled from the reference's sources by
``oracle/ref.mk`` (test infrastructure only).

``search`` runs the reference's op sequence of ``searchNeighbors_cpp``
(py/compactHashing/cppSearch.py:60-98) on the structure the oracle builds:
query positions wrapped once, sorted reference positions wrapped twice
(cppSearch.py:221), int32 offsets from a cumsum of the counts, then
``j = sortIndex[j]``.  It pins the oracle's traversal, distance, predicate and
row order against the reference's own compiled code.

The library exists only where the reference's sources were present at build
time; ``available()`` says whether it was built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_OUT = os.path.join(_HERE, "_ref")
_LIB_PATH = os.path.join(_OUT, "libtcr_ref.so")
REF_SRC = "/root/reference/src/torchCompactRadius/cppSrc"
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> Optional[str]:
    """Compile the reference's ops with oracle/ref.mk when its sources are
    present; returns the library path, or None when they are absent."""
    if not os.path.isdir(REF_SRC):
        return _LIB_PATH if os.path.exists(_LIB_PATH) else None
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-j3", "-f", os.path.join(_HERE, "ref.mk")] + (["-B"] if force else []),
                       check=True, cwd=os.path.dirname(_HERE))
    return _LIB_PATH


def available() -> bool:
    return os.path.exists(_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (loads libtorch_cpu / libc10 the library links)
        L = ctypes.CDLL(_LIB_PATH)
        common = [_f32p, _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _f32p, _f32p, ctypes.c_int64,
                  _i32p, ctypes.c_int64, _i32p, _i32p, ctypes.c_int64, ctypes.c_double, _f32p, _f32p, _u8p,
                  ctypes.c_int32]
        L.ref_count_neighbors.restype = ctypes.c_int
        L.ref_count_neighbors.argtypes = common + [_i32p]
        L.ref_build_neighbor_list.restype = ctypes.c_int
        L.ref_build_neighbor_list.argtypes = [_i32p, _i32p, ctypes.c_int64] + common + [_i64p, _i64p]
        L.ref_compute_hash_indices.restype = ctypes.c_int
        L.ref_compute_hash_indices.argtypes = [_i32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, _i32p]
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_set_threads.restype = None
        L.ref_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return ctypes.cast(None, t) if a is None else a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().ref_last_error().decode())


MODES = {"symmetric": 0, "gather": 1, "scatter": 2, "superSymmetric": 3}


class Structure:
    """The reference-layout structure arguments of the ops."""

    def __init__(self, sorted_pos, sorted_sup, hash_table, cell_table, num_cells, hcell, dmin, dmax, periodic):
        self.sorted_pos = np.ascontiguousarray(sorted_pos, np.float32)
        self.sorted_sup = None if sorted_sup is None else np.ascontiguousarray(sorted_sup, np.float32)
        self.hash_table = np.ascontiguousarray(hash_table, np.int32)
        self.cell_table = np.ascontiguousarray(cell_table, np.int32)
        self.num_cells = np.ascontiguousarray(num_cells, np.int32)
        self.hcell = float(np.float32(hcell))
        self.dmin = np.ascontiguousarray(dmin, np.float32)
        self.dmax = np.ascontiguousarray(dmax, np.float32)
        self.periodic = np.ascontiguousarray(np.asarray(periodic, dtype=bool).astype(np.uint8))

    def args(self, qpos, qsup, mode, search_range):
        qpos = np.ascontiguousarray(qpos, np.float32)
        qsup = None if qsup is None else np.ascontiguousarray(qsup, np.float32)
        nq, dim = qpos.shape
        return (qpos, qsup), (_p(qpos, _f32p), _p(qsup, _f32p), nq, dim, search_range,
                              _p(self.sorted_pos, _f32p), _p(self.sorted_sup, _f32p), self.sorted_pos.shape[0],
                              _p(self.hash_table, _i32p), self.hash_table.shape[0], _p(self.num_cells, _i32p),
                              _p(self.cell_table, _i32p), self.cell_table.shape[0], self.hcell,
                              _p(self.dmin, _f32p), _p(self.dmax, _f32p), _p(self.periodic, _u8p), MODES[mode])


def count_neighbors(st: Structure, qpos, qsup, mode: str, search_range: int = 1) -> np.ndarray:
    keep, a = st.args(qpos, qsup, mode, search_range)
    out = np.zeros(keep[0].shape[0], np.int32)
    _check(lib().ref_count_neighbors(*a, _p(out, _i32p)))
    return out


def build_neighbor_list(st: Structure, counts, offsets, length: int, qpos, qsup, mode: str, search_range: int = 1):
    keep, a = st.args(qpos, qsup, mode, search_range)
    counts = np.ascontiguousarray(counts, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int32)
    i = np.zeros(length, np.int64)
    j = np.zeros(length, np.int64)
    _check(lib().ref_build_neighbor_list(_p(counts, _i32p), _p(offsets, _i32p), int(length), *a, _p(i, _i64p),
                                         _p(j, _i64p)))
    return i, j


def compute_hash_indices(cells, L: int) -> np.ndarray:
    cells = np.ascontiguousarray(cells, np.int32)
    out = np.zeros(cells.shape[0], np.int32)
    _check(lib().ref_compute_hash_indices(_p(cells, _i32p), cells.shape[0], cells.shape[1], int(L), _p(out, _i32p)))
    return out


def set_threads(n: int) -> None:
    lib().ref_set_threads(int(n))


def prepare(qpos, qsup, rpos=None, rsup=None, *, domain_min=None, domain_max=None, periodic=None, mode="gather",
            hash_map_length=4096, search_range=1):
    """The op arguments of a search: the oracle's structure of the inputs
    (sorted positions wrapped twice, cppSearch.py:221), the query positions
    wrapped once (neighborhood.py:202-203) and the sort index.  Returns a
    dict for ``run_ops``."""
    import oracle

    qpos = np.ascontiguousarray(qpos, np.float32)
    nq, dim = qpos.shape
    if rpos is None:
        rpos, rsup = qpos, qsup
    rpos = np.ascontiguousarray(rpos, np.float32)
    if domain_min is None:
        mn, mx, per = np.zeros(dim, np.float32), np.ones(dim, np.float32), [False] * dim
    else:
        mn, mx = np.asarray(domain_min, np.float32), np.asarray(domain_max, np.float32)
        per = [bool(p) for p in (periodic if periodic is not None else [False] * dim)]
    o = oracle.radius_search(qpos, qsup, None if rpos is qpos else rpos, None if rpos is qpos else rsup,
                             domain_min=None if domain_min is None else mn,
                             domain_max=None if domain_min is None else mx, periodic=per, mode=mode,
                             hash_map_length=hash_map_length, search_range=search_range, want_rows=False,
                             want_structure=True)
    q1 = oracle.wrap(qpos, mn, mx, per)               # radiusSearch's once-wrap (neighborhood.py:202-203)
    r2 = oracle.wrap(oracle.wrap(rpos, mn, mx, per), mn, mx, per)  # + cppSearch.py:221
    si = o.sort_index.astype(np.int64)
    st = Structure(r2[si], None if rsup is None else np.asarray(rsup, np.float32)[si], o.hash_table,
                   o.cell_table, o.num_cells, o.hcell, mn, mx, per)
    qs = None if qsup is None else np.ascontiguousarray(qsup, np.float32)
    return {"st": st, "q1": np.ascontiguousarray(q1), "qs": qs, "si": si, "mode": mode,
            "search_range": search_range}


def run_ops(a):
    """countNeighbors, the int32 offsets of its cumsum, buildNeighborList and
    j = sortIndex[j] (cppSearch.py:60-98) on prepared arguments.  Returns
    (counts int32[nq], i int64[P], j int64[P])."""
    st, q1, qs, si, mode, sr = a["st"], a["q1"], a["qs"], a["si"], a["mode"], a["search_range"]
    nq = q1.shape[0]
    counts = count_neighbors(st, q1, qs, mode, sr)
    offs = np.zeros(nq, np.int32)
    if nq > 1:
        offs[1:] = np.cumsum(counts, dtype=np.int64)[:-1].astype(np.int32)
    length = int(offs[-1]) + int(counts[-1]) if nq else 0
    i, j = build_neighbor_list(st, counts, offs, length, q1, qs, mode, sr)
    return counts, i, si[j] if length else j


def search(qpos, qsup, rpos=None, rsup=None, **kw):
    """The reference's count / offsets / fill / sortIndex sequence
    (cppSearch.py:60-98) over the oracle's structure of the same inputs.
    Returns (counts int32[nq], i int64[P], j int64[P])."""
    return run_ops(prepare(qpos, qsup, rpos, rsup, **kw))
