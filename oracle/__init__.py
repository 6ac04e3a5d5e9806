"""CPU oracle for the compact-hashing radius search (test infrastructure only).

``oracle/tcr_oracle.c`` restates torchCompactRadius 0.5.5's
``radiusSearch(..., algorithm='compact')`` in plain C + OpenMP; this module is
its ctypes binding plus numpy helpers.  ``oracle/naive.py`` restates the
brute-force ``algorithm='naive'`` search in numpy.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it,
as the checker; the product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .naive import naive_mod, radius_naive, radius_naive_fixed, radius_search_naive, wrap_once  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtcr_oracle.so")
_lib = None

MODES = {"symmetric": 0, "gather": 1, "scatter": 2, "superSymmetric": 3}

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_compute_h.restype = ctypes.c_float
        L.orc_compute_h.argtypes = [_f32p, _f32p, ctypes.c_float, ctypes.c_int]
        L.orc_hash_index.restype = ctypes.c_int32
        L.orc_hash_index.argtypes = [_i32p, ctypes.c_int, ctypes.c_uint32]
        L.orc_linear_index.restype = ctypes.c_int32
        L.orc_linear_index.argtypes = [_i32p, _i32p, ctypes.c_int]
        L.orc_offsets.restype = ctypes.c_int32
        L.orc_offsets.argtypes = [ctypes.c_int32, ctypes.c_int, _i32p]
        L.orc_wrap.restype = None
        L.orc_wrap.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, _f32p, ctypes.c_uint32]
        L.orc_pair_distance.restype = ctypes.c_float
        L.orc_pair_distance.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_int]
        L.orc_grid_support.restype = ctypes.c_float
        L.orc_grid_support.argtypes = [_f32p, ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_int]
        L.orc_search_create.restype = ctypes.c_void_p
        L.orc_search_create.argtypes = [_f32p, _f32p, ctypes.c_int64, _f32p, _f32p, ctypes.c_int64,
                                        ctypes.c_int, _f32p, _f32p, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_uint32, ctypes.c_int32, ctypes.c_int]
        L.orc_search_total.restype = ctypes.c_int64
        L.orc_search_total.argtypes = [ctypes.c_void_p]
        L.orc_search_counts.restype = _i32p
        L.orc_search_counts.argtypes = [ctypes.c_void_p]
        L.orc_search_map.restype = ctypes.c_void_p
        L.orc_search_map.argtypes = [ctypes.c_void_p]
        L.orc_search_hcell.restype = ctypes.c_float
        L.orc_search_hcell.argtypes = [ctypes.c_void_p]
        L.orc_search_ncells.restype = ctypes.c_int64
        L.orc_search_ncells.argtypes = [ctypes.c_void_p]
        L.orc_search_fill.restype = None
        L.orc_search_fill.argtypes = [ctypes.c_void_p, _i64p, _i64p]
        L.orc_search_free.restype = None
        L.orc_search_free.argtypes = [ctypes.c_void_p]
        L.orc_map_export.restype = None
        L.orc_map_export.argtypes = [ctypes.c_void_p, _i32p, _f32p, _i32p, _i32p, _i32p]
        L.orc_brute_counts.restype = None
        L.orc_brute_counts.argtypes = [_f32p, _f32p, ctypes.c_int64, _f32p, _f32p, ctypes.c_int64,
                                       ctypes.c_int, _f32p, _f32p, ctypes.c_uint32, ctypes.c_int, _i32p]
        L.orc_small.restype = None
        L.orc_small.argtypes = [_f32p, _f32p, ctypes.c_int64, _f32p, _f32p, ctypes.c_int64, ctypes.c_int, _f32p,
                                _f32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_float, _i32p, _i64p,
                                _i64p, _i64p]
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray], t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _mask(periodic, dim) -> int:
    if periodic is None:
        return 0
    if isinstance(periodic, (bool, np.bool_)):
        periodic = [bool(periodic)] * dim
    m = 0
    for d, p in enumerate(periodic):
        if p:
            m |= 1 << d
    return m


def compute_h(mn, mx, hmax) -> np.float32:
    mn, mx = _f32(mn), _f32(mx)
    return np.float32(lib().orc_compute_h(_p(mn, _f32p), _p(mx, _f32p), float(np.float32(hmax)), mn.shape[0]))


def hash_index(cell, L) -> int:
    c = np.ascontiguousarray(np.asarray(cell, dtype=np.int32))
    return int(lib().orc_hash_index(_p(c, _i32p), c.shape[0], int(L)))


def linear_index(cell, num_cells) -> int:
    c = np.ascontiguousarray(np.asarray(cell, dtype=np.int32))
    n = np.ascontiguousarray(np.asarray(num_cells, dtype=np.int32))
    return int(lib().orc_linear_index(_p(c, _i32p), _p(n, _i32p), c.shape[0]))


def offsets(search_range: int, dim: int) -> np.ndarray:
    n = (2 * search_range + 1) ** dim
    out = np.zeros((n, dim), dtype=np.int32)
    lib().orc_offsets(search_range, dim, _p(out, _i32p))
    return out


def wrap(pos, mn, mx, periodic) -> np.ndarray:
    p = _f32(pos).copy()
    dim = p.shape[1]
    mn, mx = _f32(mn), _f32(mx)
    lib().orc_wrap(_p(p, _f32p), p.shape[0], dim, _p(mn, _f32p), _p(mx, _f32p), _mask(periodic, dim))
    return p


def pair_distance(xi, xj, mn, mx, periodic) -> np.float32:
    xi, xj, mn, mx = _f32(xi), _f32(xj), _f32(mn), _f32(mx)
    dim = xi.shape[0]
    return np.float32(lib().orc_pair_distance(_p(xi, _f32p), _p(xj, _f32p), _p(mn, _f32p), _p(mx, _f32p),
                                              _mask(periodic, dim), dim))


@dataclass
class OracleResult:
    counts: np.ndarray       # int32 [Nq]
    row: Optional[np.ndarray]  # int64 [P]
    col: np.ndarray          # int64 [P]
    hcell: float
    num_cells: np.ndarray    # int32 [dim]
    ncells: int
    sort_index: Optional[np.ndarray] = None
    hash_table: Optional[np.ndarray] = None
    cell_table: Optional[np.ndarray] = None

    @property
    def indptr(self) -> np.ndarray:
        ip = np.zeros(self.counts.shape[0] + 1, dtype=np.int64)
        np.cumsum(self.counts, out=ip[1:])
        return ip.astype(np.int32)


def radius_search(qpos, qsup, rpos=None, rsup=None, *, domain_min=None, domain_max=None,
                  periodic=None, mode="gather", hash_map_length=4096, search_range=1,
                  want_rows=True, want_structure=False, threads: int = 0) -> OracleResult:
    """radiusSearch(..., algorithm='compact') on the CPU.  ``domain_min=None``
    means the reference's default domain [0,1]^d, non-periodic."""
    qpos = _f32(qpos)
    nq, dim = qpos.shape
    qsup = None if qsup is None else _f32(qsup)
    self_search = rpos is None
    if not self_search:
        rpos = _f32(rpos)
        rsup = None if rsup is None else _f32(rsup)
    nr = nq if self_search else rpos.shape[0]
    if domain_min is None:
        mn = np.zeros(dim, np.float32)
        mx = np.ones(dim, np.float32)
        mask = 0
    else:
        mn, mx = _f32(domain_min), _f32(domain_max)
        mask = _mask(periodic, dim)
    L = lib()
    h = L.orc_search_create(_p(qpos, _f32p), _p(qsup, _f32p), nq,
                            _p(None if self_search else rpos, _f32p), _p(None if self_search else rsup, _f32p),
                            nr, dim, _p(mn, _f32p), _p(mx, _f32p), mask, MODES[mode],
                            int(hash_map_length), int(search_range), int(threads))
    if not h:
        raise ValueError("oracle rejected the arguments")
    try:
        total = L.orc_search_total(h)
        counts = np.ctypeslib.as_array(L.orc_search_counts(h), shape=(nq,)).copy() if nq else np.zeros(0, np.int32)
        col = np.empty(total, dtype=np.int64)
        row = np.empty(total, dtype=np.int64) if want_rows else None
        L.orc_search_fill(h, _p(row, _i64p), _p(col, _i64p))
        m = L.orc_search_map(h)
        ncells = L.orc_search_ncells(h)
        nc = np.zeros(dim, np.int32)
        res = OracleResult(counts=counts, row=row, col=col, hcell=float(L.orc_search_hcell(h)),
                           num_cells=nc, ncells=int(ncells))
        si = np.zeros(nr, np.int32) if want_structure else None
        ht = np.zeros((int(hash_map_length), 2), np.int32) if want_structure else None
        ct = np.zeros((int(ncells), 3), np.int32) if want_structure else None
        L.orc_map_export(m, _p(si, _i32p), _p(None, _f32p), _p(ct, _i32p), _p(ht, _i32p), _p(nc, _i32p))
        res.sort_index, res.hash_table, res.cell_table = si, ht, ct
        return res
    finally:
        L.orc_search_free(h)


def brute_counts(qpos, qsup, rpos=None, rsup=None, *, domain_min=None, domain_max=None, periodic=None,
                 mode="gather") -> np.ndarray:
    qpos = _f32(qpos)
    nq, dim = qpos.shape
    qsup = None if qsup is None else _f32(qsup)
    if rpos is None:
        rpos, rsup = qpos, qsup
    rpos = _f32(rpos)
    rsup = None if rsup is None else _f32(rsup)
    if domain_min is None:
        mn, mx, mask = np.zeros(dim, np.float32), np.ones(dim, np.float32), 0
    else:
        mn, mx, mask = _f32(domain_min), _f32(domain_max), _mask(periodic, dim)
    out = np.zeros(nq, np.int32)
    lib().orc_brute_counts(_p(qpos, _f32p), _p(qsup, _f32p), nq, _p(rpos, _f32p), _p(rsup, _f32p), rpos.shape[0],
                           dim, _p(mn, _f32p), _p(mx, _f32p), mask, MODES[mode], _p(out, _i32p))
    return out


def radius_search_small(qpos, qsup, rpos=None, rsup=None, *, domain_min=None, domain_max=None, periodic=None,
                        mode: str = "gather", support_override: Optional[float] = None):
    """radiusSearch(..., algorithm='small') as the reference's commented-out
    op defines it (cpp/neighborhoodSmall_cpu.cpp): default domain [0,1]^d
    non-periodic; with a domain, positions wrapped once (util.py:375-387);
    (ii, jj) int64 in (query, reference) order, modDistance < support
    (fixed: modDistance2 < h*h)."""
    qpos = _f32(qpos)
    nq, dim = qpos.shape
    if rpos is None:
        rpos, rsup = qpos, qsup
    rpos = _f32(rpos)
    if domain_min is None:
        mn, mx, mask = np.zeros(dim, np.float32), np.ones(dim, np.float32), 0
    else:
        mn, mx, mask = _f32(domain_min), _f32(domain_max), _mask(periodic, dim)
        qpos = wrap(qpos, mn, mx, periodic)
        rpos = wrap(rpos, mn, mx, periodic)
    qs = None if qsup is None else _f32(qsup)
    rs = None if rsup is None else _f32(rsup)
    fixed = support_override is not None
    h = float(support_override) if fixed else 0.0
    args = (_p(qpos, _f32p), _p(qs, _f32p), nq, _p(rpos, _f32p), _p(rs, _f32p), rpos.shape[0], dim,
            _p(mn, _f32p), _p(mx, _f32p), mask, MODES[mode], int(fixed), ctypes.c_float(h))
    counts = np.zeros(nq, np.int32)
    lib().orc_small(*args, _p(counts, _i32p), None, None, None)
    offs = np.zeros(nq + 1, np.int64)
    np.cumsum(counts, out=offs[1:])
    ii = np.zeros(int(offs[-1]), np.int64)
    jj = np.zeros(int(offs[-1]), np.int64)
    lib().orc_small(*args, None, _p(offs, _i64p), _p(ii, _i64p), _p(jj, _i64p))
    return ii, jj
