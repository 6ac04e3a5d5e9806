"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/tcr.h declares, and its host-only entry points (workspace sizing,
argument validation) behave without touching a GPU."""
import ctypes
import os
import re

import pytest

from torchcompactradius_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "tcr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tcr_[a-z_]+)\s*\(", txt)))


def test_library_loads_and_exports_header_symbols():
    L = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 9
    for s in syms:
        assert hasattr(L, s), f"missing export {s}"
    assert set(syms) == set(_lib.EXPORTED)


def test_version_string():
    assert b"gfx950" in _lib.load().tcr_version()


def _cloud(**kw):
    c = _lib.TcrCloud()
    c.dim, c.mode, c.periodic_mask, c.search_range = 3, 1, 0, 1
    for d in range(3):
        c.domain_min[d], c.domain_max[d] = -1.0, 1.0
    c.hash_map_length, c.self_search = 4096, 1
    c.nq, c.nr = 1000, 1000
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_workspace_size_scales_with_points():
    L = _lib.load()
    small = L.tcr_search_workspace_size(ctypes.byref(_cloud(nq=1000)), 1)
    big = L.tcr_search_workspace_size(ctypes.byref(_cloud(nq=100000)), 1)
    assert small > 0 and big > small * 50


@pytest.mark.parametrize("field,value,msg", [("dim", 4, "dim"), ("mode", 7, "mode"),
                                             ("search_range", 0, "search_range"),
                                             ("hash_map_length", 0, "hashMapLength"),
                                             ("periodic_mask", 8, "periodic")])
def test_invalid_cloud_rejected(field, value, msg):
    L = _lib.load()
    assert L.tcr_search_workspace_size(ctypes.byref(_cloud(**{field: value})), 1) == 0
    assert msg in L.tcr_last_error().decode()


def test_search_count_rejects_small_workspace():
    L = _lib.load()
    c = _cloud(nq=10)
    total = ctypes.c_int64()
    rc = L.tcr_search_count(ctypes.byref(c), 1, 16, 16, None, None, 64, 8, 64, ctypes.byref(total), None)
    assert rc == _lib.TCR_ERR_SPACE
    assert "workspace" in L.tcr_last_error().decode()


def test_compute_hash_indices_rejects_bad_dim():
    L = _lib.load()
    assert L.tcr_compute_hash_indices(None, 0, 5, 16, None, None) == _lib.TCR_ERR_ARG
