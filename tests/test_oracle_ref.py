"""Pins the oracle against the reference's own compiled ops.

oracle/_ref/libtcr_ref.so is torchCompactRadius 0.5.5's
cpp/neighborhoodDynamic_cpu.cpp and cpp/hashing_cpu.cpp compiled from the
reference's sources (oracle/ref.mk).  On the structure the oracle builds for
the same inputs, the reference's countNeighbors / int32 offsets /
buildNeighborList / sortIndex sequence (py/compactHashing/cppSearch.py:60-98,
``oracle.ref.search``) must give the oracle's counts, rows and columns, in
order, for every mode, dimension and periodicity, with hash collisions,
points outside the domain, non-finite coordinates, searchRange 2 and the
workloads; computeHashIndices must match the oracle's hash.
"""
import numpy as np
import pytest

import oracle
from oracle import ref
from torchcompactradius_amd import synthetic

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref (the reference's own ops) is not built")


def _both(*args, **kw):
    c, i, j = ref.search(*args, **kw)
    o = oracle.radius_search(*args, **kw)
    np.testing.assert_array_equal(c, o.counts)
    np.testing.assert_array_equal(i, o.row)
    np.testing.assert_array_equal(j, o.col)
    return o


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("pmask", [0, 1, 7])
def test_random_clouds(dim, mode, pmask):
    rng = np.random.default_rng(500 + dim * 31 + pmask * 5 + len(mode))
    per = [bool((pmask >> d) & 1) for d in range(dim)]
    q = rng.uniform(-1, 1, (700, dim)).astype(np.float32)
    r = rng.uniform(-1, 1, (500, dim)).astype(np.float32)
    hq = rng.uniform(0.05, 0.25, 700).astype(np.float32)
    hr = rng.uniform(0.05, 0.25, 500).astype(np.float32)
    mn, mx = [-1.0] * dim, [1.0] * dim
    for L in (97, 4096):
        _both(q, hq, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=L)
        _both(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=L)


def test_plain_2d_1024_workload():
    w = synthetic.plain_2d_1024()
    o = _both(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic, mode=w.mode,
              hash_map_length=w.hash_map_length)
    assert 30000 < o.col.shape[0] < 36000


def test_periodic_3d_reduced_workload():
    w = synthetic.periodic_3d(n=20)
    _both(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic, mode=w.mode,
          hash_map_length=w.hash_map_length)


def test_domain_none_points_outside_and_collisions():
    rng = np.random.default_rng(6)
    q = rng.uniform(-0.7, 1.8, (1200, 2)).astype(np.float32)
    h = np.full(1200, 0.09, np.float32)
    for L in (1, 7, 4096):
        _both(q, h, hash_map_length=L)


@pytest.mark.parametrize("periodic", [False, True])
def test_non_finite_and_huge(periodic):
    rng = np.random.default_rng(8)
    q = rng.uniform(-1, 1, (900, 3)).astype(np.float32)
    q[5, 0] = np.nan
    q[17, 2] = np.inf
    q[40, 0] = -np.inf
    q[99, 1] = np.float32(1e30)
    h = np.full(900, 0.2, np.float32)
    _both(q, h, domain_min=[-1] * 3, domain_max=[1] * 3, periodic=[periodic] * 3, mode="symmetric",
          hash_map_length=331)


def test_search_range_two():
    rng = np.random.default_rng(10)
    q = rng.uniform(0, 1, (800, 2)).astype(np.float32)
    h = rng.uniform(0.05, 0.1, 800).astype(np.float32)
    _both(q, h, domain_min=[0, 0], domain_max=[1, 1], periodic=[True, False], mode="gather", hash_map_length=64,
          search_range=2)


def test_wrap_boundary_values_and_small_periodic_grids():
    lo, hi = np.float32(0.1), np.float32(0.7)
    v = np.concatenate([b + np.arange(-30, 31, dtype=np.float32) * np.spacing(b) for b in (lo, hi)])
    rng = np.random.default_rng(11)
    q = np.stack([v, rng.uniform(0.1, 0.7, v.shape[0]).astype(np.float32)], 1).astype(np.float32)
    _both(q, np.full(q.shape[0], 0.05, np.float32), domain_min=[lo, lo], domain_max=[hi, hi],
          periodic=[True, True], hash_map_length=64)
    p = rng.uniform(0, 1, (300, 2)).astype(np.float32)
    _both(p, np.full(300, 0.4, np.float32), domain_min=[0, 0], domain_max=[1, 1], periodic=[True, False],
          hash_map_length=16)


def test_compute_hash_indices():
    rng = np.random.default_rng(12)
    for dim in (1, 2, 3):
        cells = rng.integers(-2 ** 31, 2 ** 31 - 1, (400, dim)).astype(np.int32)
        cells[:50] = rng.integers(-3, 50, (50, dim))
        for L in (1, 97, 4096, 1 << 20):
            got = ref.compute_hash_indices(cells, L)
            want = np.array([oracle.hash_index(c, L) for c in cells], np.int32)
            np.testing.assert_array_equal(got, want)
