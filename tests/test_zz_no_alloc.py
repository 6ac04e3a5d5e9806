"""tcr_search_noalloc (the entry point whose workspace is the library's own
device array) against tcr_search_launch + tcr_search_finish with a
caller-provided workspace: bit-exact offsets and pairs on two calls in a
row, the second with a smaller input, a batch of clouds, and the rejection
of a search larger than the array."""
import ctypes

import numpy as np
import pytest
import torch

from torchcompactradius_amd import _lib, synthetic
from torchcompactradius_amd.search import CloudSpec, SearchPlan, _stream
from torchcompactradius_amd._lib import check, ptr

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _spec(w, rows=True):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    mask = sum(1 << d for d, p in enumerate(w.periodic) if p)
    return CloudSpec(qpos=t(w.pos), qsup=t(w.sup), rpos=None, rsup=None, domain_min=list(w.domain_min),
                     domain_max=list(w.domain_max), periodic=mask, mode=w.mode, hash_map_length=w.hash_map_length)


def _noalloc(plan, capacity, index_dtype=torch.int64, rows=True):
    L = _lib.load()
    offsets = torch.full((plan.nq + 1,), -7, dtype=torch.int64, device=DEV)
    col = torch.full((capacity,), -7, dtype=index_dtype, device=DEV)
    row = torch.full((capacity,), -7, dtype=index_dtype, device=DEV) if rows else None
    total = ctypes.c_int64(-1)
    rc = L.tcr_search_noalloc(plan.clouds, plan.n, ptr(plan.qpos), ptr(plan.qsup), ptr(plan.rpos), ptr(plan.rsup),
                              ptr(offsets), ptr(row), ptr(col), 8 if index_dtype == torch.int64 else 4, capacity,
                              ctypes.byref(total), _stream())
    torch.cuda.synchronize()
    return rc, int(total.value), offsets, row, col


def _same(plan, want_rows, index_dtype=torch.int64):
    ref = plan.run()
    torch.cuda.synchronize()
    rc, total, offsets, row, col = _noalloc(plan, ref.total + 100, index_dtype, want_rows)
    check(rc)
    assert total == ref.total
    assert torch.equal(offsets, ref.offsets)
    assert torch.equal(col[:total], ref.col)
    if want_rows:
        assert torch.equal(row[:total], ref.row)
    return total


def test_noalloc_matches_launch_twice_then_smaller():
    big = synthetic.periodic_3d(n=36)      # 46,656 points
    small = synthetic.periodic_3d(n=17)    # 4,913 points
    print("case: noalloc periodic-3d n=36 then n=17", flush=True)
    pb = SearchPlan([_spec(big)], want_rows=True)
    t1 = _same(pb, True)
    ps = SearchPlan([_spec(small)], want_rows=True)
    t2 = _same(ps, True)
    assert t1 > 40 * big.pos.shape[0] and 0 < t2 < t1
    # int32 indices, CSR-style (no rows), on the small one again
    ps32 = SearchPlan([_spec(small)], want_rows=False, index_dtype=torch.int32)
    _same(ps32, False, torch.int32)


def test_noalloc_batch_of_clouds():
    specs = []
    rng = np.random.default_rng(3)
    for k in range(5):
        dim = 2 + (k % 2)
        n = int(rng.integers(500, 4000))
        pos = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
        sup = rng.uniform(0.05, 0.15, n).astype(np.float32)
        specs.append(CloudSpec(qpos=torch.from_numpy(pos).to(DEV), qsup=torch.from_numpy(sup).to(DEV), rpos=None,
                               rsup=None, domain_min=[-1.0] * dim, domain_max=[1.0] * dim, periodic=k % 4,
                               mode=("gather", "symmetric", "scatter", "superSymmetric", "gather")[k],
                               hash_map_length=1024))
    print("case: noalloc batch of 5 clouds", flush=True)
    _same(SearchPlan(specs, want_rows=True), True)


def test_noalloc_capacity_and_size_limits():
    w = synthetic.periodic_3d(n=12)
    plan = SearchPlan([_spec(w)], want_rows=True)
    ref = plan.run()
    torch.cuda.synchronize()
    print("case: noalloc capacity too small", flush=True)
    rc, total, offsets, row, col = _noalloc(plan, ref.total // 2)
    assert rc == _lib.TCR_ERR_SPACE and total == ref.total
    assert torch.equal(offsets, ref.offsets)
    assert (col == -7).all()  # the fill did not run
    print("case: noalloc search beyond the scratch array", flush=True)
    huge = synthetic.periodic_3d(n=60)     # 216,000 points: more workspace than the array holds
    ph = SearchPlan([_spec(huge)], want_rows=False)
    rc, total, *_ = _noalloc(ph, 16, rows=False)
    assert rc == _lib.TCR_ERR_SPACE and total == -1
    assert b"scratch array" in _lib.load().tcr_last_error()
