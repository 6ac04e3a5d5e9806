"""GPU parity of radiusSearch(algorithm='mlm', mlmSemantics='reference') --
the reference's own multi-level semantics on libtcr_hip.so's tcr_mlmref_*
kernels -- against its restatement oracle/mlm_ref.py: identical COO pairs
and order, the CSR with its length-Nr offsets, every mode (scatter as the
transposed gather search), 1-3 dimensions, periodic and not, hash-map
lengths with collisions, points outside the domain, and the cases where the
reference's output is undefined (both refuse them)."""
import numpy as np
import pytest
import torch

from oracle import mlm_ref as M
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch

pytestmark = pytest.mark.gpu
DEV = "cuda"
F32 = np.float32


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _gpu(q, hq, r, hr, mn, mx, per, mode, L, fmt):
    dom = DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                            torch.tensor(per, dtype=torch.bool), len(mn))
    out = radiusSearch(PointCloud(_t(q), _t(hq)), None if r is None else PointCloud(_t(r), _t(hr)), mode=mode,
                       domain=dom, hashMapLength=L, algorithm="mlm", format=fmt, mlmSemantics="reference")
    torch.cuda.synchronize()
    return out


def _same(out, want, fmt):
    if fmt == "coo":
        np.testing.assert_array_equal(out.row.cpu().numpy(), want[1])
        np.testing.assert_array_equal(out.col.cpu().numpy(), want[2])
    else:
        np.testing.assert_array_equal(out.indices.cpu().numpy(), want[1])
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), want[2])
        np.testing.assert_array_equal(out.rowEntries.cpu().numpy(), want[3])
        assert out.indptr.dtype == torch.int32


def _check(q, hq, r, hr, mn, mx, per, mode, L, fmt):
    """The GPU result equals the restatement's, or both refuse the case (the
    reference's output is undefined there)."""
    try:
        want = M.search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=L,
                        fmt=fmt)
    except ValueError:
        with pytest.raises((ValueError, RuntimeError)):
            _gpu(q, hq, r, hr, mn, mx, per, mode, L, fmt)
        return False
    _same(_gpu(q, hq, r, hr, mn, mx, per, mode, L, fmt), want, fmt)
    return True


def _cloud(rng, n, dim, spread=8.0, lo=0.0, hi=1.0):
    p = rng.uniform(lo, hi, (n, dim)).astype(F32)
    h = (0.035 * np.exp2(rng.uniform(0, np.log2(spread), n))).astype(F32)
    return p, h


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "symmetric", "superSymmetric", "scatter"])
@pytest.mark.parametrize("periodic", [False, True])
def test_reference_mlm_self_search(dim, mode, periodic):
    rng = np.random.default_rng(600 + 7 * dim + 3 * len(mode) + periodic)
    n = {1: 300, 2: 700, 3: 700}[dim]
    # superSymmetric rows are built at the level of the synced support; at a
    # level of 2 cells per dimension the C remainder visits a cell twice and
    # the rows outgrow their counts (undefined in the reference): a narrower
    # support ratio keeps its levels at 16 cells and more
    p, h = _cloud(rng, n, dim, spread=1.6 if mode == "superSymmetric" else 8.0)
    mn, mx, per = [0.0] * dim, [1.0] * dim, [periodic] * dim
    print(f"case: mlm reference dim={dim} mode={mode} periodic={periodic}", flush=True)
    for fmt in (("coo", "csr") if dim == 2 else ("coo",)):
        defined = _check(p, h, None, None, mn, mx, per, mode, 4096, fmt)
        # a periodic superSymmetric search pairs points across a periodic face
        # that the one-sided C-remainder stepping visits from one side only
        assert defined or (periodic and mode == "superSymmetric")


@pytest.mark.parametrize("L", [1, 7, 331])
def test_hash_collisions_and_points_outside(L):
    """Hash-map lengths that put many cells in one bucket, and points below
    the domain minimum (negative cells: the C remainder keeps them negative
    and their Morton codes decode elsewhere)."""
    rng = np.random.default_rng(700 + L)
    p, h = _cloud(rng, 700, 2, lo=-0.15, hi=1.0)
    mn, mx, per = [0.0, 0.0], [1.0, 1.0], [False, False]
    print(f"case: mlm reference L={L}", flush=True)
    assert _check(p, h, None, None, mn, mx, per, "gather", L, "coo")


def test_three_d_levels_and_a_shifted_domain():
    rng = np.random.default_rng(710)
    p, h = _cloud(rng, 1500, 3, spread=8.0, lo=-0.6, hi=1.4)
    mn, mx = [-0.6, -0.6, -0.6], [1.4, 1.4, 1.4]
    print("case: mlm reference 3-D multi-level shifted domain", flush=True)
    assert _check(p, h, None, None, mn, mx, [True, False, True], "gather", 97, "coo")
    _check(p, h, None, None, mn, mx, [False] * 3, "superSymmetric", 97, "coo")


def test_query_subset_of_reference_and_undefined_cases():
    """Queries that are the first Nq reference points: their rows fit the
    reference counts.  More queries than reference points, or a query set
    whose rows outgrow those counts: the reference's output is undefined and
    both refuse."""
    rng = np.random.default_rng(720)
    r, hr = _cloud(rng, 800, 2)
    mn, mx, per = [0.0, 0.0], [1.0, 1.0], [False, False]
    print("case: mlm reference non-self", flush=True)
    for fmt in ("coo", "csr"):
        assert _check(r[:500], hr[:500], r, hr, mn, mx, per, "gather", 4096, fmt)
    q, hq = _cloud(rng, 300, 2)
    for args in ((r, hr, r[:200], hr[:200]), (q, hq * 3, r, hr)):
        assert not _check(*args, mn, mx, per, "gather", 4096, "coo")
