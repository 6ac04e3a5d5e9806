"""GPU parity of the multi-level search (algorithm='mlm'): the same pairs in
the same order as the compact search and the oracle, for variable supports
in every mode that bounds the radius by the reference supports, 1-3 dims,
periodic and not; inputs it does not take fall back to the compact search."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch, synthetic
from torchcompactradius_amd._lib import Unsupported
from torchcompactradius_amd.search import CloudSpec, run_mlm

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _search(pos, sup, mn, mx, per, mode, algorithm, fmt="coo", rpos=None, rsup=None, index_dtype=torch.int64):
    dom = DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                            torch.tensor(per), len(mn))
    ref = None if rpos is None else PointCloud(_t(rpos), _t(rsup))
    out = radiusSearch(PointCloud(_t(pos), _t(sup)), ref, mode=mode, domain=dom, hashMapLength=4096,
                       algorithm=algorithm, format=fmt, indexDtype=index_dtype)
    torch.cuda.synchronize()
    return out


def _spec(pos, sup, mn, mx, per, mode):
    return CloudSpec(qpos=_t(pos), qsup=_t(sup), rpos=None, rsup=None, domain_min=list(mn), domain_max=list(mx),
                     periodic=sum(1 << d for d in range(len(per)) if per[d]), mode=mode, hash_map_length=4096)


def test_adaptive_cloud_matches_oracle_and_compact():
    w = synthetic.adaptive_3d(n=30000, seed=3)
    print("case: adaptive 30000 symmetric", flush=True)
    spec = _spec(w.pos, w.sup, w.domain_min, w.domain_max, w.periodic, "symmetric")
    r = run_mlm(spec, want_rows=True)          # raises if it fell back
    torch.cuda.synchronize()
    o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode="symmetric", hash_map_length=4096)
    np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
    np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)
    c = _search(w.pos, w.sup, w.domain_min, w.domain_max, w.periodic, "symmetric", "compact")
    assert torch.equal(c.row, r.row) and torch.equal(c.col, r.col)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("pmask", [0, 5])
def test_modes_dims_periodicity(dim, mode, pmask):
    rng = np.random.default_rng(700 + dim * 10 + pmask + len(mode))
    n = 4000
    per = [bool((pmask >> d) & 1) for d in range(dim)]
    pos = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    base = {1: 0.004, 2: 0.03, 3: 0.05}[dim]
    sup = (base * np.exp2(rng.uniform(0, 3, n))).astype(np.float32)   # ratio up to 8
    mn, mx = [-1.0] * dim, [1.0] * dim
    print(f"case: mlm dim={dim} mode={mode} pmask={pmask}", flush=True)
    r = run_mlm(_spec(pos, sup, mn, mx, per, mode), want_rows=True)
    torch.cuda.synchronize()
    o = oracle.radius_search(pos, sup, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=4096)
    np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
    np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)


def test_int32_csr_and_levels_capped():
    w = synthetic.adaptive_3d(n=20000, seed=9, ratio=8.0)
    print("case: mlm int32 csr", flush=True)
    for levels in (1, 2, 4):
        r = run_mlm(_spec(w.pos, w.sup, w.domain_min, w.domain_max, w.periodic, "symmetric"), want_rows=False,
                    index_dtype=torch.int32, max_levels=levels)
        torch.cuda.synchronize()
        o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max,
                                 periodic=w.periodic, mode="symmetric", hash_map_length=4096, want_rows=False)
        assert r.col.dtype == torch.int32
        np.testing.assert_array_equal(r.col.cpu().numpy().astype(np.int64), o.col)
        np.testing.assert_array_equal(r.offsets.cpu().numpy(), o.indptr.astype(np.int64))


def test_unsupported_inputs_fall_back_to_compact():
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    sup = rng.uniform(0.05, 0.3, 3000).astype(np.float32)
    mn, mx = [-1.0] * 3, [1.0] * 3
    print("case: mlm gather falls back", flush=True)
    with pytest.raises(Unsupported):
        run_mlm(_spec(pos, sup, mn, mx, [False] * 3, "gather"), want_rows=True)
    out_pos = pos.copy()
    out_pos[:10] *= 1.5                       # outside [-1, 1]: outside the grid
    with pytest.raises(Unsupported):
        run_mlm(_spec(out_pos, sup, mn, mx, [False] * 3, "symmetric"), want_rows=True)
    for p, mode in ((pos, "gather"), (out_pos, "symmetric")):
        a = _search(p, sup, mn, mx, [False] * 3, mode, "mlm")
        o = oracle.radius_search(p, sup, domain_min=mn, domain_max=mx, periodic=[False] * 3, mode=mode,
                                 hash_map_length=4096)
        np.testing.assert_array_equal(a.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(a.col.cpu().numpy(), o.col)


def test_non_self_search():
    rng = np.random.default_rng(12)
    q = rng.uniform(-1, 1, (2500, 3)).astype(np.float32)
    r = rng.uniform(-1, 1, (3500, 3)).astype(np.float32)
    hq = (0.05 * np.exp2(rng.uniform(0, 3, 2500))).astype(np.float32)
    hr = (0.05 * np.exp2(rng.uniform(0, 3, 3500))).astype(np.float32)
    mn, mx = [-1.0] * 3, [1.0] * 3
    for mode in ("symmetric", "scatter"):
        print(f"case: mlm non-self {mode}", flush=True)
        spec = CloudSpec(qpos=_t(q), qsup=_t(hq), rpos=_t(r), rsup=_t(hr), domain_min=mn, domain_max=mx,
                         periodic=5, mode=mode, hash_map_length=4096)
        a = run_mlm(spec, want_rows=True)
        torch.cuda.synchronize()
        o = oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=[True, False, True],
                                 mode=mode, hash_map_length=4096)
        np.testing.assert_array_equal(a.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(a.col.cpu().numpy(), o.col)


@pytest.mark.parametrize("mode", ["scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("dim", [2, 3])
def test_distances_at_the_support(mode, dim):
    """A lattice whose pair distances equal the supports or lie a rounding
    step away from them: the fp32 estimate of a cloud without periodic
    dimensions leaves these to the reference formula, which must match the
    oracle bit for bit."""
    rng = np.random.default_rng(880 + dim + len(mode))
    g = np.arange(-7, 8, dtype=np.float32) * np.float32(0.125)       # exact in fp32
    grid = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), axis=-1).reshape(-1, dim).astype(np.float32)
    # supports exactly lattice distances (1, sqrt(2), 2 spacings) and their fp32 neighbours
    cand = np.array([0.125, 0.25, np.sqrt(2) * 0.125, 0.375], np.float32)
    sup = cand[rng.integers(0, cand.size, grid.shape[0])]
    step = rng.integers(-1, 2, grid.shape[0])
    sup = np.where(step < 0, np.nextafter(sup, np.float32(0)), np.where(step > 0, np.nextafter(sup, np.float32(1)), sup))
    sup = sup.astype(np.float32)
    mn, mx = [-1.0] * dim, [1.0] * dim
    per = [False] * dim
    print(f"case: mlm support ties mode={mode} dim={dim}", flush=True)
    r = run_mlm(_spec(grid, sup, mn, mx, per, mode), want_rows=True)
    torch.cuda.synchronize()
    o = oracle.radius_search(grid, sup, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=4096)
    assert o.row.shape[0] > 1000
    np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
    np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)


@pytest.mark.parametrize("cap", [1, 4, 2048])
def test_slot_capacity_and_overflow(cap, monkeypatch):
    """The count pass keeps each row's keys in a slot of MLM_SLOT_CAP; rows
    longer than it send the fill back to a second traversal.  Either way the
    output is the oracle's."""
    import torchcompactradius_amd.search as S
    monkeypatch.setattr(S, "MLM_SLOT_CAP", cap)
    rng = np.random.default_rng(990 + cap)
    pos = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    sup = (0.06 * np.exp2(rng.uniform(0, 3, 3000))).astype(np.float32)
    mn, mx, per = [-1.0] * 3, [1.0] * 3, [False] * 3
    print(f"case: mlm slot cap={cap}", flush=True)
    for fmt_rows, dt in ((True, torch.int64), (False, torch.int32)):
        r = run_mlm(_spec(pos, sup, mn, mx, per, "symmetric"), want_rows=fmt_rows, index_dtype=dt)
        torch.cuda.synchronize()
        o = oracle.radius_search(pos, sup, domain_min=mn, domain_max=mx, periodic=per, mode="symmetric",
                                 hash_map_length=4096)
        assert o.counts.max() > 4
        if fmt_rows:
            np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(r.col.cpu().numpy().astype(np.int64), o.col)


@pytest.mark.parametrize("case", ["four_periodic_cells", "wrap_strays", "dense_runs"])
def test_multilevel_edge_inputs_match_oracle(case):
    """Inputs at the edges of the multi-level traversal: a periodic dimension
    of 4 hCell cells (neighbour columns wrap onto each other's images),
    points just below the domain minimum whose once-wrapped copy (the query)
    differs from the twice-wrapped one (the reference), and a dense core of
    small supports inside a sparse cloud of large ones (groups searched by
    sub-cell)."""
    rng = np.random.default_rng({"four_periodic_cells": 901, "wrap_strays": 902, "dense_runs": 903}[case])
    mn, mx = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    if case == "four_periodic_cells":
        n, per = 3000, [True, False, True]
        pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        sup = (0.12 * np.exp2(rng.uniform(0, 2, n))).astype(np.float32)
        sup[0] = 0.5                                     # hMax 0.5: 4 cells per dimension
    elif case == "wrap_strays":
        n, per = 4000, [True, True, True]
        pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        pos[:200, 0] = np.float32(-1.0) - np.float32(1e-8) * rng.integers(1, 4, 200).astype(np.float32)
        pos[200:400, 2] = np.nextafter(np.float32(-1.0), np.float32(-2.0))
        sup = (0.05 * np.exp2(rng.uniform(0, 3, n))).astype(np.float32)
    else:
        n, per = 20000, [False, False, False]
        c = rng.normal(0.0, 0.05, (n // 2, 3))
        u = rng.uniform(-1, 1, (n - n // 2, 3))
        pos = np.clip(np.concatenate([c, u]), -0.999, 0.999).astype(np.float32)
        sup = np.where(np.abs(pos).max(axis=1) < 0.15, 0.02, 0.16).astype(np.float32)
    print(f"case: mlm edge inputs {case}", flush=True)
    for mode in ("symmetric", "superSymmetric", "scatter"):
        r = run_mlm(_spec(pos, sup, mn, mx, per, mode), want_rows=True)
        torch.cuda.synchronize()
        o = oracle.radius_search(pos, sup, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                                 hash_map_length=4096)
        np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)
