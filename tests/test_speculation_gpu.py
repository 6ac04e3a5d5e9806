"""GPU parity of repeated searches on one workspace.

A repeated search on the same input tensors reuses its plan (workspace,
offsets, descriptors), and the library launches only the radix passes, list
kernels and per-query kernel the previous search on that workspace needed.
When a search needs more (more cells after its supports shrink, points
outside the domain, crowded cells, a grid the cell kernels do not take), the
library repeats it with every launch before the pair total is returned.  Each search here is checked against the oracle on the
values it ran on.
"""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _domain(mn, mx, periodic):
    return DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                             torch.tensor(periodic, dtype=torch.bool), len(mn))


def _check(pos_t, sup_t, dom, mn, mx, periodic, fmt, L=4096):
    out = radiusSearch(PointCloud(pos_t, sup_t), None, mode="gather", domain=dom, hashMapLength=L,
                       algorithm="compact", format=fmt)
    torch.cuda.synchronize()
    o = oracle.radius_search(pos_t.cpu().numpy(), sup_t.cpu().numpy(), domain_min=mn, domain_max=mx,
                             periodic=periodic, mode="gather", hash_map_length=L)
    if fmt == "coo":
        np.testing.assert_array_equal(out.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(out.col.cpu().numpy(), o.col)
    else:
        np.testing.assert_array_equal(out.indices.cpu().numpy(), o.col)
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), o.indptr)
    return int(o.col.shape[0])


@pytest.mark.parametrize("fmt", ["csr", "coo"])
def test_supports_shrink_between_searches(fmt):
    """Few cells (one radix pass), then many (two, then three passes): the
    search after each change needs passes the one before it did not."""
    rng = np.random.default_rng(11)
    n = 20000
    mn, mx, periodic = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [True, True, False]
    pos = torch.from_numpy(rng.uniform(-1, 1, (n, 3)).astype(np.float32)).to(DEV)
    sup = torch.full((n,), 0.3, dtype=torch.float32, device=DEV)   # 6^3 cells
    dom = _domain(mn, mx, periodic)
    for h in (0.3, 0.3, 0.3, 0.05, 0.05, 0.012, 0.012, 0.3):       # 40^3 cells, then 166^3 (> 2^18)
        print(f"case: shrink fmt={fmt} h={h}", flush=True)
        sup.fill_(h)
        _check(pos, sup, dom, mn, mx, periodic, fmt)


def test_strays_and_crowded_cells_appear():
    """Searches without listed cells or queries, then points outside the
    domain (stray queries) and a dense cluster (cells over the register
    chunk), then back."""
    rng = np.random.default_rng(12)
    n = 30000
    mn, mx, periodic = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [False, True, False]
    base = rng.uniform(-0.95, 0.95, (n, 3)).astype(np.float32)
    pos = torch.from_numpy(base).to(DEV)
    sup = torch.full((n,), 0.08, dtype=torch.float32, device=DEV)
    dom = _domain(mn, mx, periodic)
    for step in range(3):
        print(f"case: regular step {step}", flush=True)
        _check(pos, sup, dom, mn, mx, periodic, "csr")
    moved = base.copy()
    moved[:50, 0] = np.float32(-1.0) - np.float32(0.01) * np.arange(1, 51, dtype=np.float32)  # outside in x
    moved[50:3050] = (rng.normal(0.0, 0.02, (3000, 3))).astype(np.float32)                   # a dense cluster
    pos.copy_(torch.from_numpy(moved))
    for step in range(2):
        print(f"case: strays and crowded cells step {step}", flush=True)
        p = _check(pos, sup, dom, mn, mx, periodic, "csr")
        assert p > 0
    pos.copy_(torch.from_numpy(base))
    print("case: back to regular", flush=True)
    _check(pos, sup, dom, mn, mx, periodic, "csr")


def test_plan_graph_replay_after_speculative_searches():
    """A captured graph (every launch) replayed on a workspace whose last
    search was speculative: the replay's results stand alone."""
    from torchcompactradius_amd.search import CloudSpec, SearchPlan
    rng = np.random.default_rng(13)
    n = 8000
    mn, mx = [-1.0, -1.0], [1.0, 1.0]
    pos = torch.from_numpy(rng.uniform(-1, 1, (n, 2)).astype(np.float32)).to(DEV)
    sup = torch.full((n,), 0.05, dtype=torch.float32, device=DEV)
    spec = CloudSpec(pos, sup, None, None, mn, mx, 0b11, "gather", 4096)
    plan = SearchPlan([spec], want_rows=True)
    o = oracle.radius_search(pos.cpu().numpy(), sup.cpu().numpy(), domain_min=mn, domain_max=mx,
                             periodic=[True, True], mode="gather", hash_map_length=4096)
    for _ in range(3):
        r = plan.run()
        np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)
    plan.capture(int(o.col.shape[0]) + 64)
    for _ in range(2):
        r = plan.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)
        np.testing.assert_array_equal(r.row.cpu().numpy(), o.row)
        r = plan.run()
        np.testing.assert_array_equal(r.col.cpu().numpy(), o.col)
    plan.release_graph()


def test_query_cells_lose_the_cell_path():
    """A search whose query set is not the reference set: several searches
    on the query-cell path, then supports large enough that a periodic
    dimension has fewer than 5 cells (the per-query kernel takes the
    queries), then back."""
    rng = np.random.default_rng(14)
    nq, nr = 6000, 7000
    mn, mx, periodic = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [True, False, True]
    q = torch.from_numpy(rng.uniform(-1, 1, (nq, 3)).astype(np.float32)).to(DEV)
    r = torch.from_numpy(rng.uniform(-1, 1, (nr, 3)).astype(np.float32)).to(DEV)
    hq = torch.full((nq,), 0.1, dtype=torch.float32, device=DEV)
    hr = torch.full((nr,), 0.1, dtype=torch.float32, device=DEV)
    dom = _domain(mn, mx, periodic)
    for h in (0.1, 0.1, 0.1, 0.45, 0.45, 0.1):      # 20 cells per dimension, then 4
        print(f"case: query cells h={h}", flush=True)
        hq.fill_(h)
        hr.fill_(h)
        out = radiusSearch(PointCloud(q, hq), PointCloud(r, hr), mode="gather", domain=dom, hashMapLength=4096,
                           algorithm="compact", format="csr")
        torch.cuda.synchronize()
        o = oracle.radius_search(q.cpu().numpy(), hq.cpu().numpy(), r.cpu().numpy(), hr.cpu().numpy(),
                                 domain_min=mn, domain_max=mx, periodic=periodic, mode="gather",
                                 hash_map_length=4096)
        np.testing.assert_array_equal(out.indices.cpu().numpy(), o.col)
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), o.indptr)


def test_fill_repeated_after_one_count():
    """The fill pass launched several times after one count pass (each
    launch of a cell kernel takes its cells from a counter the previous
    launch on that counter has put back)."""
    import ctypes
    from torchcompactradius_amd import search as S
    from torchcompactradius_amd._lib import check
    rng = np.random.default_rng(15)
    n = 12000
    mn, mx = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    pos = torch.from_numpy(rng.uniform(-1, 1, (n, 3)).astype(np.float32)).to(DEV)
    sup = torch.full((n,), 0.1, dtype=torch.float32, device=DEV)
    spec = S.CloudSpec(pos, sup, None, None, mn, mx, 0b101, "gather", 4096)
    plan = S.SearchPlan([spec], want_rows=True)
    o = oracle.radius_search(pos.cpu().numpy(), sup.cpu().numpy(), domain_min=mn, domain_max=mx,
                             periodic=[True, False, True], mode="gather", hash_map_length=4096)
    stream = S._stream()
    total = ctypes.c_int64(0)
    check(plan.L.tcr_search_count(*plan._args(), S.ptr(plan.offsets), ctypes.byref(total), stream))
    P = int(total.value)
    assert P == o.col.shape[0]
    for k in range(3):
        print(f"case: fill {k} after one count", flush=True)
        row, col = plan._fill_exact(P, stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(col.cpu().numpy(), o.col)
        np.testing.assert_array_equal(row.cpu().numpy(), o.row)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_repeated_searches_by_dimension(dim):
    """Repeated searches of one cloud in 1, 2 and 3 dimensions (the later
    ones on the launch history of the earlier), then moved points; every
    search against the oracle."""
    rng = np.random.default_rng(20 + dim)
    n = {1: 5000, 2: 20000, 3: 30000}[dim]
    mn, mx = [-1.0] * dim, [1.0] * dim
    periodic = [True, False, True][:dim]
    pos = torch.from_numpy(rng.uniform(-1, 1, (n, dim)).astype(np.float32)).to(DEV)
    h = {1: 0.002, 2: 0.03, 3: 0.1}[dim]   # 1000, 66^2 and 20^3 cells
    sup = torch.full((n,), h, dtype=torch.float32, device=DEV)
    dom = _domain(mn, mx, periodic)
    for step in range(4):
        print(f"case: repeated dim={dim} step {step}", flush=True)
        _check(pos, sup, dom, mn, mx, periodic, "csr")
    pos.copy_(torch.from_numpy(rng.uniform(-1, 1, (n, dim)).astype(np.float32)))
    print(f"case: repeated dim={dim} moved", flush=True)
    _check(pos, sup, dom, mn, mx, periodic, "coo")


def test_crowded_cell_appears():
    """Searches of a regular cloud, then of the same cloud with 3000 points
    gathered into one cell (a cell far over the register chunk)."""
    rng = np.random.default_rng(24)
    n = 20000
    mn, mx, periodic = [-1.0, -1.0], [1.0, 1.0], [True, True]
    base = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    pos = torch.from_numpy(base).to(DEV)
    sup = torch.full((n,), 0.05, dtype=torch.float32, device=DEV)
    dom = _domain(mn, mx, periodic)
    for step in range(2):
        print(f"case: crowded regular step {step}", flush=True)
        _check(pos, sup, dom, mn, mx, periodic, "csr")
    crowded = base.copy()
    crowded[:3000] = (np.float32(0.3) + rng.normal(0.0, 1e-4, (3000, 2))).astype(np.float32)
    pos.copy_(torch.from_numpy(crowded))
    for step in range(3):
        print(f"case: crowded cell step {step}", flush=True)
        _check(pos, sup, dom, mn, mx, periodic, "csr")
