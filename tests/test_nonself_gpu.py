"""GPU parity of searches whose query set is not the reference set, which
run as query cells (the queries sorted into the reference grid's cells, one
wavefront per cell, candidates from the reference structure): against the
oracle, bit-exact, for queries outside the grid and on the periodic bounds
(left to the per-query path one by one), crowded cells (more candidates than
one register chunk, more than 64 queries), grids the cell kernels do not take
(fewer than 5 cells in a periodic dimension: the per-query kernel takes the
whole cloud), reference points outside the grid (aliased cells, the
bucket-checked lookup) and a 200k x 150k search."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _check(q, hq, r, hr, mn, mx, per, mode, L=4096, fmt="coo"):
    dom = DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                            torch.tensor(per, dtype=torch.bool), len(mn))
    out = radiusSearch(PointCloud(_t(q), _t(hq)), PointCloud(_t(r), _t(hr)), mode=mode, domain=dom,
                       hashMapLength=L, algorithm="compact", format=fmt)
    torch.cuda.synchronize()
    o = oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                             hash_map_length=L)
    if fmt == "coo":
        np.testing.assert_array_equal(out.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(out.col.cpu().numpy(), o.col)
    else:
        np.testing.assert_array_equal(out.indices.cpu().numpy(), o.col)
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), o.indptr)
    return o


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mode", ["gather", "symmetric", "scatter", "superSymmetric"])
@pytest.mark.parametrize("L", [7, 4096])
def test_queries_outside_and_on_the_bounds(dim, mode, L):
    rng = np.random.default_rng(40 + dim * 10 + L % 10 + len(mode))
    per = [True] + [False] * (dim - 1)
    mn, mx = [-1.0] * dim, [1.0] * dim
    r = rng.uniform(-1, 1, (3000, dim)).astype(np.float32)
    q = rng.uniform(-1.15, 1.15, (2500, dim)).astype(np.float32)   # some beyond the non-periodic bounds
    q[:50, 0] = np.float32(-1.0) - np.float32(1e-8) * rng.integers(1, 4, 50).astype(np.float32)
    q[50:100, 0] = np.float32(1.0)
    q[100:150, -1] = np.float32(1.0)                              # on the max of a non-periodic dim
    q[150:200, -1] = np.nextafter(np.float32(-1.0), np.float32(-2.0))
    h = {2: 0.09, 3: 0.2}[dim]
    hq = rng.uniform(0.6 * h, h, q.shape[0]).astype(np.float32)
    hr = rng.uniform(0.6 * h, h, r.shape[0]).astype(np.float32)
    print(f"case: nonself bounds dim={dim} mode={mode} L={L}", flush=True)
    _check(q, hq, r, hr, mn, mx, per, mode, L)


@pytest.mark.parametrize("periodic", [False, True])
def test_crowded_cells(periodic):
    """Clustered references (cells with more than 448 candidates) and
    clustered queries (cells with more than 64 queries)."""
    rng = np.random.default_rng(91 + periodic)
    dim = 3
    per = [periodic] * dim
    mn, mx = [-1.0] * dim, [1.0] * dim
    r = np.concatenate([rng.normal(0.3, 0.03, (4000, dim)), rng.uniform(-1, 1, (3000, dim))]).astype(np.float32)
    q = np.concatenate([rng.normal(0.3, 0.02, (3000, dim)), rng.uniform(-1, 1, (2000, dim))]).astype(np.float32)
    r = np.clip(r, -0.999, 0.999).astype(np.float32)
    q = np.clip(q, -0.999, 0.999).astype(np.float32)
    hq = np.full(q.shape[0], 0.12, np.float32)
    hr = np.full(r.shape[0], 0.1, np.float32)
    print(f"case: nonself crowded periodic={periodic}", flush=True)
    for mode in ("gather", "symmetric"):
        _check(q, hq, r, hr, mn, mx, per, mode, fmt="csr")


@pytest.mark.parametrize("ncells", [3, 4, 5])
def test_few_periodic_cells(ncells):
    rng = np.random.default_rng(120 + ncells)
    dim = 2
    h = np.float32(2.0 / ncells * 0.999)
    r = rng.uniform(-1, 1, (800, dim)).astype(np.float32)
    q = rng.uniform(-1, 1, (700, dim)).astype(np.float32)
    print(f"case: nonself {ncells} periodic cells", flush=True)
    _check(q, np.full(700, h, np.float32), r, np.full(800, h, np.float32), [-1.0] * dim, [1.0] * dim,
           [True, True], "gather")


def test_reference_points_outside_the_grid():
    """Reference points beyond the non-periodic bounds: their cells' linear
    indices alias in-grid cells, found only through the bucket check."""
    rng = np.random.default_rng(130)
    r = rng.uniform(-1.3, 1.3, (4000, 2)).astype(np.float32)
    q = rng.uniform(-1.0, 1.0, (3000, 2)).astype(np.float32)
    print("case: nonself references outside the grid", flush=True)
    for L in (1, 97):
        _check(q, np.full(3000, 0.08, np.float32), r, np.full(4000, 0.08, np.float32), [-1.0, -1.0], [1.0, 1.0],
               [False, False], "gather", L)


def test_200k_by_150k():
    rng = np.random.default_rng(140)
    n = 58  # 58^3 ~ 195k jittered grid references
    dx = 2.0 / n
    g = (np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * dx - 1.0
    r = (g + rng.uniform(-0.1 * dx, 0.1 * dx, g.shape)).astype(np.float32)
    q = rng.uniform(-1, 1, (150000, 3)).astype(np.float32)
    h = np.float32((3 * 50 * dx ** 3 / (4 * np.pi)) ** (1 / 3))
    print("case: nonself 195k x 150k", flush=True)
    o = _check(q, np.full(q.shape[0], h, np.float32), r, np.full(r.shape[0], h, np.float32), [-1.0] * 3,
               [1.0] * 3, [True] * 3, "gather", L=1 << 20, fmt="csr")
    assert o.col.shape[0] > 5_000_000
