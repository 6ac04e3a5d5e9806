"""GPU parity: the HIP search (through libtcr_hip.so) against the CPU oracle,
bit-exact on indices (pair sets and order) for every mode, dimension and
periodicity, the edge cases of the reference path, and the workloads."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch
from torchcompactradius_amd import synthetic

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def gpu_search(q, hq, r=None, hr=None, *, mn=None, mx=None, periodic=None, mode="gather", L=4096, fmt="coo",
               structure=False):
    dom = None
    if mn is not None:
        dom = DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                                torch.tensor(periodic, dtype=torch.bool), len(mn))
    ref = None if r is None else PointCloud(_t(r), _t(hr))
    out = radiusSearch(PointCloud(_t(q), _t(hq)), ref, mode=mode, domain=dom, hashMapLength=L,
                       algorithm="compact", format=fmt, returnStructure=structure)
    torch.cuda.synchronize()
    return out


def assert_same(res_gpu, res_orc, fmt="coo"):
    if fmt == "coo":
        row = res_gpu.row.cpu().numpy()
        col = res_gpu.col.cpu().numpy()
        assert res_gpu.row.dtype == torch.int64 and res_gpu.col.dtype == torch.int64
        np.testing.assert_array_equal(row, res_orc.row)
        np.testing.assert_array_equal(col, res_orc.col)
    else:
        np.testing.assert_array_equal(res_gpu.indices.cpu().numpy(), res_orc.col)
        assert res_gpu.indptr.dtype == torch.int32 and res_gpu.rowEntries.dtype == torch.int64
        np.testing.assert_array_equal(res_gpu.indptr.cpu().numpy(), res_orc.indptr)
        np.testing.assert_array_equal(res_gpu.rowEntries.cpu().numpy(), res_orc.counts.astype(np.int64))


def test_plain_2d_1024_workload():
    w = synthetic.plain_2d_1024()
    print("case: plain-2d-1024", flush=True)
    g = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, mode=w.mode,
                   L=w.hash_map_length, fmt="coo")
    o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode=w.mode, hash_map_length=w.hash_map_length)
    assert_same(g, o)
    assert 30000 < o.col.shape[0] < 36000


@pytest.mark.parametrize("n", [20, 40])
def test_periodic_3d_reduced(n):
    w = synthetic.periodic_3d(n=n)
    print(f"case: periodic-3d n={n}", flush=True)
    g = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, mode=w.mode,
                   L=w.hash_map_length, fmt="csr")
    o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode=w.mode, hash_map_length=w.hash_map_length, want_rows=False)
    assert_same(g, o, "csr")


@pytest.mark.slow
def test_periodic_3d_1m_full():
    w = synthetic.periodic_3d(n=100)
    print("case: periodic-3d-1M", flush=True)
    g = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, mode=w.mode,
                   L=w.hash_map_length, fmt="csr")
    o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode=w.mode, hash_map_length=w.hash_map_length, want_rows=False)
    assert_same(g, o, "csr")


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("pmask", [0, 1, 7])
def test_modes_dims_periodicity(dim, mode, pmask):
    rng = np.random.default_rng(1000 + dim * 100 + pmask * 7 + len(mode))
    nq, nr = 700, 900
    per = [bool((pmask >> d) & 1) for d in range(dim)]
    q = rng.uniform(-1, 1, (nq, dim)).astype(np.float32)
    r = rng.uniform(-1, 1, (nr, dim)).astype(np.float32)
    hq = rng.uniform(0.05, 0.2, nq).astype(np.float32)
    hr = rng.uniform(0.05, 0.2, nr).astype(np.float32)
    mn, mx = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    print(f"case: dim={dim} mode={mode} pmask={pmask}", flush=True)
    for selfs in (False, True):
        g = gpu_search(q, hq, None if selfs else r, None if selfs else hr, mn=mn, mx=mx, periodic=per, mode=mode,
                       L=331)
        o = oracle.radius_search(q, hq, None if selfs else r, None if selfs else hr, domain_min=mn, domain_max=mx,
                                 periodic=per, mode=mode, hash_map_length=331)
        assert_same(g, o)


def test_domain_none_and_points_outside():
    """domain=None means [0,1]^d non-periodic; points outside the domain get
    unclamped cells whose linear indices can alias in-domain cells."""
    rng = np.random.default_rng(5)
    q = rng.uniform(-0.7, 1.8, (1500, 2)).astype(np.float32)
    h = np.full(1500, 0.09, np.float32)
    print("case: domain=None", flush=True)
    for L in (1, 7, 4096):
        g = gpu_search(q, h, L=L)
        o = oracle.radius_search(q, h, hash_map_length=L)
        assert_same(g, o)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("mode", ["gather", "symmetric"])
def test_non_finite_positions(dim, periodic, mode):
    """NaN and +-inf coordinates (no pairs: every distance to them is NaN,
    their cells are INT32_MIN and alias through the hash) and a huge finite
    one (out of the grid, or wrapped back into a periodic domain), in a
    self-search and with a separate query set."""
    rng = np.random.default_rng(30 + dim)
    n = 2000
    q = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    q[5, 0] = np.nan
    q[17, dim - 1] = np.inf
    q[40, 0] = -np.inf
    q[41, :] = np.nan
    q[99, 1] = np.float32(1e30)
    h = np.full(n, 0.12, np.float32)
    mn, mx = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    per = [periodic] * dim
    print(f"case: non-finite dim={dim} periodic={periodic} mode={mode}", flush=True)
    for fmt in ("coo", "csr"):
        g = gpu_search(q, h, mn=mn, mx=mx, periodic=per, mode=mode, fmt=fmt)
        o = oracle.radius_search(q, h, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=4096)
        assert_same(g, o, fmt)
    r = rng.uniform(-1, 1, (1500, dim)).astype(np.float32)
    r[3, 0] = np.nan
    r[8, dim - 1] = -np.inf
    hr = np.full(1500, 0.12, np.float32)
    g = gpu_search(q, h, r, hr, mn=mn, mx=mx, periodic=per, mode=mode)
    o = oracle.radius_search(q, h, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                             hash_map_length=4096)
    assert_same(g, o)


def test_wrap_boundary_values():
    """Values that once-wrap onto max and twice-wrap onto min (the reference
    wraps the reference set twice and the query set once)."""
    lo, hi = np.float32(0.1), np.float32(0.7)
    vals = []
    for b in (lo, hi, lo - (hi - lo), hi + (hi - lo)):
        vals.append(b + np.arange(-40, 41, dtype=np.float32) * np.spacing(b))
    v = np.concatenate(vals).astype(np.float32)
    rng = np.random.default_rng(9)
    q = np.stack([v, rng.uniform(0.1, 0.7, v.shape[0]).astype(np.float32)], 1)
    h = np.full(q.shape[0], 0.05, np.float32)
    mn, mx = np.array([lo, lo], np.float32), np.array([hi, hi], np.float32)
    print("case: wrap boundaries", flush=True)
    g = gpu_search(q, h, mn=mn, mx=mx, periodic=[True, True], L=64)
    o = oracle.radius_search(q, h, domain_min=mn, domain_max=mx, periodic=[True, True], hash_map_length=64)
    assert_same(g, o)


def test_small_periodic_grids_visit_cells_twice():
    """Two or three cells in a periodic dim: the stencil reaches the same
    cell more than once and the reference emits duplicate pairs."""
    rng = np.random.default_rng(12)
    q = rng.uniform(0, 1, (300, 2)).astype(np.float32)
    h = np.full(300, 0.4, np.float32)
    print("case: small periodic grid", flush=True)
    g = gpu_search(q, h, mn=[0, 0], mx=[1, 1], periodic=[True, False], L=16)
    o = oracle.radius_search(q, h, domain_min=[0, 0], domain_max=[1, 1], periodic=[True, False], hash_map_length=16)
    assert_same(g, o)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ncells", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("mode", ["gather", "symmetric", "superSymmetric"])
def test_few_cells_in_periodic_dims(dim, ncells, mode):
    """Periodic dimensions only a few cells wide: candidates reach the query
    through the image across the boundary, and the cell-centred estimate has
    to pick the reference's minimal image."""
    rng = np.random.default_rng(40 + dim * 10 + ncells + len(mode))
    n = 2500 if dim == 3 else 1200
    q = rng.uniform(0, 1, (n, dim)).astype(np.float32)
    q[:60, 0] = rng.uniform(0, 0.02, 60).astype(np.float32)      # hugging both faces
    q[60:120, 0] = rng.uniform(0.98, 1.0, 60).astype(np.float32)
    hmax = 0.999 / ncells
    h = (hmax * rng.uniform(0.6, 1.0, n)).astype(np.float32)
    h[0] = np.float32(hmax)
    per = [True, False, True][:dim] if dim == 3 else [True, True]
    print(f"case: few periodic cells dim={dim} ncells={ncells} mode={mode}", flush=True)
    g = gpu_search(q, h, mn=[0] * dim, mx=[1] * dim, periodic=per, mode=mode, L=1024)
    o = oracle.radius_search(q, h, domain_min=[0] * dim, domain_max=[1] * dim, periodic=per, mode=mode,
                             hash_map_length=1024)
    assert_same(g, o)


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [False, True])
def test_clustered_big_cells_and_member_batches(mode, periodic):
    """Dense clusters: cells whose stencils hold more candidates than one
    register chunk (the chunked big-cell kernel), a cell of more than one
    wavefront of points with empty neighbours (several member batches in the
    small-cell kernels), and sparse background points."""
    rng = np.random.default_rng(21 + len(mode) + (7 if periodic else 0))
    bg = rng.uniform(-1, 1, (2000, 3))
    ball = rng.normal(0.3, 0.03, (1500, 3))
    lone = np.array([-0.75, -0.75, 0.75]) + rng.uniform(-0.03, 0.03, (100, 3))
    q = np.concatenate([bg, ball, lone]).astype(np.float32)
    q = np.clip(q, -0.999, 0.999)
    h = rng.uniform(0.08, 0.12, q.shape[0]).astype(np.float32)
    per = [periodic] * 3
    print(f"case: clustered mode={mode} periodic={periodic}", flush=True)
    g = gpu_search(q, h, mn=[-1] * 3, mx=[1] * 3, periodic=per, mode=mode, L=4096)
    o = oracle.radius_search(q, h, domain_min=[-1] * 3, domain_max=[1] * 3, periodic=per, mode=mode,
                             hash_map_length=4096)
    assert_same(g, o)
    assert o.counts.max() > 448  # the ball's rows exceed one register chunk


def test_empty_and_tiny():
    rng = np.random.default_rng(4)
    r = rng.uniform(-1, 1, (64, 3)).astype(np.float32)
    hr = np.full(64, 0.3, np.float32)
    print("case: empty query", flush=True)
    g = gpu_search(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), r, hr, mn=[-1] * 3, mx=[1] * 3,
                   periodic=[False] * 3)
    assert g.row.numel() == 0
    print("case: single point", flush=True)
    g = gpu_search(r[:1], hr[:1], mn=[-1] * 3, mx=[1] * 3, periodic=[True] * 3)
    assert g.row.tolist() == [0] and g.col.tolist() == [0]


def test_structure_matches_oracle():
    w = synthetic.periodic_3d(n=24)
    print("case: returnStructure", flush=True)
    g, hm = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, mode=w.mode,
                       L=4096, fmt="csr", structure=True)
    o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode=w.mode, hash_map_length=4096, want_rows=False, want_structure=True)
    np.testing.assert_array_equal(hm.sortIndex.cpu().numpy(), o.sort_index)
    np.testing.assert_array_equal(hm.numCells.cpu().numpy(), o.num_cells)
    assert np.float32(hm.hCell.item()) == np.float32(o.hcell)
    ht = hm.hashTable.cpu().numpy()
    np.testing.assert_array_equal(ht[:, 1], o.hash_table[:, 1])
    np.testing.assert_array_equal(ht[:, 0] == -1, o.hash_table[:, 0] == -1)
    ct = hm.sortedCellTable.cpu().numpy()
    assert ct.shape == o.cell_table.shape
    # same cells per bucket (order inside a bucket is free)
    for b in np.nonzero(o.hash_table[:, 1])[0][:500]:
        s0, n0 = o.hash_table[b]
        s1, n1 = ht[b]
        assert sorted(map(tuple, o.cell_table[s0:s0 + n0])) == sorted(map(tuple, ct[s1:s1 + n1]))


def test_deterministic_repeat():
    w = synthetic.periodic_3d(n=30)
    print("case: determinism", flush=True)
    a = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, L=1 << 16, fmt="csr")
    b = gpu_search(w.pos, w.sup, mn=w.domain_min, mx=w.domain_max, periodic=w.periodic, L=1 << 16, fmt="csr")
    assert torch.equal(a.indices, b.indices) and torch.equal(a.indptr, b.indptr)


def _strays_in_dense_cells(periodic):
    """A dense ball (cells with far more than 448 stencil candidates) and 60
    points just below domain_min in x that land among the ball's cell
    members: in a non-periodic domain their cell -1 aliases the ball's cell
    (nx-1, cy-1, cz) through the linear index; in a periodic one they wrap
    onto max once (the query copy) and onto min twice (the sorted copy), so
    they sort into the ball's cell (0, cy, cz).  The domain [0.1, 0.7]^3 is
    one where 0.1 - 1 ulp once-wraps onto exactly 0.7."""
    rng = np.random.default_rng(77 + periodic)
    bg = rng.uniform(-1, 1, (1500, 3))
    c = np.array([-0.95, 0.1, 0.1]) if periodic else np.array([0.93, 0.1, 0.1])
    ball = np.clip(c + rng.normal(0, 0.03, (2500, 3)), -0.999, 0.999)
    k = 60
    st = np.empty((k, 3))
    st[:, 1:] = c[1:] + rng.uniform(-0.03, 0.03, (k, 2))
    if not periodic:
        st[:, 1] += 0.125
    near = st.copy()
    near[:, 0] = -0.97  # in-domain neighbours, so that the strays' rows are not empty
    pts = np.concatenate([bg, ball, near, st])
    q = (np.float32(0.1) + (pts + 1) * 0.3).astype(np.float32)
    below = np.nextafter(np.float32(0.1), np.float32(0))
    q[-k:, 0] = below - (0 if periodic else np.arange(k, dtype=np.float32) * np.float32(3e-8))
    h = (0.3 * rng.uniform(0.08, 0.12, q.shape[0])).astype(np.float32)
    return q, h, k


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("mode", ["gather", "superSymmetric"])
def test_strays_among_big_cell_members(periodic, mode):
    """Stray members of cells too big for the small-cell kernels are searched
    by the per-query kernel (their rows were once left unwritten)."""
    q, h, k = _strays_in_dense_cells(periodic)
    mn, mx, per = [0.1] * 3, [0.7] * 3, [periodic] * 3
    print(f"case: strays among big-cell members periodic={periodic} mode={mode}", flush=True)
    o = oracle.radius_search(q, h, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=4096)
    assert o.counts.max() > 448 and o.counts[-k:].min() > 0
    for fmt in ("coo", "csr"):
        g = gpu_search(q, h, mn=mn, mx=mx, periodic=per, mode=mode, L=4096, fmt=fmt)
        assert_same(g, o if fmt == "coo" else oracle.radius_search(
            q, h, domain_min=mn, domain_max=mx, periodic=per, mode=mode, hash_map_length=4096, want_rows=False),
            fmt)


@pytest.mark.parametrize("case", [(2, (False, False), (-1.0, -1.0), (1.0, 1.0)),
                                  (2, (True, True), (0.1, 0.1), (0.7, 0.7)),
                                  (3, (False, False, False), (-1.0,) * 3, (1.0,) * 3),
                                  (3, (True, False, True), (0.1, -1.0, 0.1), (0.7, 1.0, 0.7))])
@pytest.mark.parametrize("mode", ["gather", "scatter"])
def test_supports_one_ulp_from_the_distance(case, mode):
    """Self-search over islands that each hold one pair (xi, xj) at exact
    distance d (the exact-rational model of modDistance) and three copies of
    the point whose support decides it, with supports one fp32 ulp below d,
    d and one ulp above: the compact path's filtered estimate must hand every
    such pair to the reference formula and decide it as the model does, in
    periodic and non-periodic dimensions."""
    import exact_arith as ex
    dim, per, mn, mx = case
    mn32, mx32 = np.asarray(mn, np.float32), np.asarray(mx, np.float32)
    rng = np.random.default_rng(7100 + dim + 10 * sum(per) + len(mode))
    n_side = 20 if dim == 2 else 8
    xi, xj, ds = ex.island_pairs(rng, dim, list(per), mn32, mx32, n_side,
                                 lambda p: oracle.wrap(p, mn32, mx32, list(per)))
    n = xi.shape[0]
    sup3 = np.array([ex.supports_around(d) for d in ds], np.float32)          # [n, 3]
    small = np.array([np.float32(float(d)) * np.float32(0.25) for d in ds], np.float32)
    # points: 3 copies of the deciding point, then its partner
    decider, partner = (xi, xj) if mode == "gather" else (xj, xi)
    pos = np.concatenate([np.repeat(decider, 3, axis=0), partner]).astype(np.float32)
    sup = np.concatenate([sup3.reshape(-1), small]).astype(np.float32)
    print(f"case: one-ulp supports dim={dim} periodic={per} mode={mode} pairs={n}", flush=True)
    g = gpu_search(pos, sup, mn=mn32, mx=mx32, periodic=list(per), mode=mode, L=8192)
    o = oracle.radius_search(pos, sup, domain_min=mn32, domain_max=mx32, periodic=list(per), mode=mode,
                             hash_map_length=8192)
    assert_same(g, o)
    row, col = g.row.cpu().numpy(), g.col.cpu().numpy()
    pairs = set(zip(row.tolist(), col.tolist()))
    got = np.zeros((n, 3), bool)
    for t in range(n):
        for k in range(3):
            q, r = (3 * t + k, 3 * n + t)
            got[t, k] = ((q, r) in pairs) if mode == "gather" else ((r, q) in pairs)
    want = np.array([[ex.inside(ds[t], sup3[t, k]) for k in range(3)] for t in range(n)])
    np.testing.assert_array_equal(got, want)
    assert want[:, 2].all() and not want[:, :2].any()
