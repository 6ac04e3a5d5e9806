"""Pins the CPU oracle (oracle/tcr_oracle.c) before it is trusted as the
parity checker of the HIP library.

Checks, per oracle function:
  * search counts  -- the reference's own known answers (scripts/test.py:84-119)
  * pair sets      -- O(N^2) brute force with the same distance, all modes
  * lattice counts -- closed forms (57 neighbours at h/dx = 2.2854 in 3-D)
  * row order      -- the ordering contract, recomputed from geometry
  * wrap           -- torch.remainder (the op the reference calls)
  * compute_h      -- torch fp32 ops (the ops the reference calls) and the
                      configuration values 0.2 / 10x10 and 0.046512 / 43^3
  * hash, linear index, offsets -- hand-computed vectors
"""
import math

import numpy as np
import pytest
import torch

import oracle


def reference_test_data(nx=32, dim=2, target=50):
    """generateNeighborTestData of scripts/test.py:31-68 (maxDomain_0 = 1)."""
    mn = torch.tensor([-1.0] * dim)
    mx = torch.tensor([1.0] * dim)
    ext = mx - mn
    dx = float(torch.min(ext).item()) / nx
    ny = int(1 // dx)
    h = math.sqrt(target * dx ** 2 / math.pi) if dim == 2 else (target * dx ** dim / math.pi * 3 / 4) ** (1 / 3)
    pos = [torch.linspace(float(mn[d]) + dx / 2, float(mx[d]) - dx / 2, int((ext[d] - dx) / dx) + 1) for d in range(dim)]
    P = torch.stack(torch.meshgrid(*pos, indexing="xy"), dim=-1).reshape(-1, dim)
    ys = [torch.linspace(-0.5 + dx / 2, 0.5 - dx / 2, ny) for _ in range(dim)]
    Y = torch.stack(torch.meshgrid(*ys, indexing="xy"), dim=-1).reshape(-1, dim)
    hP = torch.full((P.shape[0],), h, dtype=torch.float32)
    hY = torch.full((Y.shape[0],), float(hP[0]), dtype=torch.float32)
    return Y.numpy(), hY.numpy(), P.numpy(), hP.numpy(), mn.numpy(), mx.numpy()


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [True, False])
def test_reference_known_answers(mode, periodic):
    Y, hY, P, hP, mn, mx = reference_test_data()
    full = oracle.radius_search(P, hP, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2, mode=mode,
                                hash_map_length=4096)
    if periodic:  # scripts/test.py:89-94
        assert full.col.shape[0] == 46080
        assert full.counts.min() == full.counts.max() == 45
    else:         # scripts/test.py:96-106
        assert full.col.shape[0] == 41580
        assert full.counts.min() == 15 and full.counts.max() == 45
    red = oracle.radius_search(Y, hY, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2, mode=mode,
                               hash_map_length=4096)
    assert red.col.shape[0] == 11520                      # scripts/test.py:110-115
    assert red.counts.min() == red.counts.max() == 45


def _brute_pairs(q, hq, r, hr, mn, mx, per, mode):
    """Pair set by exhaustive search with the oracle's own distance."""
    qw = oracle.wrap(q, mn, mx, per)
    rw = oracle.wrap(oracle.wrap(r, mn, mx, per), mn, mx, per)
    out = set()
    for i in range(q.shape[0]):
        for j in range(r.shape[0]):
            d = oracle.pair_distance(qw[i], rw[j], mn, mx, per)
            hi, hj = hq[i], hr[j]
            ok = {"gather": d < hi, "scatter": d < hj, "symmetric": d < np.float32((hi + hj) / np.float32(2)),
                  "superSymmetric": d < (hj if hi < hj else hi)}[mode]
            if ok:
                out.add((i, j))
    return out


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [False, True])
def test_pair_sets_vs_brute_force(dim, mode, periodic):
    rng = np.random.default_rng(100 + dim * 10 + (mode == "scatter") * 3 + periodic)
    nq, nr = 90, 120
    q = rng.uniform(-1, 1, (nq, dim)).astype(np.float32)
    r = rng.uniform(-1, 1, (nr, dim)).astype(np.float32)
    hq = rng.uniform(0.1, 0.35, nq).astype(np.float32)
    hr = rng.uniform(0.1, 0.35, nr).astype(np.float32)
    mn, mx = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    per = [periodic] * dim
    res = oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                               hash_map_length=97)
    got = set(zip(res.row.tolist(), res.col.tolist()))
    assert len(got) == res.col.shape[0]  # no duplicates (>= 5 cells per periodic dim here)
    assert got == _brute_pairs(q, hq, r, hr, mn, mx, per, mode)
    bc = oracle.brute_counts(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
    np.testing.assert_array_equal(bc, res.counts)


def test_lattice_closed_form_3d():
    """Unjittered periodic lattice: every point has exactly the 57 lattice
    vectors with |v|^2 <= 5 (h/dx = 2.2854, h^2/dx^2 = 5.22)."""
    n = 24
    dx = 2.0 / n
    g = -1 + (np.arange(n) + 0.5) * dx
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    h = np.full(P.shape[0], 2.2854 * dx, np.float32)
    res = oracle.radius_search(P, h, domain_min=[-1] * 3, domain_max=[1] * 3, periodic=[True] * 3, mode="gather",
                               hash_map_length=1 << 14, want_rows=False)
    expect = sum(1 for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4) if a * a + b * b + c * c <= 5)
    assert expect == 57
    assert (res.counts == 57).all()


def test_row_order_contract():
    """Within a row: stencil offset order (dim 0 slowest), then ascending
    original reference index (stable sort inside a cell)."""
    rng = np.random.default_rng(7)
    P = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    h = np.full(3000, 0.2, np.float32)
    res = oracle.radius_search(P, h, domain_min=[-1] * 3, domain_max=[1] * 3, periodic=[True, False, True],
                               mode="gather", hash_map_length=4096, want_structure=True)
    hc = np.float32(res.hcell)
    W = oracle.wrap(oracle.wrap(P, [-1] * 3, [1] * 3, [True, False, True]), [-1] * 3, [1] * 3, [True, False, True])
    Q = oracle.wrap(P, [-1] * 3, [1] * 3, [True, False, True])
    cq = np.floor((Q - np.float32(-1)) / hc).astype(np.int64)
    cr = np.floor((W - np.float32(-1)) / hc).astype(np.int64)
    nc = res.num_cells.astype(np.int64)
    off = cr[res.col] - cq[res.row]
    off = np.where(off > 1, off - nc, np.where(off < -1, off + nc, off))  # periodic dims wrap
    assert np.abs(off).max() <= 1
    oidx = (off[:, 0] + 1) * 9 + (off[:, 1] + 1) * 3 + (off[:, 2] + 1)
    key = oidx * 10 ** 7 + res.col
    for i in range(0, 3000, 7):
        s, e = res.indptr[i], res.indptr[i + 1]
        assert np.all(np.diff(key[s:e]) > 0)


def test_wrap_matches_torch_remainder():
    rng = np.random.default_rng(3)
    for lo, hi in [(-1.0, 1.0), (0.1, 0.7), (-3.5, 12.25), (0.0, 1.0), (-1e-3, 2e-3)]:
        ext = np.float32(hi) - np.float32(lo)
        base = np.concatenate([rng.uniform(lo - 3 * ext, hi + 3 * ext, 200000),
                               [lo, hi, 0.0, (lo + hi) / 2, lo - ext, hi + ext]]).astype(np.float32)
        # +-3000 ulps around the boundaries (values that once-wrap onto max)
        near = []
        for b in np.float32([lo, hi, 0.0, lo - ext, hi + ext]):
            near.append(b + np.arange(-3000, 3001, dtype=np.float32) * np.spacing(b if b != 0 else np.float32(1e-30)))
        vals = np.concatenate([base] + near).astype(np.float32).reshape(-1, 1)
        got = oracle.wrap(vals, [lo], [hi], [True])[:, 0]
        t = torch.from_numpy(vals[:, 0])
        mn, mx = torch.tensor(lo, dtype=torch.float32), torch.tensor(hi, dtype=torch.float32)
        want = (torch.remainder(t - mn, mx - mn) + mn).numpy()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def _compute_h_torch(mn, mx, hmax):
    """The op sequence of compute_h (py/util.py:40-79) run through torch fp32."""
    qMin, qMax, s = torch.tensor(mn), torch.tensor(mx), torch.tensor(hmax)
    qExtent = qMax - qMin
    numCells = torch.floor(qExtent / s)
    h = qExtent / numCells
    if torch.any(qExtent / h - numCells > 0):
        numCells -= 1
        h = qExtent / numCells
    if torch.any(torch.ceil(qExtent / h) > qExtent / h):
        h = h * (1e-4 + 1)
    return torch.max(h).item()


def test_compute_h_matches_torch_ops():
    rng = np.random.default_rng(11)
    for _ in range(3000):
        dim = int(rng.integers(1, 4))
        mn = rng.uniform(-5, 1, dim).astype(np.float32)
        mx = (mn + rng.uniform(0.5, 7, dim)).astype(np.float32)
        hmax = np.float32(rng.uniform(0.01, 0.4) * (mx - mn).min())
        a = float(oracle.compute_h(mn, mx, hmax))
        b = _compute_h_torch(mn, mx, hmax)
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32), (mn, mx, hmax, a, b)


def test_compute_h_workload_values():
    # 2-D: h = sqrt(32 dx^2 / pi), dx = 1/16 -> hCell 0.2 (10 x 10 cells)
    h1 = np.float32(math.sqrt(32 * (2 / 32) ** 2 / math.pi))
    hc = oracle.compute_h([-1, -1], [1, 1], h1)
    assert abs(hc - 0.2) < 1e-6
    assert math.ceil(2 / hc) == 10
    # 3-D periodic 1M: h = cbrt(150 dx^3 / (4 pi)), dx = 0.02 -> 0.046512, 43^3
    h2 = np.float32((3 * 50 * 0.02 ** 3 / (4 * math.pi)) ** (1 / 3))
    hc2 = oracle.compute_h([-1] * 3, [1] * 3, h2)
    assert abs(hc2 - 2 / 43) < 1e-6 and hc2 >= h2
    assert int(np.ceil(np.float32(2) / np.float32(hc2))) == 43


def test_hash_and_linear_index_vectors():
    P = [73856093, 19349663, 83492791]
    L = 4096
    assert oracle.hash_index([5], L) == 5
    assert oracle.hash_index([-1], L) == (2 ** 32 - 1) % L
    assert oracle.hash_index([3, 7], L) == (3 * P[0] + 7 * P[1]) % 2 ** 32 % L
    assert oracle.hash_index([-1, 2, 9], 1 << 20) == ((2 ** 32 - 1) * P[0] + 2 * P[1] + 9 * P[2]) % 2 ** 32 % (1 << 20)
    assert oracle.hash_index([100, 200, 300], 97) == ((100 * P[0] + 200 * P[1] + 300 * P[2]) % 2 ** 32) % 97
    assert oracle.linear_index([2, 3, 4], [10, 20, 30]) == 2 + 3 * 10 + 4 * 200
    assert oracle.linear_index([-1, 3], [10, 20]) == -1 + 30
    # int32 wrap-around
    big = oracle.linear_index([1, 1, 70000], [70000, 70000, 70000])
    assert big == np.int32(np.uint32((1 + 70000 + 70000 * 70000 * 70000) % 2 ** 32).astype(np.int64) - (2 ** 32 if (1 + 70000 + 70000 ** 3) % 2 ** 32 >= 2 ** 31 else 0))


def test_offsets_order():
    o = oracle.offsets(1, 2).tolist()
    assert o == [[-1, -1], [-1, 0], [-1, 1], [0, -1], [0, 0], [0, 1], [1, -1], [1, 0], [1, 1]]
    o3 = oracle.offsets(1, 3)
    assert o3.shape == (27, 3) and o3[0].tolist() == [-1, -1, -1] and o3[1].tolist() == [-1, -1, 0]
    assert o3[13].tolist() == [0, 0, 0] and o3[26].tolist() == [1, 1, 1]
    assert oracle.offsets(2, 1)[:, 0].tolist() == [-2, -1, 0, 1, 2]


def test_empty_inputs():
    P = np.zeros((0, 3), np.float32)
    h = np.zeros(0, np.float32)
    R = np.random.default_rng(1).uniform(-1, 1, (50, 3)).astype(np.float32)
    res = oracle.radius_search(P, h, R, np.full(50, 0.3, np.float32), domain_min=[-1] * 3, domain_max=[1] * 3,
                               periodic=[False] * 3, mode="gather", hash_map_length=64)
    assert res.col.shape[0] == 0 and res.counts.shape[0] == 0
