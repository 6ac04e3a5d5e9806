"""GPU parity of radiusSearch(algorithm='naive') -- the reference's default,
brute-force algorithm (searchAlgorithms/radiusNaive.py) -- through
libtcr_hip.so's tcr_naive_count / tcr_naive_fill, against the reference's own
vectors (tests/golden/naive_cases.npz), its lattice known answers and the
numpy oracle (oracle/naive.py); bit-exact on indices and order."""
import numpy as np
import pytest
import torch

import oracle
from test_naive_oracle import golden_cases
from test_oracle import reference_test_data
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _dom(mn, mx, per):
    return DomainDescription(torch.tensor(mn), torch.tensor(mx), torch.tensor(list(per)), len(mn))


def test_golden_vectors():
    cases = golden_cases()
    print(f"case: naive golden vectors ({len(cases)})", flush=True)
    for c in cases:
        out = radiusSearch(PointCloud(_t(c["x"]), _t(c["hx"])), PointCloud(_t(c["y"]), _t(c["hy"])),
                           mode=str(c["mode"]), domain=_dom(c["mn"], c["mx"], c["per"].tolist()), algorithm="naive")
        np.testing.assert_array_equal(out.row.cpu().numpy(), c["ii"])
        np.testing.assert_array_equal(out.col.cpu().numpy(), c["jj"])


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [True, False])
def test_reference_known_answers(mode, periodic):
    Y, hY, P, hP, mn, mx = reference_test_data()
    dom = _dom(mn, mx, [periodic] * 2)
    print(f"case: naive lattice mode={mode} periodic={periodic}", flush=True)
    full = radiusSearch(PointCloud(_t(P), _t(hP)), PointCloud(_t(P), _t(hP)), mode=mode, domain=dom,
                        algorithm="naive")
    counts = torch.bincount(full.row, minlength=P.shape[0]).cpu().numpy()
    if periodic:
        assert full.row.shape[0] == 46080 and counts.min() == counts.max() == 45
    else:
        assert full.row.shape[0] == 41580 and counts.min() == 15 and counts.max() == 45
    if mode != "superSymmetric":
        red = radiusSearch(PointCloud(_t(Y), _t(hY)), PointCloud(_t(P), _t(hP)), mode=mode, domain=dom,
                           algorithm="naive")
        counts = torch.bincount(red.row, minlength=Y.shape[0]).cpu().numpy()
        assert red.row.shape[0] == 11520 and counts.min() == counts.max() == 45


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("fmt", ["coo", "csr"])
def test_vs_oracle(dim, mode, fmt):
    rng = np.random.default_rng(900 + dim * 7 + len(mode) + (fmt == "csr"))
    mn = np.array([-0.4, 0.2, -3.0][:dim], np.float32)
    mx = np.array([1.3, 0.9, -1.5][:dim], np.float32)
    ext = mx - mn
    nx = 1500
    ny = nx if mode == "superSymmetric" else 1811
    x = (mn + rng.uniform(-0.3, 1.3, (nx, dim)) * ext).astype(np.float32)
    y = (mn + rng.uniform(-0.3, 1.3, (ny, dim)) * ext).astype(np.float32)
    h = float(ext.min()) * (0.2 if dim == 3 else 0.1)
    hx = (h * rng.uniform(0.5, 1.5, nx)).astype(np.float32)
    hy = (h * rng.uniform(0.5, 1.5, ny)).astype(np.float32)
    per = [bool(v) for v in rng.random(dim) < 0.6]
    print(f"case: naive vs oracle dim={dim} mode={mode} fmt={fmt} per={per}", flush=True)
    out = radiusSearch(PointCloud(_t(x), _t(hx)), PointCloud(_t(y), _t(hy)), mode=mode, domain=_dom(mn, mx, per),
                       algorithm="naive", format=fmt)
    ii, jj = oracle.radius_search_naive(x, hx, y, hy, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
    assert ii.shape[0] > 1000
    if fmt == "coo":
        np.testing.assert_array_equal(out.row.cpu().numpy(), ii)
        np.testing.assert_array_equal(out.col.cpu().numpy(), jj)
    else:
        # coo_to_csr(isSorted=True) of the reference-ordered pairs
        nj = np.bincount(ii, minlength=nx)
        np.testing.assert_array_equal(out.indices.cpu().numpy(), jj)
        np.testing.assert_array_equal(out.rowEntries.cpu().numpy(), nj)
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), np.concatenate([[0], np.cumsum(nj)]).astype(np.int32))
        assert out.indptr.dtype == torch.int32 and out.rowEntries.dtype == torch.int64


def test_self_search_default_domain_int32_and_fixed_support():
    rng = np.random.default_rng(77)
    x = rng.uniform(-0.1, 1.1, (2500, 2)).astype(np.float32)
    hx = rng.uniform(0.02, 0.05, 2500).astype(np.float32)
    print("case: naive self default domain", flush=True)
    out = radiusSearch(PointCloud(_t(x), _t(hx)), None, mode="symmetric", algorithm="naive")
    ii, jj = oracle.radius_search_naive(x, hx, None, None, mode="symmetric")
    np.testing.assert_array_equal(out.row.cpu().numpy(), ii)
    np.testing.assert_array_equal(out.col.cpu().numpy(), jj)
    out32 = radiusSearch(PointCloud(_t(x), _t(hx)), None, mode="symmetric", algorithm="naive",
                         indexDtype=torch.int32)
    assert out32.row.dtype == torch.int32
    np.testing.assert_array_equal(out32.col.cpu().numpy(), jj)
    print("case: naive fixed support", flush=True)
    dom = _dom(np.zeros(2, np.float32), np.ones(2, np.float32), [True, False])
    fx = radiusSearch(PointCloud(_t(x), None), PointCloud(_t(x[::2].copy()), None), supportOverride=0.04,
                      domain=dom, algorithm="naive")
    fi, fj = oracle.radius_search_naive(x, None, x[::2], None, domain_min=np.zeros(2), domain_max=np.ones(2),
                                        periodic=[True, False], support_override=0.04)
    assert fi.shape[0] > 1000
    np.testing.assert_array_equal(fx.row.cpu().numpy(), fi)
    np.testing.assert_array_equal(fx.col.cpu().numpy(), fj)


def test_empty_and_errors():
    x = torch.rand((50, 3), device=DEV)
    h = torch.full((50,), 0.2, device=DEV)
    e = torch.zeros((0, 3), device=DEV)
    eh = torch.zeros(0, device=DEV)
    print("case: naive empty / errors", flush=True)
    for q, r in (((e, eh), (x, h)), ((x, h), (e, eh))):
        out = radiusSearch(PointCloud(*q), PointCloud(*r), algorithm="naive")
        assert out.row.numel() == 0 and out.col.numel() == 0
    with pytest.raises(ValueError):   # superSymmetric: torch.max(hx, hy) needs equal lengths
        radiusSearch(PointCloud(x, h), PointCloud(x[:20].contiguous(), h[:20].contiguous()), mode="superSymmetric",
                     algorithm="naive")
    with pytest.raises(ValueError):   # gather without query supports
        radiusSearch(PointCloud(x, None), PointCloud(x, h), mode="gather", algorithm="naive")


# ---- algorithm='small': oracle.radius_search_small (C brute force, pinned
# against the compact oracle and the lattice known answers)

@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("fmt", ["coo", "csr"])
def test_small_vs_oracle(dim, mode, fmt):
    rng = np.random.default_rng(1300 + dim * 7 + len(mode) + (fmt == "csr"))
    mn = np.array([-0.4, 0.2, -3.0][:dim], np.float32)
    mx = np.array([1.3, 0.9, -1.5][:dim], np.float32)
    ext = mx - mn
    nx, ny = 1400, 1733
    x = (mn + rng.uniform(-0.3, 1.3, (nx, dim)) * ext).astype(np.float32)
    y = (mn + rng.uniform(-0.3, 1.3, (ny, dim)) * ext).astype(np.float32)
    h = float(ext.min()) * (0.2 if dim == 3 else 0.1)
    hx = (h * rng.uniform(0.5, 1.5, nx)).astype(np.float32)
    hy = (h * rng.uniform(0.5, 1.5, ny)).astype(np.float32)
    per = [bool(v) for v in rng.random(dim) < 0.6]
    print(f"case: small vs oracle dim={dim} mode={mode} fmt={fmt} per={per}", flush=True)
    out = radiusSearch(PointCloud(_t(x), _t(hx)), PointCloud(_t(y), _t(hy)), mode=mode, domain=_dom(mn, mx, per),
                       algorithm="small", format=fmt)
    ii, jj = oracle.radius_search_small(x, hx, y, hy, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
    assert ii.shape[0] > 1000
    if fmt == "coo":
        np.testing.assert_array_equal(out.row.cpu().numpy(), ii)
        np.testing.assert_array_equal(out.col.cpu().numpy(), jj)
    else:
        nj = np.bincount(ii, minlength=nx)
        np.testing.assert_array_equal(out.indices.cpu().numpy(), jj)
        np.testing.assert_array_equal(out.rowEntries.cpu().numpy(), nj)
        np.testing.assert_array_equal(out.indptr.cpu().numpy(), np.concatenate([[0], np.cumsum(nj)]).astype(np.int32))


@pytest.mark.parametrize("periodic", [True, False])
def test_small_known_answers_and_fixed(periodic):
    Y, hY, P, hP, mn, mx = reference_test_data()
    dom = _dom(mn, mx, [periodic] * 2)
    print(f"case: small lattice periodic={periodic}", flush=True)
    for override in (None, float(hP[0])):
        sup = None if override is not None else _t(hP)
        full = radiusSearch(PointCloud(_t(P), sup), None, supportOverride=override, domain=dom, algorithm="small")
        counts = torch.bincount(full.row, minlength=P.shape[0]).cpu().numpy()
        if periodic:
            assert full.row.shape[0] == 46080 and counts.min() == counts.max() == 45
        else:
            assert full.row.shape[0] == 41580 and counts.min() == 15 and counts.max() == 45
        ii, jj = oracle.radius_search_small(P, hP, None, None, domain_min=mn, domain_max=mx,
                                            periodic=[periodic] * 2, mode="gather", support_override=override)
        np.testing.assert_array_equal(full.row.cpu().numpy(), ii)
        np.testing.assert_array_equal(full.col.cpu().numpy(), jj)


@pytest.mark.parametrize("algorithm", ["naive", "small"])
def test_non_finite_inputs(algorithm):
    """NaN / inf positions and supports, zero and negative supports: every
    comparison with NaN fails, inf distances exceed every finite support, and
    the kernels agree with the restatements."""
    rng = np.random.default_rng(55)
    n = 600
    x = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    h = rng.uniform(0.05, 0.15, n).astype(np.float32)
    x[3, 0] = np.nan
    x[7, 1] = np.inf
    x[11, 2] = -np.inf
    h[5] = np.nan
    h[9] = np.inf
    h[13] = 0.0
    h[17] = -0.1
    mn, mx = np.zeros(3, np.float32), np.ones(3, np.float32)
    per = [True, False, True]
    print(f"case: {algorithm} non-finite inputs", flush=True)
    for mode in ("gather", "scatter", "symmetric", "superSymmetric"):
        out = radiusSearch(PointCloud(_t(x), _t(h)), None, mode=mode, domain=_dom(mn, mx, per), algorithm=algorithm)
        if algorithm == "naive":
            ii, jj = oracle.radius_search_naive(x, h, None, None, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
        else:
            ii, jj = oracle.radius_search_small(x, h, None, None, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
        np.testing.assert_array_equal(out.row.cpu().numpy(), ii)
        np.testing.assert_array_equal(out.col.cpu().numpy(), jj)
