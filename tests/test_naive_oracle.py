"""Pins the brute-force restatement (oracle/naive.py) of the reference's
algorithm='naive' search (searchAlgorithms/radiusNaive.py):

  * golden vectors written by the reference function itself
    (tests/golden/make_naive_golden.py): asymmetric periodic domains, mixed
    periodicity, all modes, 1-3 dimensions;
  * the reference's own known answers for its lattices, which its test
    script asserts for 'naive' as for 'compact' (scripts/test.py:84-137);
  * the compact oracle (pinned separately): without periodic dimensions both
    compute the same fp32 distance, so the pair sets differ only by ties
    (distance == support: '<=' here, '<' there); naive's superSymmetric
    support is max(hx[a], hy[a]), the gather support of a self-search;
  * the periodic 'mod' on [0, 1] leaves |x| < 1 unchanged (its documented
    quirk: (|x| + 0) % 1 + 0 == |x|), and on [-a, a] it is the minimum image.
"""
import os

import numpy as np
import pytest

import oracle
from test_oracle import reference_test_data

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "naive_cases.npz")


def golden_cases():
    z = np.load(GOLDEN)
    n = int(z["n_cases"])
    keys = ("x", "y", "hx", "hy", "mn", "mx", "per", "mode", "ii", "jj")
    return [{k: z[f"c{i}_{k}"] for k in keys} for i in range(n)]


def test_golden_vectors():
    cases = golden_cases()
    assert len(cases) == 48
    for c in cases:
        ii, jj = oracle.radius_search_naive(c["x"], c["hx"], c["y"], c["hy"], domain_min=c["mn"],
                                            domain_max=c["mx"], periodic=c["per"].tolist(), mode=str(c["mode"]))
        np.testing.assert_array_equal(ii, c["ii"])
        np.testing.assert_array_equal(jj, c["jj"])


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [True, False])
def test_reference_known_answers(mode, periodic):
    Y, hY, P, hP, mn, mx = reference_test_data()
    ii, jj = oracle.radius_search_naive(P, hP, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2,
                                        mode=mode)
    counts = np.bincount(ii, minlength=P.shape[0])
    if periodic:
        assert ii.shape[0] == 46080 and counts.min() == counts.max() == 45
    else:
        assert ii.shape[0] == 41580 and counts.min() == 15 and counts.max() == 45
    if mode != "superSymmetric":   # needs as many queries as references
        ii, jj = oracle.radius_search_naive(Y, hY, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2,
                                            mode=mode)
        counts = np.bincount(ii, minlength=Y.shape[0])
        assert ii.shape[0] == 11520 and counts.min() == counts.max() == 45


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
def test_matches_compact_oracle_up_to_ties(dim, mode):
    rng = np.random.default_rng(500 + dim)
    n = 700
    q = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    r = q if mode == "superSymmetric" else rng.uniform(-1, 1, (n + 37, dim)).astype(np.float32)
    h = 0.25 if dim < 3 else 0.35
    hq = (h * rng.uniform(0.6, 1.4, q.shape[0])).astype(np.float32)
    hr = hq if r is q else (h * rng.uniform(0.6, 1.4, r.shape[0])).astype(np.float32)
    mn, mx = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    ii, jj = oracle.radius_search_naive(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=[False] * dim,
                                        mode=mode)
    naive = set(zip(ii.tolist(), jj.tolist()))
    # reference order: by reference index, then query index
    key = jj * (q.shape[0] + 1) + ii
    assert np.all(np.diff(key) > 0)
    # radiusNaive's superSymmetric support is torch.max(hx, hy) elementwise
    # (radiusNaive.py:29): max(hx[a], hy[a]) with the query index on both
    # sides, i.e. the gather support of a self-search
    cmode = "gather" if mode == "superSymmetric" else mode
    o = oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=[False] * dim, mode=cmode)
    compact = set(zip(o.row.tolist(), o.col.tolist()))
    assert compact <= naive
    for i, j in naive - compact:
        d = oracle.pair_distance(q[i], r[j], mn, mx, [False] * dim)
        thr = {"gather": hq[i], "scatter": hr[j], "symmetric": np.float32((hq[i] + hr[j]) / np.float32(2)),
               "superSymmetric": hq[i]}[mode]
        assert d == thr


def test_mod_quirk_closed_forms():
    x = np.linspace(-0.999, 0.999, 2001).astype(np.float32)
    np.testing.assert_array_equal(oracle.naive_mod(x, 0.0, 1.0), x)
    a = np.float32(1.0)
    y = np.linspace(-1.999, 1.999, 4001).astype(np.float32)
    want = np.where(np.abs(y) > a, y - np.sign(y) * np.float32(2) * a, y).astype(np.float32)
    np.testing.assert_array_equal(oracle.naive_mod(y, -1.0, 1.0), want)


# ---- algorithm='small' (cpp/neighborhoodSmall_cpu.cpp, commented out of the
# reference build): oracle.radius_search_small, the C brute force over the
# compact path's distance and predicate.

@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("periodic", [False, True])
def test_small_matches_compact_oracle(dim, mode, periodic):
    """Same distance and predicate as the compact oracle; inside the domain a
    once-wrapped reference point is its twice-wrapped self, so the pair sets
    agree exactly; the order is (query, reference)."""
    rng = np.random.default_rng(700 + dim * 5 + periodic)
    q = rng.uniform(-0.999, 0.999, (500, dim)).astype(np.float32)
    r = rng.uniform(-0.999, 0.999, (620, dim)).astype(np.float32)
    h = 0.25 if dim < 3 else 0.35
    hq = (h * rng.uniform(0.6, 1.4, 500)).astype(np.float32)
    hr = (h * rng.uniform(0.6, 1.4, 620)).astype(np.float32)
    mn, mx = np.full(dim, -1, np.float32), np.full(dim, 1, np.float32)
    per = [periodic] * dim
    ii, jj = oracle.radius_search_small(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
    key = ii * (r.shape[0] + 1) + jj
    assert np.all(np.diff(key) > 0)
    o = oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode)
    assert ii.shape[0] > 1000
    assert set(zip(ii.tolist(), jj.tolist())) == set(zip(o.row.tolist(), o.col.tolist()))
    np.testing.assert_array_equal(np.bincount(ii, minlength=q.shape[0]), o.counts)


@pytest.mark.parametrize("periodic", [True, False])
def test_small_reference_known_answers(periodic):
    """scripts/test.py:126-137 asserts its lattice counts for 'small' with
    per-point and with fixed supports."""
    Y, hY, P, hP, mn, mx = reference_test_data()
    for fixed in (None, float(hP[0])):
        ii, _ = oracle.radius_search_small(P, hP, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2,
                                           mode="gather", support_override=fixed)
        counts = np.bincount(ii, minlength=P.shape[0])
        if periodic:
            assert ii.shape[0] == 46080 and counts.min() == counts.max() == 45
        else:
            assert ii.shape[0] == 41580 and counts.min() == 15 and counts.max() == 45
        ii, _ = oracle.radius_search_small(Y, hY, P, hP, domain_min=mn, domain_max=mx, periodic=[periodic] * 2,
                                           mode="gather", support_override=fixed)
        assert ii.shape[0] == 11520
