"""Pins the oracle's pair-distance rounding with exact arithmetic.

For 1,200+ (query, reference, support) triples in 2-D and 3-D, periodic and
non-periodic, on a domain symmetric about 0 and on one that is not, each
support lies at the pair's distance or one fp32 ulp either side of it.  The
expected decision comes from the exact-rational model of modDistance
(cpp/common.h:95-129, tests/exact_arith.py), which rounds exactly where the
reference does.  The oracle must decide every triple as the model does; the
same set must flip at least one decision under each plausible mistake the
model lists (fp32 moduloOp, fp32 squares, an fp64 accumulator, fp64
differences, '<=' for '<'), and none under the fused fp64 accumulate, which
is provably the same arithmetic.  Where oracle/_ref is built, the
reference's own countNeighbors decides the same triples too.
"""
import numpy as np
import pytest

import oracle
from oracle import ref
import exact_arith as ex

CASES = [
    (2, [False, False], [-1.0, -1.0], [1.0, 1.0]),
    (2, [True, True], [-1.0, -1.0], [1.0, 1.0]),
    (2, [True, False], [0.1, -0.3], [0.7, 0.9]),
    (3, [False, False, False], [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]),
    (3, [True, True, True], [0.1, 0.1, 0.1], [0.7, 0.7, 0.7]),
    (3, [True, False, True], [-1.0, -0.5, -2.0], [1.0, 1.5, 0.25]),
]
PAIRS = 70  # x 3 supports x 6 cases = 1,260 triples


def _triples(case_idx):
    dim, per, mn, mx = CASES[case_idx]
    rng = np.random.default_rng(4200 + case_idx)
    mn32, mx32 = np.asarray(mn, np.float32), np.asarray(mx, np.float32)
    xi, xj, ds = ex.boundary_pairs(rng, dim, per, mn32, mx32, PAIRS, lambda p: oracle.wrap(p, mn32, mx32, per))
    return dim, per, mn32, mx32, xi, xj, ds


@pytest.fixture(scope="module")
def all_triples():
    return [_triples(c) for c in range(len(CASES))]


@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_oracle_decides_as_exact_model(case_idx, all_triples):
    dim, per, mn, mx, xi, xj, ds = all_triples[case_idx]
    n_in = 0
    for t in range(xi.shape[0]):
        od = oracle.pair_distance(xi[t], xj[t], mn, mx, per)
        assert ex.fr(od) == ds[t], f"pair {t}: oracle distance {od!r} != exact {float(ds[t])!r}"
        for h in ex.supports_around(ds[t]):
            want = ex.inside(ds[t], h)
            assert bool(od < h) == want
            n_in += want
    assert n_in == xi.shape[0]  # exactly the support one ulp above the distance is in


def test_mistakes_flip_boundary_decisions(all_triples):
    flips = {v: 0 for v in ex.VARIANTS + ("fma64",)}
    for dim, per, mn, mx, xi, xj, ds in all_triples:
        for t in range(xi.shape[0]):
            for v in flips:
                dv = ex.dist(xi[t], xj[t], mn, mx, per, None if v == "le" else v)
                for h in ex.supports_around(ds[t]):
                    flips[v] += ex.inside(dv, h, v) != ex.inside(ds[t], h)
    print(flips)
    for v in ex.VARIANTS:
        assert flips[v] > 0, f"the boundary set does not discriminate {v}"
    assert flips["fma64"] == 0


def test_exact_model_known_values():
    """Hand-checkable anchors of the model: a 3-4-5 triangle (exact), a
    periodic wrap across [0.1, 0.7] (0.65 vs 0.15 is 0.1 apart through the
    boundary, in fp32 arithmetic), and sqrt(2) correctly rounded."""
    f = np.float32
    assert ex.dist([f(0.0), f(0.0)], [f(0.75), f(1.0)], [f(-2)] * 2, [f(2)] * 2, [False, False]) == ex.fr(1.25)
    d = ex.dist([f(0.65)], [f(0.15)], [f(0.1)], [f(0.7)], [True])
    assert abs(float(d) - 0.1) < 1e-6
    assert ex.sqrt_f32(ex.Fraction(2)) == ex.fr(np.float32(np.sqrt(np.float64(2))))
    # ties to even at fp32 precision: 1 + 2^-24 rounds to 1, 1 + 3*2^-24 to 1 + 2^-22
    assert ex.f32(1 + ex.Fraction(1, 2 ** 24)) == 1
    assert ex.f32(1 + ex.Fraction(3, 2 ** 24)) == 1 + ex.Fraction(1, 2 ** 22)


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref (the reference's own ops) is not built")
@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_reference_ops_decide_as_exact_model(case_idx, all_triples):
    """The reference's compiled countNeighbors on one (query, reference) pair
    per triple agrees with the exact model."""
    dim, per, mn, mx, xi, xj, ds = all_triples[case_idx]
    for t in range(0, xi.shape[0], 3):
        for h in ex.supports_around(ds[t]):
            c, _, _ = ref.search(xi[t:t + 1], np.array([h], np.float32), xj[t:t + 1], np.array([h], np.float32),
                                 domain_min=mn, domain_max=mx, periodic=per, mode="gather", hash_map_length=64)
            assert int(c[0]) == int(ex.inside(ds[t], h))
