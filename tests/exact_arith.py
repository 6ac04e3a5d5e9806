"""Exact-rational model of the reference's pair distance (test helper).

``dist`` evaluates ``modDistance`` (cpp/common.h:106-129) on fp32 inputs with
``fractions.Fraction``, rounding at exactly the points the reference's CPU
build rounds (x86-64, -O2, no FMA contraction):

* ``x_i - x_j``: fp32 subtraction (non-periodic dims, and ``p - q`` inside
  ``moduloOp``);
* ``moduloOp(p, q, h) = ((p-q) + h/2.0) - floor(((p-q) + h/2.0) / h) * h
  - h/2.0``: every operation in fp64, with ``h = (float)(max - min)``;
* ``sum += diff * diff``: the square and the add in fp64, the result stored
  back to the fp32 accumulator;
* ``std::sqrt(sum)`` on a float: the correctly rounded fp32 square root.

``VARIANTS`` are plausible mistakes of an implementation (each is checked to
flip at least one boundary decision, so the boundary set discriminates them):
``mod32`` (moduloOp in fp32), ``sq32`` (fp32 squares), ``acc64`` (the sum
kept in fp64 across dimensions, rounded once), ``diff64`` (the coordinate
differences in fp64), ``le`` (``<=`` instead of ``<``).  ``fma64`` (the
square fused into the fp64 add) is provably equivalent: every ``diff`` is an
fp32 value, so ``diff * diff`` is exact in fp64 and fusing it changes nothing.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

F0 = Fraction(0)


def _round(x: Fraction, p: int, emin: int) -> Fraction:
    """Round to the nearest binary float with a p-bit significand (ties to
    even); emin is the smallest normal exponent (subnormals below it)."""
    if x == 0:
        return F0
    s = -1 if x < 0 else 1
    a = -x if x < 0 else x
    n, d = a.numerator, a.denominator
    e = n.bit_length() - d.bit_length()
    if (n << max(0, -e)) < (d << max(0, e)):  # a < 2^e
        e -= 1
    e = max(e, emin)
    q = e - (p - 1)  # exponent of the last significand bit
    if q >= 0:
        num, den = n, d << q
    else:
        num, den = n << -q, d
    m, r = divmod(num, den)
    if 2 * r > den or (2 * r == den and m & 1):
        m += 1
    return s * (Fraction(m) * (Fraction(2) ** q))


def f32(x: Fraction) -> Fraction:
    return _round(x, 24, -126)


def f64(x: Fraction) -> Fraction:
    return _round(x, 53, -1022)


def fr(v) -> Fraction:
    return Fraction(float(v))


def floor_fr(x: Fraction) -> Fraction:
    return Fraction(math.floor(x))


def sqrt_f32(s: Fraction) -> Fraction:
    """Correctly rounded fp32 square root of a non-negative fp32 value."""
    if s == 0:
        return F0
    r = np.float32(math.sqrt(float(s)))
    for c in (r, np.nextafter(r, np.float32(0)), np.nextafter(r, np.float32(np.inf))):
        c = np.float32(c)
        up = fr(np.nextafter(c, np.float32(np.inf))) - fr(c)
        dn = fr(c) - fr(np.nextafter(c, np.float32(0)))
        lo, hi = fr(c) - dn / 2, fr(c) + up / 2
        if lo * lo < s < hi * hi:  # no fp32 square root lies on a midpoint
            return fr(c)
    raise AssertionError("sqrt search failed")


def modulo_op(p: Fraction, q: Fraction, h: Fraction, variant=None) -> Fraction:
    """moduloOp (cpp/common.h:95-104) with p - q in fp32 and the rest fp64."""
    if variant == "mod32":
        pq = f32(p - q)
        h2 = f32(h / 2)
        t = f32(pq + h2)
        k = floor_fr(f32(t / h))
        return f32(f32(t - f32(k * h)) - h2)
    pq = f64(p - q) if variant == "diff64" else f32(p - q)
    h2 = h / 2  # h / 2.0 of an fp32 value: exact in fp64
    t = f64(pq + h2)
    k = floor_fr(f64(t / h))
    return f64(f64(t - f64(k * h)) - h2)


def dist(xi, xj, mn, mx, periodic, variant=None) -> Fraction:
    """modDistance (cpp/common.h:106-129) of fp32 points xi, xj."""
    s = F0
    for d in range(len(xi)):
        p, q = fr(xi[d]), fr(xj[d])
        if periodic[d]:
            h = f32(fr(mx[d]) - fr(mn[d]))
            diff = modulo_op(p, q, h, variant)
        else:
            diff = f64(p - q) if variant == "diff64" else f32(p - q)
        if variant == "sq32":
            s = f32(s + f32(diff * diff))
        elif variant == "acc64":
            s = f64(s + f64(diff * diff))
        elif variant == "fma64":
            s = f32(f64(s + diff * diff))
        else:
            s = f32(f64(s + f64(diff * diff)))
    if variant == "acc64":
        s = f32(s)
    return sqrt_f32(s)


def inside(d: Fraction, h, variant=None) -> bool:
    """The gather/scatter predicate dist < h (strict, neighborhoodDynamic.h:63-70)."""
    return d <= fr(h) if variant == "le" else d < fr(h)


VARIANTS = ("mod32", "sq32", "acc64", "diff64", "le")


def boundary_pairs(rng, dim, periodic, mn, mx, n_pairs, wrap):
    """n_pairs point pairs (xi, xj) in the domain, each within ~0.02 of
    extent of each other, a third of them across a periodic boundary where
    there is one; positions are fixed points of `wrap` (the reference's
    once/twice wrap leaves them unchanged).  Returns xi, xj fp32 [n, dim] and
    the exact distances."""
    mn = np.asarray(mn, np.float64)
    mx = np.asarray(mx, np.float64)
    ext = mx - mn
    xi = np.empty((n_pairs, dim), np.float32)
    xj = np.empty((n_pairs, dim), np.float32)
    k = 0
    while k < n_pairs:
        a = mn + rng.uniform(0.05, 0.95, dim) * ext
        step = rng.normal(0, 1, dim)
        step *= rng.uniform(0.005, 0.03) * ext.min() / np.linalg.norm(step)
        if k % 3 == 0 and any(periodic):
            d0 = int(np.nonzero(periodic)[0][rng.integers(0, int(np.sum(periodic)))])
            a[d0] = mx[d0] - rng.uniform(0, 0.01) * ext[d0]
            step[d0] = abs(step[d0])
        b = a + step
        for d in range(dim):
            if periodic[d] and b[d] >= mx[d]:
                b[d] -= ext[d]
        pa, pb = np.float32(a), np.float32(b)
        pair = np.stack([pa, pb])
        if not np.array_equal(wrap(pair), pair):
            continue
        if np.any(pair < np.float32(mn)) or np.any(pair >= np.float32(mx)):
            continue
        xi[k], xj[k] = pa, pb
        k += 1
    ds = [dist(xi[t], xj[t], mn.astype(np.float32), mx.astype(np.float32), periodic) for t in range(n_pairs)]
    return xi, xj, ds


def supports_around(d: Fraction):
    """fp32 supports one ulp below, at and one ulp above an fp32 distance."""
    v = np.float32(float(d))
    return [np.nextafter(v, np.float32(0)), v, np.nextafter(v, np.float32(np.inf))]


def island_pairs(rng, dim, periodic, mn, mx, n_side, wrap):
    """One pair per island of an n_side^dim grid over the domain, islands
    ten pair lengths apart (no pair between islands for supports near the
    pair length); the islands of the last row of a periodic dimension put
    their pair across that boundary.  Positions are fixed points of `wrap`.
    Returns xi, xj fp32 [n, dim] and the exact distances."""
    mn = np.asarray(mn, np.float64)
    mx = np.asarray(mx, np.float64)
    ext = mx - mn
    sp = ext / n_side
    plen = 0.08 * sp.min()
    idx = np.stack(np.meshgrid(*[np.arange(n_side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    xi = np.empty((idx.shape[0], dim), np.float32)
    xj = np.empty_like(xi)
    for k, g in enumerate(idx):
        while True:
            a = mn + (g + 0.5) * sp + rng.uniform(-0.1, 0.1, dim) * sp
            step = rng.normal(0, 1, dim)
            step *= plen * rng.uniform(0.7, 1.0) / np.linalg.norm(step)
            across = [d for d in range(dim) if periodic[d] and g[d] == n_side - 1]
            for d in across:
                a[d] = mx[d] - rng.uniform(0.0, 0.5) * plen
                step[d] = abs(step[d]) + 0.3 * plen
            b = a + step
            for d in range(dim):
                if periodic[d] and b[d] >= mx[d]:
                    b[d] -= ext[d]
            pair = np.stack([np.float32(a), np.float32(b)]).astype(np.float32)
            if np.array_equal(wrap(pair), pair) and np.all(pair >= np.float32(mn)) and np.all(pair < np.float32(mx)):
                break
        xi[k], xj[k] = pair
    mn32, mx32 = mn.astype(np.float32), mx.astype(np.float32)
    ds = [dist(xi[t], xj[t], mn32, mx32, periodic) for t in range(xi.shape[0])]
    return xi, xj, ds
