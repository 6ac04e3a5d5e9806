"""CPU restatement of the reference's brute-force search (test infrastructure
only): searchAlgorithms/radiusNaive.py of torchCompactRadius 0.5.5, the
``algorithm='naive'`` branch of radiusSearch (neighborhood.py:226-229,
268-271).  numpy fp32 arithmetic in the order the torch ops take it (CPU):

* periodic coordinates wrapped by torch.remainder (fmod, then + divisor when
  the signs differ), once by radiusSearch's caller (util.py:375-387) and once
  more inside radiusNaive (:12-13);
* per dimension the fp32 difference x_a - y_b (:15), through radiusNaive's
  own ``mod`` (:4-6) in periodic dimensions, whose Python-float arguments
  reach the fp32 tensor ops rounded to fp32;
* squares summed in dimension order (torch.sum over the last axis), sqrtf;
* ``<=`` against the mode's support (:19-32), pairs ordered by reference
  index, then query index (the row-major flattening of the (ny, nx) mask,
  :17 and :36-37).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

f32 = np.float32


def torch_remainder(a: np.ndarray, b: np.float32) -> np.ndarray:
    """torch.remainder on fp32 with a fp32 divisor."""
    a = np.asarray(a, f32)
    b = f32(b)
    with np.errstate(invalid="ignore"):
        r = np.fmod(a, b).astype(f32)
        fix = (r != 0) & ((r < 0) != (b < 0))
    return np.where(fix, (r + b).astype(f32), r).astype(f32)


def wrap_once(p: np.ndarray, mn, mx, periodic) -> np.ndarray:
    """remainder(x - min, max - min) + min in periodic dimensions, fp32
    tensor arithmetic (util.py:387, radiusNaive.py:12-13)."""
    p = np.array(p, f32, copy=True)
    for d in range(p.shape[1]):
        if periodic[d]:
            lo, hi = f32(mn[d]), f32(mx[d])
            p[:, d] = (torch_remainder((p[:, d] - lo).astype(f32), f32(hi - lo)) + lo).astype(f32)
    return p


def naive_mod(x: np.ndarray, mn: float, mx: float) -> np.ndarray:
    """radiusNaive.py:4-6: where |x| > (max - min)/2, sgn(x) * ((|x| + min) %
    (max - min) + min), else x.  min/max arrive as Python floats (the fp32
    values); (max - min) and its half are formed in double and rounded to
    fp32 where they meet the fp32 tensors."""
    x = np.asarray(x, f32)
    mnd, mxd = float(mn), float(mx)
    half = f32((mxd - mnd) / 2.0)
    ext = f32(mxd - mnd)
    lo = f32(mnd)
    ax = np.abs(x)
    r = (torch_remainder((ax + lo).astype(f32), ext) + lo).astype(f32)
    with np.errstate(invalid="ignore"):
        s = np.where(x > 0, f32(1), np.where(x < 0, f32(-1), f32(0))).astype(f32)
        v = (s * r).astype(f32)
        return np.where(ax > half, v, x).astype(f32)


def _sq_dist(x: np.ndarray, y: np.ndarray, mn, mx, periodic) -> np.ndarray:
    """(ny, nx) fp32 sum over dimensions of the squared (mod) differences."""
    s = None
    for d in range(x.shape[1]):
        diff = (x[None, :, d] - y[:, None, d]).astype(f32)
        if periodic[d]:
            diff = naive_mod(diff, mn[d], mx[d])
        sq = (diff * diff).astype(f32)
        s = sq if s is None else (s + sq).astype(f32)
    return s


def radius_naive(x, y, hx, hy, periodic: Sequence[bool], mn, mx, mode: str = "gather"):
    """radiusNaive (radiusNaive.py:8-39): (ii, jj) int64, query and reference
    indices of every pair with distance <= the mode's support."""
    x, y = wrap_once(x, mn, mx, periodic), wrap_once(y, mn, mx, periodic)
    with np.errstate(invalid="ignore"):
        dist = np.sqrt(_sq_dist(x, y, mn, mx, periodic)).astype(f32)
        nx, ny = x.shape[0], y.shape[0]
        if mode == "gather":
            thr = np.broadcast_to(np.asarray(hx, f32)[None, :], (ny, nx))
        elif mode == "scatter":
            thr = np.broadcast_to(np.asarray(hy, f32)[:, None], (ny, nx))
        elif mode == "symmetric":
            thr = ((np.asarray(hx, f32)[None, :] + np.asarray(hy, f32)[:, None]).astype(f32) / f32(2)).astype(f32)
        elif mode == "superSymmetric":
            a, b = np.asarray(hx, f32), np.asarray(hy, f32)
            if a.shape != b.shape:
                raise ValueError("superSymmetric needs as many query as reference points (torch.max(hx, hy))")
            m = np.where(np.isnan(a) | np.isnan(b), f32(np.nan), np.maximum(a, b)).astype(f32)
            thr = np.broadcast_to(m[None, :], (ny, nx))
        else:
            raise ValueError("mode must be one of gather, scatter, symmetric, or superSymmetric")
        adj = dist <= thr
    jj, ii = np.nonzero(adj)          # row-major: reference index, then query index
    return ii.astype(np.int64), jj.astype(np.int64)


def radius_naive_fixed(x, y, h: float, periodic: Sequence[bool], mn, mx):
    """radiusNaiveFixed (radiusNaive.py:41-56): squared distance <= h**2, h**2
    formed in double and rounded to fp32 at the comparison."""
    x, y = wrap_once(x, mn, mx, periodic), wrap_once(y, mn, mx, periodic)
    with np.errstate(invalid="ignore"):
        adj = _sq_dist(x, y, mn, mx, periodic) <= f32(float(h) ** 2)
    jj, ii = np.nonzero(adj)
    return ii.astype(np.int64), jj.astype(np.int64)


def radius_search_naive(qpos, qsup, rpos=None, rsup=None, *, domain_min=None, domain_max=None,
                        periodic=None, mode: str = "gather", support_override: Optional[float] = None):
    """radiusSearch(..., algorithm='naive') (neighborhood.py:186-271): default
    domain [0,1]^d non-periodic; with a domain the caller wraps periodic
    coordinates once before radiusNaive wraps them again."""
    qpos = np.asarray(qpos, f32)
    dim = qpos.shape[1]
    if rpos is None:
        rpos, rsup = qpos, qsup
    rpos = np.asarray(rpos, f32)
    if domain_min is None:
        mn, mx, per = np.zeros(dim, f32), np.ones(dim, f32), [False] * dim
        x, y = qpos, rpos
    else:
        mn, mx = np.asarray(domain_min, f32), np.asarray(domain_max, f32)
        per = list(periodic) if periodic is not None else [False] * dim
        x, y = wrap_once(qpos, mn, mx, per), wrap_once(rpos, mn, mx, per)
    if support_override is not None:
        return radius_naive_fixed(x, y, support_override, per, mn, mx)
    return radius_naive(x, y, qsup, rsup, per, mn, mx, mode)
