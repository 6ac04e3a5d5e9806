"""Launch-shape autotuner of the batched search.

Times a ``SearchPlan`` under candidate launch shapes (workgroups of the
wavefront-per-cell count/fill kernels, workgroups of the stencil kernel)
with the library's own HIP-event phase timing, keeps the fastest, and caches
the choice per batch shape.  The hash map length is not tuned: it is part of
the search's definition (it decides which aliased cells a lookup sees), so
it stays what the caller asked for.
"""
from __future__ import annotations

import ctypes
import statistics
from typing import Dict, Iterable, Optional, Tuple

from . import _lib
from ._lib import TcrTuning

_cache: Dict[tuple, Tuple[int, int]] = {}

CELL_CANDIDATES = (0, 256, 512, 1024, 2048, 4096)
STENCIL_CANDIDATES = (0, 512, 1024, 4096)


def _time_plan(plan, tuning: Tuple[int, int], reps: int) -> float:
    L = _lib.load()
    plan.tuning = tuning
    L.tcr_timing_collect(None, None, None, 0, None)
    L.tcr_timing_enable(1)
    try:
        for _ in range(reps):
            plan.run()
    finally:
        L.tcr_timing_enable(0)
    n = max(reps, 1)
    cms = (ctypes.c_float * n)()
    fms = (ctypes.c_float * n)()
    got = ctypes.c_int32(0)
    L.tcr_timing_collect(cms, fms, None, n, ctypes.byref(got))
    tot = [max(c, 0.0) + max(f, 0.0) for c, f in zip(cms[:got.value], fms[:got.value])]
    return statistics.median(tot) if tot else float("inf")


def autotune(plan, reps: int = 3, cell_candidates: Iterable[int] = CELL_CANDIDATES,
             stencil_candidates: Iterable[int] = STENCIL_CANDIDATES, force: bool = False) -> Tuple[int, int]:
    """Pick (cell_workgroups, stencil_workgroups) for this plan's batch shape
    (coordinate descent: cell kernels first, then the stencil kernel); the
    result is cached per shape and installed on the plan."""
    if not force and plan.key in _cache:
        plan.tuning = _cache[plan.key]
        return plan.tuning
    plan.tuning = (0, 0)
    plan.run()  # warm-up: learns the pair count so the timed runs take the back-to-back path
    best_c, best_t = 0, float("inf")
    for c in cell_candidates:
        t = _time_plan(plan, (c, 0), reps)
        if t < best_t:
            best_c, best_t = c, t
    best_s = 0
    for s in stencil_candidates:
        if s == 0:
            continue
        t = _time_plan(plan, (best_c, s), reps)
        if t < best_t:
            best_s, best_t = s, t
    _cache[plan.key] = (best_c, best_s)
    plan.tuning = (best_c, best_s)
    return plan.tuning


_installed: list = ["unset"]  # the launch shape last set through install()


def install(tuning: Optional[Tuple[int, int]]) -> None:
    """Set the library's launch shape (None: automatic); a repeat of the
    shape last set is a no-op."""
    key = None if tuning is None else (int(tuning[0]), int(tuning[1]))
    if _installed[0] == key:
        return
    L = _lib.load()
    if key is None:
        _lib.check(L.tcr_set_tuning(None))
    else:
        t = TcrTuning(*key)
        _lib.check(L.tcr_set_tuning(ctypes.byref(t)))
    _installed[0] = key
