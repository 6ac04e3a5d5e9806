"""Launch-shape autotuner of the batched search.

Times a ``SearchPlan`` under candidate launch shapes with the library's own
HIP-event phase timing, keeps the fastest, and caches the choice per batch
shape and device -- in memory and in a JSON file, so that a later process
reuses it without timing again.  A launch shape is

    (cell_workgroups, stencil_workgroups, cell_threads, cells_per_wave)

-- the workgroups of the wavefront-per-cell count/fill kernels, the
workgroups of the stencil kernel, the threads per workgroup of the cell
kernels (64, 128 or 256) and the run of consecutive cells a wavefront takes
in turn; 0 means the library's default for each.  The hash map length is not
tuned: it is part of the search's definition (it decides which aliased cells
a lookup sees), so it stays what the caller asked for.

The shape is installed per host thread (the library keeps one per thread),
so plans tuned differently can run on different threads.

Cache file: ``$TCR_AUTOTUNE_CACHE`` if set (an empty value disables the
file), else ``~/.cache/torchcompactradius_amd/autotune.json``.
"""
from __future__ import annotations

import ctypes
import json
import os
import statistics
import tempfile
import threading
from typing import Dict, Iterable, Optional, Tuple

from . import _lib
from ._lib import TcrTuning

Shape = Tuple[int, int, int, int]

CELL_CANDIDATES = (0, 256, 512, 1024, 2048, 4096)
STENCIL_CANDIDATES = (0, 512, 1024, 4096)
THREADS_CANDIDATES = (0, 128, 64)
RUN_CANDIDATES = (0, 2, 4, 8)

_cache: Dict[str, Shape] = {}
_file_loaded = [False]
_lock = threading.Lock()


def cache_path() -> Optional[str]:
    """The persisted cache's path, or None when the file is disabled."""
    env = os.environ.get("TCR_AUTOTUNE_CACHE")
    if env is not None:
        return env or None
    return os.path.join(os.path.expanduser("~"), ".cache", "torchcompactradius_amd", "autotune.json")


def normalize(tuning) -> Optional[Shape]:
    """A full 4-value shape from None or a tuple of 2 to 4 values (missing
    trailing values are 0, the defaults)."""
    if tuning is None:
        return None
    t = tuple(int(v) for v in tuning)
    if not 2 <= len(t) <= 4:
        raise ValueError("a launch shape has 2 to 4 values")
    return (t + (0, 0))[:4]


def shape_key(plan, device_name: str) -> str:
    """Cache key of a plan: the device and the batch shape (clouds, points,
    dimensions and modes of the first clouds)."""
    return json.dumps([device_name, [int(v) if isinstance(v, int) else [list(x) for x in v] for v in plan.key]])


def _device_name(plan) -> str:
    import torch
    return torch.cuda.get_device_name(plan.device)


def load_file(path: Optional[str] = None) -> Dict[str, Shape]:
    """Entries of the cache file (empty when it is absent or unreadable)."""
    path = path if path is not None else cache_path()
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path, "r", encoding="utf-8") as f:
            raw = json.load(f)
        return {k: normalize(v) for k, v in raw.items() if isinstance(v, list)}
    except (OSError, ValueError):
        return {}


def save_file(entries: Dict[str, Shape], path: Optional[str] = None) -> None:
    """Merge entries into the cache file (written to a temporary file, then
    renamed over the old one)."""
    path = path if path is not None else cache_path()
    if not path:
        return
    merged = load_file(path)
    merged.update(entries)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    try:
        with os.fdopen(fd, "w", encoding="utf-8") as f:
            json.dump({k: list(v) for k, v in merged.items()}, f, indent=0, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        if os.path.exists(tmp):
            os.unlink(tmp)


def _lookup(key: str) -> Optional[Shape]:
    with _lock:
        if not _file_loaded[0]:
            for k, v in load_file().items():
                _cache.setdefault(k, v)
            _file_loaded[0] = True
        return _cache.get(key)


def _time_plan(plan, tuning: Shape, reps: int) -> float:
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
             stencil_candidates: Iterable[int] = STENCIL_CANDIDATES,
             threads_candidates: Iterable[int] = THREADS_CANDIDATES,
             run_candidates: Iterable[int] = RUN_CANDIDATES, force: bool = False,
             persist: bool = True) -> Shape:
    """Pick the launch shape for this plan's batch shape by coordinate
    descent (threads per workgroup and cell workgroups of the cell kernels,
    then the cells per wavefront, then the stencil workgroups); the result is
    cached per device and batch shape (with ``persist``, also in the cache
    file) and installed on the plan."""
    key = shape_key(plan, _device_name(plan))
    if not force:
        hit = _lookup(key)
        if hit is not None:
            plan.tuning = hit
            return hit
    plan.tuning = (0, 0, 0, 0)
    plan.run()  # warm-up: learns the pair count so the timed runs take the back-to-back path
    best, best_t = (0, 0, 0, 0), float("inf")
    for th in threads_candidates:
        for c in cell_candidates:
            shape = (c, 0, th, 0)
            t = _time_plan(plan, shape, reps)
            if t < best_t:
                best, best_t = shape, t
    for r in run_candidates:
        if r == 0:
            continue
        shape = (best[0], 0, best[2], r)
        t = _time_plan(plan, shape, reps)
        if t < best_t:
            best, best_t = shape, t
    for s in stencil_candidates:
        if s == 0:
            continue
        shape = (best[0], s, best[2], best[3])
        t = _time_plan(plan, shape, reps)
        if t < best_t:
            best, best_t = shape, t
    with _lock:
        _cache[key] = best
    if persist:
        save_file({key: best})
    plan.tuning = best
    return best


_installed = threading.local()  # the launch shape this thread last set through install()


def install(tuning) -> None:
    """Set the calling thread's launch shape in the library (None:
    automatic); a repeat of the shape this thread last set is a no-op."""
    key = normalize(tuning)
    if getattr(_installed, "key", "unset") == key:
        return
    L = _lib.load()
    if key is None:
        _lib.check(L.tcr_set_tuning(None))
    else:
        t = TcrTuning(*key)
        _lib.check(L.tcr_set_tuning(ctypes.byref(t)))
    _installed.key = key
