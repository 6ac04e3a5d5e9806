"""Serving loop: continuous batching of radius-search requests on one GPU.

Requests arrive over time (e.g. from many running simulations).  A single
host loop

* moves arrived requests into a FIFO queue,
* admits queued requests into a new batch while the estimated HBM footprint
  of everything in flight stays under the budget and a launch slot (a set
  of resident batch buffers) is free -- footprint = workspace of the request's points + the
  predicted pairs (from that simulation's previous step) x index bytes,
* launches each batch asynchronously as one batched search (all clouds in
  one launch chain, outputs sized from the predictions, no host round trip
  between count and fill) on its slot's stream, ordered after the previous
  batch: each batch fills the device, and the host completes one while the
  next runs,
* completes batches as their events fire, frees their budget, and records
  each request's latency (completion - arrival).

A request served alone reuses its simulation's resident plan -- workspace,
offsets and output buffers kept across timesteps, optionally a captured HIP
graph of the whole search -- after its positions are copied into the plan's
buffer.  Each simulation's positions stay resident on the device and advance
in place; its previous pair count sizes the next step's outputs.

The scheduling core (``Scheduler``) is independent of the device; the GPU
backend (``GpuBackend``) runs batches through the library's batched search
entry points.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional, Sequence

import numpy as np


@dataclass
class Request:
    rid: int
    run: int                      # simulation id
    step: int                     # timestep of that simulation
    arrival: float                # seconds from the start of the stream
    n: int                        # points
    predicted_pairs: int
    payload: object = None        # backend data (position snapshot)
    admitted: float = -1.0
    done: float = -1.0
    result: object = None
    inputs: object = None         # position snapshot, kept when the backend keeps results


@dataclass
class Batch:
    requests: List[Request]
    footprint: int
    handle: object = None


class Scheduler:
    """Admission and continuous batching over an abstract backend with
    ``footprint(req)``, ``launch(batch) -> handle``, ``ready(handle)``,
    ``complete(batch)`` (results into the requests) and ``slots``."""

    def __init__(self, backend, budget_bytes: float = 200e9, max_batch: int = 64,
                 clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep):
        self.backend = backend
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.clock = clock
        self.sleep = sleep
        self.in_use = 0
        self.peak_in_use = 0
        self.batches: List[int] = []       # sizes of the launched batches

    def _admit(self, queue: Deque[Request], now: float) -> Optional[Batch]:
        reqs, fp = [], 0
        while queue and len(reqs) < self.max_batch:
            f = self.backend.footprint(queue[0])
            if f > self.budget:
                raise ValueError(f"request {queue[0].rid} needs {f} bytes, more than the whole budget")
            if self.in_use + fp + f > self.budget:
                break
            r = queue.popleft()
            r.admitted = now
            reqs.append(r)
            fp += f
        if not reqs:
            return None
        self.in_use += fp
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        self.batches.append(len(reqs))
        return Batch(reqs, fp)

    def run(self, requests: Sequence[Request]) -> List[Request]:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: Deque[Request] = collections.deque()
        inflight: Deque[Batch] = collections.deque()
        t0 = self.clock()
        done: List[Request] = []

        def finish(b: Batch):
            self.backend.complete(b)
            t = self.clock() - t0
            for r in b.requests:
                r.done = t
            self.in_use -= b.footprint
            done.extend(b.requests)

        while pending or queue or inflight:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                self.backend.arrive(r)
                queue.append(r)
            progressed = False
            while inflight and self.backend.ready(inflight[0].handle):
                finish(inflight.popleft())
                progressed = True
            if queue and len(inflight) < self.backend.slots:
                b = self._admit(queue, now)
                if b is not None:
                    b.handle = self.backend.launch(b)
                    inflight.append(b)
                    progressed = True
            if progressed:
                continue
            if inflight and (len(inflight) >= self.backend.slots or not pending or queue):
                # nothing can be admitted now: wait for the oldest batch
                self.backend.wait(inflight[0].handle)
                finish(inflight.popleft())
            elif pending:
                # poll finely while batches run (their completion is the
                # latency), coarsely when only the next arrival is awaited
                dt = pending[0].arrival - (self.clock() - t0)
                if dt > 0:
                    self.sleep(min(dt, 5e-5 if inflight else 1e-3))
        return done


def latency_stats(reqs: Sequence[Request]) -> Dict[str, float]:
    lat = np.array([r.done - r.arrival for r in reqs], dtype=np.float64)
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}


def poisson_requests(n_requests: int, rate_hz: float, sizes: Sequence[int], seed: int,
                     predicted_pairs: Sequence[int]) -> List[Request]:
    """Seeded Poisson arrivals; each arrival comes from a uniformly drawn
    simulation and is that simulation's next timestep."""
    rng = np.random.default_rng(seed)
    t = 0.0
    steps = [0] * len(sizes)
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        run = int(rng.integers(0, len(sizes)))
        steps[run] += 1
        out.append(Request(k, run, steps[run], t, int(sizes[run]), int(predicted_pairs[run])))
    return out


# ------------------------------------------------------------------ GPU ----

def _mix64(x: int) -> int:
    """splitmix64 finaliser: a seeded integer hash (the per-step shift of a
    simulation's displacement field)."""
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


@dataclass
class Simulation:
    """One simulated SPH run: resident positions advanced in place by seeded
    displacements of length at most ``max_disp`` per step."""
    rid: int
    pos: object                   # torch float32 [N, 3] on the device
    sup: object                   # torch float32 [N]
    h: float
    domain_min: List[float]
    domain_max: List[float]
    periodic: List[bool]
    mode: str = "gather"
    hash_map_length: int = 1 << 16
    seed: int = 0
    step: int = 0
    last_pairs: int = 0
    plan: object = None           # resident SearchPlan for requests served alone
    plan_cap: int = 0
    plan_busy: bool = False       # the resident plan is in flight (one use at a time)
    plan_out: object = None       # resident output buffer (plan_cap entries)
    max_disp: float = 0.0
    disp: object = None           # seeded displacement field (built on the first step)

    def advance(self, torch) -> None:
        """Move every point by a seeded displacement of length <= max_disp:
        the simulation's displacement field (uniform directions, uniform
        radius fractions, drawn once from its seed) rolled over the points by
        a seeded per-step shift, so each step moves each point differently
        at the cost of one add.  The sum is a new tensor: the previous
        positions stay intact as the snapshot of the request that saw them."""
        self.step += 1
        n = self.pos.shape[0]
        if self.disp is None:
            g = torch.Generator(device=self.pos.device)
            g.manual_seed(self.seed * 1000003)
            d = torch.randn((2 * n, 3), generator=g, device=self.pos.device, dtype=torch.float32)
            d = d / d.norm(dim=1, keepdim=True).clamp_min(1e-12)
            r = torch.rand((2 * n, 1), generator=g, device=self.pos.device, dtype=torch.float32)
            self.disp = d * (r * self.max_disp)
        shift = _mix64(self.seed * 1000003 + self.step) % (n + 1)
        self.pos = self.pos + self.disp[shift:shift + n]


def make_simulations(torch, n_runs: int = 64, n_min: float = 1e3, n_max: float = 1e6, seed: int = 5,
                     neighbors: float = 50.0, device: str = "cuda") -> List[Simulation]:
    """Simulations with log-uniform sizes: jittered 3-D grids in [-1,1]^3,
    periodic in every dimension, gather mode, ~``neighbors`` neighbours;
    displacements per step of at most 0.1 h."""
    from .synthetic import jittered_grid, support_for
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(n_runs):
        N = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        n = max(8, int(round(N ** (1.0 / 3.0))))
        dx = 2.0 / n
        pos = jittered_grid(n, 3, -1.0, 1.0, 0.1, rng)
        h = float(np.float32(support_for(neighbors, dx, 3)))
        sims.append(Simulation(k, torch.from_numpy(pos).to(device),
                               torch.full((pos.shape[0],), h, dtype=torch.float32, device=device), h,
                               [-1.0] * 3, [1.0] * 3, [True] * 3, seed=seed * 7919 + k, max_disp=0.1 * h,
                               last_pairs=int(pos.shape[0] * neighbors)))
    return sims


class GpuBackend:
    """Runs batches of requests as batched searches on ``slots`` launch
    slots -- a HIP stream and a set of resident batch buffers each, one batch
    in flight per slot -- in launch order.  ``batched=False`` gives the per-request plain baseline:
    one radiusSearch per request, one at a time, on the current stream."""

    def __init__(self, torch, sims: Sequence[Simulation], slots: int = 2, batched: bool = True,
                 index_dtype=None, keep_inputs: bool = False, resident_graph: bool = False):
        self.torch = torch
        self.sims = list(sims)
        self.batched = batched
        self.slots = slots if batched else 1
        self.index_dtype = index_dtype or torch.int64
        self.ib = 8 if self.index_dtype == torch.int64 else 4
        self.keep = keep_inputs
        self.resident_graph = resident_graph
        # a slot's resident buffers (concatenated positions and supports,
        # workspace, pinned first-pair readback) are reused by every batch it
        # runs.  Each batch waits for the previous one's end event: a batch
        # fills the device, so concurrent batches would only share it and
        # finish together, idling it while the host completes them; in
        # order, batch k+1 runs while the host completes batch k (whose
        # stream synchronisation waits for batch k alone)
        self.streams = [torch.cuda.Stream() for _ in range(self.slots)] if batched else [None]
        self.free_slots = collections.deque(range(self.slots))
        self.arena = [dict() for _ in range(self.slots)]
        self.last_event = None
        self.pairs = 0
        self.alone = 0              # requests served by a resident graph replay

    # footprint: workspace (about 460 B per point, 20 B per hash bucket) +
    # positions snapshot + offsets + predicted pairs x index bytes (4% headroom)
    def footprint(self, r: Request) -> int:
        L = self.sims[r.run].hash_map_length
        return int(r.n * 480 + 24 * L + r.predicted_pairs * self.ib * 1.04 + 65536)

    def arrive(self, r: Request) -> None:
        s = self.sims[r.run]
        while s.step < r.step:
            s.advance(self.torch)
        r.predicted_pairs = max(r.predicted_pairs, s.last_pairs)
        r.payload = s.pos      # advance() never writes it again: the snapshot

    def _spec(self, r: Request, pos):
        from .search import CloudSpec, periodic_mask
        s = self.sims[r.run]
        return CloudSpec(qpos=pos, qsup=s.sup, rpos=None, rsup=None, domain_min=s.domain_min,
                         domain_max=s.domain_max, periodic=periodic_mask(s.periodic, 3), mode=s.mode,
                         hash_map_length=s.hash_map_length)

    def launch(self, b: Batch):
        torch = self.torch
        if not self.batched:
            return None
        slot = self.free_slots.popleft()   # the scheduler keeps at most `slots` batches in flight
        stream = self.streams[slot]
        stream.wait_stream(torch.cuda.current_stream())
        if self.last_event is not None:
            stream.wait_event(self.last_event)
        # the requests hold their snapshots until complete() has synchronised
        # this stream, so the snapshots need no record_stream
        with torch.cuda.stream(stream):
            if len(b.requests) == 1 and not self.sims[b.requests[0].run].plan_busy:
                handle = self._launch_alone(b.requests[0], stream)
            else:
                handle = self._launch_batch(b, stream, slot)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.last_event = ev
        return (handle, ev, stream, slot)

    def _take(self, slot: int, name: str, n: int, dtype):
        """A view of n elements of the slot's resident buffer `name`, grown
        (by a quarter beyond the request) when too small."""
        buf = self.arena[slot].get(name)
        if buf is None or buf.numel() < n or buf.dtype != dtype:
            buf = self.torch.empty(max(1, n + n // 4), dtype=dtype, device="cuda")
            self.arena[slot][name] = buf
        return buf[:n]

    def _launch_alone(self, r: Request, stream):
        """Resident plan of the simulation (workspace, offsets and output
        buffers kept across its timesteps): copy the snapshot into the plan's
        buffer, then one direct launch of the whole chain or, with
        ``resident_graph``, a replay of the plan's captured graph."""
        from .search import SearchPlan
        from ._lib import check, ptr
        s = self.sims[r.run]
        cap = int(r.predicted_pairs * 1.03) + 256
        if s.plan is None or cap > s.plan_cap:
            if s.plan is not None:
                s.plan.release_graph()
            buf = r.payload.clone()
            s.plan = SearchPlan([self._spec(r, buf)], want_rows=False, index_dtype=self.index_dtype)
            s.plan_cap = cap
            if self.resident_graph:
                s.plan.run()
                s.plan.capture(cap)
                s.plan_out = s.plan._graph_out[1]
            else:
                s.plan_out = s.plan._alloc(cap)[1]
        s.plan.qpos.copy_(r.payload.reshape(-1))
        s.plan_busy = True
        self.alone += 1
        if self.resident_graph:
            check(s.plan.L.tcr_graph_launch(s.plan._graph, stream.cuda_stream))
        else:
            check(s.plan.L.tcr_search_launch(*s.plan._args(), ptr(s.plan.offsets), None, ptr(s.plan_out),
                                             s.plan.index_bytes, s.plan_cap, stream.cuda_stream))
        return ("alone", s.plan, s.plan_out, s.plan_cap)

    def _launch_batch(self, b: Batch, stream, slot: int):
        """One batched search over the slot's resident buffers: the requests'
        position snapshots and supports concatenated in place (one copy
        kernel each), the cloud descriptors, the slot's workspace; fresh
        outputs (the requests keep views of them)."""
        from . import _lib
        from ._lib import MODES, TcrCloud, check, ptr
        from .search import periodic_mask
        torch = self.torch
        L = _lib.load()
        n = len(b.requests)
        clouds = (TcrCloud * n)()
        q0 = 0
        for k, r in enumerate(b.requests):
            s = self.sims[r.run]
            c = clouds[k]
            c.dim, c.mode, c.periodic_mask, c.search_range = 3, MODES[s.mode], periodic_mask(s.periodic, 3), 1
            for d in range(3):
                c.domain_min[d], c.domain_max[d] = s.domain_min[d], s.domain_max[d]
            c.hash_map_length, c.self_search = s.hash_map_length, 1
            c.q_off, c.nq, c.q_foff, c.r_off, c.nr, c.r_foff = q0, r.n, 3 * q0, 0, r.n, 0
            q0 += r.n
        qpos = self._take(slot, "qpos", 3 * q0, torch.float32)
        qsup = self._take(slot, "qsup", q0, torch.float32)
        torch.cat([r.payload.reshape(-1) for r in b.requests], out=qpos)
        torch.cat([self.sims[r.run].sup for r in b.requests], out=qsup)
        wsb = L.tcr_search_workspace_size(clouds, n)
        ws = self._take(slot, "ws", wsb, torch.uint8)
        cap = int(sum(r.predicted_pairs for r in b.requests) * 1.03) + 256
        # offsets and pair indices in one allocation (the indices start on a
        # 256-byte boundary): one allocator call per batch for its outputs
        head = (q0 + 1 + 31) // 32 * 32
        out = torch.empty(head + (cap * self.ib + 7) // 8, dtype=torch.int64, device="cuda")
        offsets = out[:q0 + 1]
        col = out[head:].view(self.index_dtype)[:cap]
        args = (clouds, n, ptr(qpos), ptr(qsup), None, None, ptr(ws), wsb)
        check(L.tcr_search_launch(*args, ptr(offsets), None, ptr(col), self.ib, cap, stream.cuda_stream))
        return ("batch", args, offsets, col, cap)

    def ready(self, handle) -> bool:
        return handle is None or handle[1].query()

    def wait(self, handle) -> None:
        if handle is not None:
            handle[1].synchronize()

    def complete(self, b: Batch) -> None:
        torch = self.torch
        if not self.batched:
            return self._serve_plain(b)
        handle, ev, stream, slot = b.handle
        with torch.cuda.stream(stream):
            if handle[0] == "alone":
                _, plan, col, cap = handle
                P = plan._finish(stream.cuda_stream)
                if P > cap:
                    _, col = plan._fill_exact(P, stream.cuda_stream)
                else:
                    col = col[:P].clone()      # the resident output buffer is reused by the next step
                self._deliver(b.requests[0], plan.offsets.clone(), 0, col, P)
                self.sims[b.requests[0].run].plan_busy = False
            else:
                import ctypes
                from . import _lib
                from ._lib import check, ptr
                L = _lib.load()
                _, args, offsets, col, cap = handle
                starts, q0 = [], 0
                for r in b.requests:
                    starts.append(q0)
                    q0 += r.n
                starts.append(q0)
                # each request's first pair: gathered on the stream before the
                # host waits, read after the one synchronisation
                idx = torch.tensor(starts, dtype=torch.int64).to(offsets.device, non_blocking=True)
                bounds_dev = offsets[idx]
                bounds_host = self.arena[slot].get("bounds")
                if bounds_host is None or bounds_host.numel() < len(starts):
                    bounds_host = torch.empty(max(len(starts), 65), dtype=torch.int64, pin_memory=True)
                    self.arena[slot]["bounds"] = bounds_host
                bounds_host = bounds_host[:len(starts)]
                bounds_host.copy_(bounds_dev, non_blocking=True)
                total = ctypes.c_int64(0)
                check(L.tcr_search_finish(args[0], args[1], args[6], args[7], ptr(offsets), ctypes.byref(total),
                                          stream.cuda_stream))
                P = int(total.value)
                if P > cap:
                    col = torch.empty(P, dtype=self.index_dtype, device="cuda")
                    check(L.tcr_search_fill(*args, ptr(offsets), None, ptr(col), self.ib, stream.cuda_stream))
                bounds = bounds_host.tolist()
                for k, r in enumerate(b.requests):
                    a, e = bounds[k], bounds[k + 1]
                    self._deliver(r, offsets[starts[k]:starts[k + 1] + 1], a, col[a:e], e - a)
        self.free_slots.append(slot)
        for r in b.requests:
            if self.keep:
                r.inputs = r.payload
            r.payload = None

    def _serve_plain(self, b: Batch) -> None:
        from . import radiusSearch
        from .util import DomainDescription, PointCloud
        torch = self.torch
        for r in b.requests:
            s = self.sims[r.run]
            dom = DomainDescription(torch.tensor(s.domain_min), torch.tensor(s.domain_max),
                                    torch.tensor(s.periodic), 3)
            out = radiusSearch(PointCloud(r.payload, s.sup), None, mode=s.mode, domain=dom,
                               hashMapLength=s.hash_map_length, algorithm="compact", format="csr",
                               indexDtype=self.index_dtype)
            self._deliver(r, out.indptr, 0, out.indices, int(out.indices.numel()))
            if self.keep:
                r.inputs = r.payload
            r.payload = None

    def _deliver(self, r: Request, indptr_src, base: int, col, P: int) -> None:
        """Hands the result to the request: CSR row pointers (int32 indptr,
        or a slice of the batch's int64 offsets minus ``base``) and columns."""
        s = self.sims[r.run]
        s.last_pairs = P
        self.pairs += P
        r.result = (indptr_src, base, col)

    def result(self, r: Request):
        """(indptr int32 [n+1], col [P]) of a completed request."""
        indptr_src, base, col = r.result
        if indptr_src.dtype != self.torch.int32:
            indptr_src = (indptr_src - base).to(self.torch.int32)
        return indptr_src, col

    def digest(self, r: Request) -> tuple:
        return result_digest(self.torch, *self.result(r))


def result_digest(torch, indptr, col) -> tuple:
    """Order-sensitive fingerprint of a CSR result: sizes plus two position-
    weighted wrapping int64 sums of the columns and of the row pointers."""
    c = col.to(torch.int64)
    i = torch.arange(c.numel(), device=c.device, dtype=torch.int64)
    p = indptr.to(torch.int64)
    j = torch.arange(p.numel(), device=p.device, dtype=torch.int64)
    a = int((c * (i * 2654435761 + 97)).sum().item()) if c.numel() else 0
    b = int(((c + 1) * (i ^ 0x5bd1e995)).sum().item()) if c.numel() else 0
    d = int((p * (j * 40503 + 11)).sum().item())
    return (int(c.numel()), int(p.numel()), a, b, d)
