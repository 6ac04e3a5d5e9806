"""Serving loop: continuous batching of radius-search requests on one GPU.

Requests arrive over time (e.g. from many running simulations).  A single
host loop

* moves arrived requests into a FIFO queue,
* admits queued requests into a new batch while the estimated HBM footprint
  of everything in flight stays under the budget and a launch slot (a HIP
  stream) is free -- footprint = workspace of the request's points + the
  predicted pairs (from that simulation's previous step) x index bytes,
* launches each batch asynchronously as one batched search (all clouds in
  one launch chain, outputs sized from the predictions, no host round trip
  between count and fill) on its own stream, so batches overlap on the device,
* completes batches as their events fire, frees their budget, and records
  each request's latency (completion - arrival).

A request served alone reuses its simulation's resident plan -- workspace,
offsets and output buffers kept across timesteps, optionally a captured HIP
graph of the whole search -- after its positions are copied into the plan's
buffer.  Each simulation's positions stay resident on the device and advance
in place; its previous pair count sizes the next step's outputs.

The scheduling core (``Scheduler``) is independent of the device; the GPU
backend (``GpuBackend``) runs batches through ``SearchPlan``.
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

    def advance(self, torch) -> None:
        """Move every point by a seeded displacement of length <= max_disp
        (uniform direction, uniform radius fraction)."""
        self.step += 1
        g = torch.Generator(device=self.pos.device)
        g.manual_seed(self.seed * 1000003 + self.step)
        d = torch.randn(self.pos.shape, generator=g, device=self.pos.device, dtype=torch.float32)
        d = d / d.norm(dim=1, keepdim=True).clamp_min(1e-12)
        r = torch.rand((self.pos.shape[0], 1), generator=g, device=self.pos.device, dtype=torch.float32)
        self.pos.add_(d * r * self.max_disp)


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
    """Runs batches of requests through batched search plans on ``slots``
    HIP streams.  ``batched=False`` gives the per-request plain baseline:
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
        self.streams = [torch.cuda.Stream() for _ in range(self.slots)] if batched else [None]
        self._next = 0
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
        r.payload = s.pos.clone()

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
        stream = self.streams[self._next % self.slots]
        self._next += 1
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            for r in b.requests:
                r.payload.record_stream(stream)
            if len(b.requests) == 1 and not self.sims[b.requests[0].run].plan_busy:
                handle = self._launch_alone(b.requests[0], stream)
            else:
                handle = self._launch_batch(b, stream)
            ev = torch.cuda.Event()
            ev.record(stream)
        return (handle, ev, stream)

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

    def _launch_batch(self, b: Batch, stream):
        from .search import SearchPlan
        from ._lib import check, ptr
        plan = SearchPlan([self._spec(r, r.payload) for r in b.requests], want_rows=False,
                          index_dtype=self.index_dtype)
        cap = int(sum(r.predicted_pairs for r in b.requests) * 1.03) + 256
        row, col = plan._alloc(cap)
        check(plan.L.tcr_search_launch(*plan._args(), ptr(plan.offsets), None, ptr(col), plan.index_bytes, cap,
                                       stream.cuda_stream))
        return ("batch", plan, col, cap)

    def ready(self, handle) -> bool:
        return handle is None or handle[1].query()

    def wait(self, handle) -> None:
        if handle is not None:
            handle[1].synchronize()

    def complete(self, b: Batch) -> None:
        torch = self.torch
        if not self.batched:
            return self._serve_plain(b)
        handle, ev, stream = b.handle
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
                _, plan, col, cap = handle
                P = plan._finish(stream.cuda_stream)
                if P > cap:
                    _, col = plan._fill_exact(P, stream.cuda_stream)
                starts, q0 = [], 0
                for r in b.requests:
                    starts.append(q0)
                    q0 += r.n
                starts.append(q0)
                idx = torch.tensor(starts, dtype=torch.int64).to(plan.offsets.device, non_blocking=True)
                bounds = plan.offsets[idx].cpu().tolist()
                for k, r in enumerate(b.requests):
                    a, e = bounds[k], bounds[k + 1]
                    self._deliver(r, plan.offsets[starts[k]:starts[k + 1] + 1], a, col[a:e], e - a)
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
