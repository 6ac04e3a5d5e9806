#!/usr/bin/env python3
"""Benchmark of the compact-hashing radius search on one MI355X.

Workload (one step = one full search): 1M-point 3-D periodic cloud (100^3
jittered grid in [-1,1]^3, ~50 neighbours, gather, CSR, hashMapLength 2^20),
inputs resident in HBM; the step runs radiusSearch(..., algorithm='compact',
format='csr') end to end: structure build, count, scan and fill (into
outputs sized from the previous step's total, so the chain runs without a
host round trip in between), one host sync for the pair total, and the CSR
tuple.  The roofline figure is the dominant kernel of the search (the one
with the larger average time of the two big passes, the wavefront-per-cell
count kernel k_count_small and the fill kernel k_fill_small), each timed
with HIP events that the library records on the search stream around it,
in a second run of as many steps right after the timed ones (the events
are kept out of the wall-clock steps); both kernels' figures and the
end-to-end fraction are reported beside it.

Prints one JSON line (rank 0).  The CPU baseline is the C/OpenMP oracle
timed on this host on the same workload (BASELINE.md's protocol: one
warm-up, median of five); the reference's own compiled count/fill ops
(oracle/_ref) are timed beside it on the same structure where they were
built.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "neighbor pairs/s + HBM GB/s (1M pts 3D); batched/served p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nq: int, dim: int, pairs: int, support_read: bool = True) -> int:
    """Inputs read once, outputs written once: positions + supports of the
    (self-search) cloud, int64 column indices, int32 indptr[N+1] and int64
    rowEntries[N]."""
    return nq * (4 * dim + (4 if support_read else 0)) + pairs * 8 + 4 * (nq + 1) + 8 * nq


def fill_bytes(nq: int, pairs: int, index_bytes: int = 8) -> int:
    """Algorithmic bytes of the fill kernel: int64 row offsets read (it reads
    no positions or supports: it replays the count pass's hit masks) and the
    column indices written."""
    return 8 * (nq + 1) + pairs * index_bytes


def count_bytes(nq: int, dim: int) -> int:
    """Algorithmic bytes of the count kernel: positions + supports of the
    (self-search) cloud read once, int32 counts written."""
    return nq * (4 * dim + 4) + 4 * nq


TRAFFIC_FILES = {"k_fill_small": os.path.join(ROOT, "profiles", "pmc_fill_periodic3d_1M.md"),
                 "k_count_small": os.path.join(ROOT, "profiles", "pmc_count_periodic3d_1M.md")}


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel from its committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (line 'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(TRAFFIC_FILES[kernel]) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except (OSError, KeyError):
        pass
    return None


def copy_probe(torch):
    """A plain device copy of known size (256 MiB read + 256 MiB written)."""
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()


def _cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it (model name, family/model/
    stepping, logical CPUs)."""
    info = {}
    n = 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    n += 1
                elif k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v
    except OSError:
        return "unknown"
    return (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, model "
            f"{info.get('model', '?')}, stepping {info.get('stepping', '?')}, {info.get('cpu MHz', '?')} MHz, "
            f"{n} logical CPUs visible)")


def host_threads() -> int:
    """Threads of the CPU legs: the cores this process may run on, capped by
    OMP_NUM_THREADS where it is set (the GPU box sets it to its CPU share
    per GPU, 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _median_after_warmup(fn, reps: int):
    fn()
    ts = []
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), out


def cpu_baseline(w, reps: int = 5, single=None):
    """The C/OpenMP oracle on this host: the full workload on host_threads()
    threads, one warm-up then the median of `reps` searches; plus one
    single-thread search of `single` (a smaller cloud of the same density, so
    that the single-thread leg stays a few seconds).  Where oracle/_ref is
    built, the reference's own compiled countNeighbors + buildNeighborList
    (with its int32 offsets and the sortIndex gather) are timed the same way
    on the oracle's structure of the same cloud."""
    import oracle
    oracle.build()
    threads = host_threads()

    def one(cloud, nthreads):
        r = oracle.radius_search(cloud.pos, cloud.sup, domain_min=cloud.domain_min, domain_max=cloud.domain_max,
                                 periodic=cloud.periodic, mode=cloud.mode, hash_map_length=cloud.hash_map_length,
                                 want_rows=False, threads=nthreads)
        return int(r.col.shape[0])

    t, P = _median_after_warmup(lambda: one(w, threads), reps)
    s1 = single if single is not None else w
    t0 = time.perf_counter()
    P1 = one(s1, 1)
    t1 = time.perf_counter() - t0
    out = {"value": P / t, "unit": "pairs/s", "cores": threads, "kind": "port",
           "sample": f"full workload ({w.pos.shape[0]} points, {P} pairs): one warm-up, then the median of {reps} "
                     f"searches, {t * 1e3:.0f} ms each; single thread: {s1.pos.shape[0]} points of the same "
                     f"density, {P1} pairs, {t1 * 1e3:.0f} ms",
           "single_thread_value": P1 / t1, "cpu_model": _cpu_model()}
    try:
        from oracle import ref
        if ref.available():
            ref.set_threads(threads)
            args = ref.prepare(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                               mode=w.mode, hash_map_length=w.hash_map_length)
            tr, res = _median_after_warmup(lambda: ref.run_ops(args), reps)
            Pr = int(res[2].shape[0])
            out["reference_ops"] = {
                "value": Pr / tr, "unit": "pairs/s", "cores": threads, "kind": "reference",
                "sample": f"the reference's compiled countNeighbors, int32 offsets, buildNeighborList and "
                          f"sortIndex gather on the full workload's structure ({Pr} pairs; the structure build, "
                          f"Python torch ops in the reference, is not timed): one warm-up, median of {reps}, "
                          f"{tr * 1e3:.0f} ms each",
                "same_pairs_as_port": Pr == P}
    except Exception as e:  # the reference leg is optional
        out["reference_ops"] = {"error": str(e)[:200]}
    return out


def timed_steps(L, step, steps: int):
    """`steps` calls of step() between two device synchronisations, timed by
    the host clock with no instrumentation inside the searches; then the
    same number of calls again with the library's HIP-event phase timing on
    (the events sit between kernels on the search stream and add a few
    microseconds each, so they are kept out of the wall-clock figure).
    Returns (ms per step, mean count phase ms, mean fill kernel ms, mean
    count kernel ms) -- NaN for a kernel that did not run."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.tcr_timing_collect(None, None, None, 0, None)
    L.tcr_timing_enable(1)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    L.tcr_timing_enable(0)
    cms, fms, kms = ((ctypes.c_float * steps)() for _ in range(3))
    got = ctypes.c_int32(0)
    L.tcr_timing_collect(cms, fms, kms, steps, ctypes.byref(got))

    def mean(a):
        v = [x for x in a[:got.value] if x >= 0]
        return statistics.mean(v) if v else float("nan")
    return (t1 - t0) * 1e3 / steps, mean(cms), mean(fms), mean(kms)


def kernel_line(name: str, nbytes: int, avg_ms: float, launches: int) -> dict:
    """One kernel's roofline figures: algorithmic bytes per launch over its
    average HIP-event time, against HBM peak, and its measured traffic."""
    ach = nbytes / (avg_ms / 1e3) / 1e9 if avg_ms == avg_ms else float("nan")
    return {"algorithmic_bytes": nbytes, "avg_ms": avg_ms, "achieved": ach, "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(name), "launches": launches}


def bench_batched(torch, reps: int = 5, scale: float = 1.0):
    """Batched workload: 256 independent clouds (log-uniform 1e3..2e5 points,
    half 2-D half 3-D, per-cloud support and periodicity).  Compares one
    radiusSearch per cloud (plain per-request launches) with the whole batch
    in one launch chain, autotuned and replayed from a captured HIP graph;
    checks that every cloud's CSR output is identical."""
    from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch, synthetic
    from torchcompactradius_amd.autotune import autotune
    from torchcompactradius_amd.search import CloudSpec, SearchPlan, periodic_mask

    ws = synthetic.batched_clouds(scale=scale)
    dev = "cuda"
    data = [(torch.from_numpy(w.pos).to(dev), torch.from_numpy(w.sup).to(dev)) for w in ws]
    doms = [DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max), torch.tensor(w.periodic),
                              w.pos.shape[1]) for w in ws]

    def plain():
        return [radiusSearch(PointCloud(p, h), None, mode=w.mode, domain=d, hashMapLength=w.hash_map_length,
                             algorithm="compact", format="csr") for (p, h), w, d in zip(data, ws, doms)]

    outs = plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    P = sum(int(o.indices.numel()) for o in outs)
    N = sum(w.pos.shape[0] for w in ws)

    specs = [CloudSpec(qpos=p, qsup=h, rpos=None, rsup=None, domain_min=[float(x) for x in w.domain_min],
                       domain_max=[float(x) for x in w.domain_max], periodic=periodic_mask(w.periodic, len(w.periodic)),
                       mode=w.mode, hash_map_length=w.hash_map_length) for (p, h), w in zip(data, ws)]
    res = {}
    for name, dt in (("int64", torch.int64), ("int32", torch.int32)):
        plan = SearchPlan(specs, want_rows=False, index_dtype=dt)
        tuning = autotune(plan)
        r = plan.run()
        plan.capture(r.total + r.total // 32 + 256)
        r = plan.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = plan.replay()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        offs = r.offsets.cpu()
        exact = r.total == P
        q0 = 0
        for o, w in zip(outs, ws):
            n = w.pos.shape[0]
            a, b = int(offs[q0]), int(offs[q0 + n])
            exact = exact and torch.equal(r.col[a:b].to(torch.int64), o.indices) and \
                torch.equal((offs[q0:q0 + n + 1] - a).to(torch.int32), o.indptr.cpu())
            q0 += n
        res[name] = {"ms": t * 1e3, "pairs_per_s": P / t, "bit_exact_vs_plain": bool(exact), "tuning": list(tuning)}
        plan.release_graph()
        del plan, r
    return {"workload": f"batched-{len(ws)}-clouds", "clouds": len(ws), "points": N, "pairs": P,
            "plain_ms": t_plain * 1e3, "plain_pairs_per_s": P / t_plain,
            "graph_int64": res["int64"], "graph_int32": res["int32"],
            "speedup_vs_plain": t_plain / (res["int64"]["ms"] / 1e3)}


def bench_served(torch, n_requests: int = 512, rates=(100000.0, 1500.0), runs: int = 64):
    """Served stream: Poisson arrivals of 3-D search requests from `runs`
    simulated SPH runs (sizes log-uniform 1e3..1e6, every step displaced by
    at most 0.1 h), admitted under a 200 GB HBM budget.  Continuous batching
    (batched launches on 2 launch slots run in order, resident per-run plans) against the same
    stream served by per-request plain launches, one at a time; every
    request's result must be identical in both (fingerprints compared)."""
    from torchcompactradius_amd.serve import (GpuBackend, Scheduler, latency_stats, make_simulations,
                                              poisson_requests)
    out = {"runs": runs, "requests": n_requests, "budget_GB": 200}
    warm = make_simulations(torch, n_runs=8, n_min=1e3, n_max=1e5, seed=99)
    for batched in (True, False):
        be = GpuBackend(torch, warm, slots=2, batched=batched)
        Scheduler(be, max_batch=64 if batched else 1).run(
            poisson_requests(32, 20000.0, [s.pos.shape[0] for s in warm], 98, [s.last_pairs for s in warm]))
    torch.cuda.synchronize()
    for rate in rates:
        rec = {}
        digests = {}
        for name, batched in (("batched", True), ("plain", False)):
            sims = make_simulations(torch, n_runs=runs, seed=5)
            npts = [s.pos.shape[0] for s in sims]
            reqs = poisson_requests(n_requests, rate, npts, 5, [s.last_pairs for s in sims])
            be = GpuBackend(torch, sims, slots=2, batched=batched)
            sch = Scheduler(be, budget_bytes=200e9, max_batch=64 if batched else 1)
            done = sch.run(reqs)
            torch.cuda.synchronize()
            wall = max(r.done for r in done) - min(r.arrival for r in done)
            st = latency_stats(done)
            st.update({"pairs": be.pairs, "pairs_per_s": be.pairs / wall, "wall_s": wall,
                       "points": int(sum(r.n for r in done)),
                       "mean_batch": float(np.mean(sch.batches)), "peak_in_use_GB": sch.peak_in_use / 1e9})
            rec[name] = st
            digests[name] = {r.rid: be.digest(r) for r in done}   # after the timed stream
            del sims, be, done
        rec["bit_exact_alone_vs_mix"] = digests["batched"] == digests["plain"]
        rec["speedup_pairs_per_s"] = rec["batched"]["pairs_per_s"] / rec["plain"]["pairs_per_s"]
        out[f"rate_{int(rate)}_per_s"] = rec
    return out


def _reference_pairs(torch, pos, sup, rows, mn, mx, mode="symmetric"):
    """Exact neighbours of a few query rows by brute force over all points,
    with the reference's pair arithmetic (fp32 difference, fp64 square added
    to an fp32 running sum, fp32 sqrt, strict '<'), in the compact row order
    (hCell stencil offset, then index).  Non-periodic domains."""
    out = []
    for i in rows.tolist():
        d = pos[i:i + 1] - pos                            # fp32 differences
        acc = torch.zeros(pos.shape[0], dtype=torch.float32, device=pos.device)
        for k in range(pos.shape[1]):
            acc = (acc.double() + d[:, k].double() ** 2).float()
        dist = torch.sqrt(acc)
        h = (sup[i] + sup) / 2.0 if mode == "symmetric" else torch.maximum(sup[i].expand_as(sup), sup)
        j = torch.nonzero(dist < h).reshape(-1)
        out.append(j)
    return out


def bench_adaptive(torch, n: int = 2_000_000, reps: int = 3):
    """Adaptive supports (70% Gaussian clusters + 30% uniform, k-NN supports
    with an 8x ratio), symmetric, non-periodic, COO: the multi-level search.
    Full-size checks: the pair set is symmetric, and sampled rows equal a
    brute-force evaluation of the reference predicate in the compact order."""
    from torchcompactradius_amd import synthetic
    from torchcompactradius_amd.search import CloudSpec, run_mlm
    w = synthetic.adaptive_3d(n=n)
    pos = torch.from_numpy(w.pos).cuda()
    sup = torch.from_numpy(w.sup).cuda()
    spec = CloudSpec(qpos=pos, qsup=sup, rpos=None, rsup=None, domain_min=[-1.0] * 3, domain_max=[1.0] * 3,
                     periodic=0, mode="symmetric", hash_map_length=w.hash_map_length)
    r = run_mlm(spec, want_rows=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = run_mlm(spec, want_rows=True)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    P = r.total
    # symmetry of the pair set
    a = torch.sort(r.row * n + r.col).values
    b = torch.sort(r.col * n + r.row).values
    symmetric = bool(torch.equal(a, b))
    del a, b
    # sampled rows against brute force, in the compact order
    g = torch.Generator().manual_seed(3)
    rows = torch.randint(0, n, (64,), generator=g)
    ref = _reference_pairs(torch, pos, sup, rows, -1.0, 1.0)
    from torchcompactradius_amd.util import compute_h
    hcell = compute_h([-1.0] * 3, [1.0] * 3, float(w.sup.max()))
    cells = torch.floor((pos + 1.0) / hcell).to(torch.int64)      # hCell cell of every point
    offs = r.offsets
    ok = True
    for i, j in zip(rows.tolist(), ref):
        got = r.col[int(offs[i]):int(offs[i + 1])]
        # compact order: stencil offset of the candidate's cell, then index
        dc = cells[j] - cells[i]
        o = (dc[:, 0] + 1) * 9 + (dc[:, 1] + 1) * 3 + (dc[:, 2] + 1)
        want = j[torch.argsort(o * n + j)]
        ok = ok and bool(torch.equal(got, want))
    return {"workload": w.name, "points": n, "pairs": P, "ms": t * 1e3, "pairs_per_s": P / t,
            "h_min": w.meta["h_min"], "h_max": w.meta["h_max"], "format": "coo", "mode": "symmetric",
            "checks": {"pair_set_symmetric": symmetric, "sampled_rows_vs_brute_force": bool(ok),
                       "sampled_rows": int(rows.numel())}}


def bench_nonself(torch, n: int = 100, reps: int = 10, warmup: int = 3):
    """Query set != reference set: 1M queries (a jittered grid of another
    seed, shifted by half a grid step) against the 1M-point jittered grid,
    uniform support for ~50 neighbours, gather mode, non-periodic, CSR --
    the search radius() and neighborSearchExisting run.  Check: sampled rows
    equal a brute-force evaluation of the reference predicate in the compact
    order (stencil offset of the reference point's cell seen from the
    query's cell, then index)."""
    from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch, synthetic
    from torchcompactradius_amd.util import compute_h
    wr = synthetic.periodic_3d(n=n, seed=2)
    wq = synthetic.periodic_3d(n=n, seed=7)
    dx = 2.0 / n
    qpos_np = ((wq.pos + np.float32(dx / 2) + np.float32(1.0)) % np.float32(2.0) - np.float32(1.0)).astype(np.float32)
    rpos = torch.from_numpy(wr.pos).cuda()
    rsup = torch.from_numpy(wr.sup).cuda()
    qpos = torch.from_numpy(qpos_np).cuda()
    qsup = torch.from_numpy(wq.sup).cuda()
    dom = DomainDescription(torch.tensor([-1.0] * 3), torch.tensor([1.0] * 3), torch.tensor([False] * 3), 3)

    def step():
        return radiusSearch(PointCloud(qpos, qsup), PointCloud(rpos, rsup), mode="gather", domain=dom,
                            hashMapLength=1 << 20, algorithm="compact", format="csr")
    # warm-up: the pair-count hint, and the caching allocator's blocks for
    # two live results (a call's result stays alive while the next runs)
    for _ in range(warmup):
        out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    P = int(out.indices.shape[0])
    # sampled rows against brute force (gather: dist < h of the query)
    hcell = compute_h([-1.0] * 3, [1.0] * 3, float(wq.sup.max()))
    rcells = torch.floor((rpos + 1.0) / hcell).to(torch.int64)
    g = torch.Generator().manual_seed(5)
    rows = torch.randint(0, qpos.shape[0], (64,), generator=g)
    ok = True
    N = rpos.shape[0]
    indptr = out.indptr.to(torch.int64)
    for i in rows.tolist():
        d = qpos[i:i + 1] - rpos
        acc = torch.zeros(N, dtype=torch.float32, device=rpos.device)
        for k in range(3):
            acc = (acc.double() + d[:, k].double() ** 2).float()
        j = torch.nonzero(torch.sqrt(acc) < qsup[i]).reshape(-1)
        qc = torch.floor((qpos[i] + 1.0) / hcell).to(torch.int64)
        dc = rcells[j] - qc
        o = (dc[:, 0] + 1) * 9 + (dc[:, 1] + 1) * 3 + (dc[:, 2] + 1)
        want = j[torch.argsort(o * N + j)]
        got = out.indices[int(indptr[i]):int(indptr[i + 1])].to(torch.int64)
        ok = ok and bool(torch.equal(got, want))
    return {"workload": f"nonself-3d-{qpos.shape[0]}x{N}", "queries": int(qpos.shape[0]), "references": N,
            "pairs": P, "ms": t * 1e3, "pairs_per_s": P / t, "mode": "gather", "format": "csr",
            "periodic": "none", "checks": {"sampled_rows_vs_brute_force": bool(ok), "sampled_rows": 64}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100, help="grid points per dimension (100 -> 1M points)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (256-cloud) comparison")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream comparison")
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive-support multi-level search")
    ap.add_argument("--no-nonself", action="store_true", help="skip the query != reference search")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--only", choices=("batched", "served", "adaptive", "nonself"),
                    help="run only this comparison leg and print its JSON object (to profile it alone)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the search does not "
              "shard across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.only:
        leg = {"batched": bench_batched, "served": bench_served, "adaptive": bench_adaptive,
               "nonself": bench_nonself}[args.only]
        print(json.dumps({"leg": args.only, **leg(torch)}), flush=True)
        return
    from torchcompactradius_amd import DomainDescription, PointCloud, _lib, radiusSearch, synthetic

    w = synthetic.periodic_3d(n=args.n)
    dev = "cuda"
    pos = torch.from_numpy(w.pos).to(dev)
    sup = torch.from_numpy(w.sup).to(dev)
    dom = DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max),
                            torch.tensor(w.periodic), 3)
    pc = PointCloud(pos, sup)

    def step():
        return radiusSearch(pc, None, mode=w.mode, domain=dom, hashMapLength=w.hash_map_length,
                            algorithm="compact", format="csr")

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    P = int(out.indices.numel())

    L = _lib.load()
    ms, count_ms, fill_ms, ck_ms = timed_steps(L, step, args.steps)
    nq = w.pos.shape[0]
    pairs_per_s = P / (ms / 1e3)
    e2e_bytes = algorithmic_bytes(nq, 3, P)
    kernels = {
        "k_count_small": kernel_line("k_count_small", count_bytes(nq, 3), ck_ms, args.steps),
        "k_fill_small": kernel_line("k_fill_small", fill_bytes(nq, P), fill_ms, args.steps),
    }
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"] if kernels[k]["avg_ms"] == kernels[k]["avg_ms"] else -1)
    if args.copy_probe:
        copy_probe(torch)

    # the same search with the int32-index option (an extension: the
    # reference returns int64), same timing
    def step32():
        return radiusSearch(pc, None, mode=w.mode, domain=dom, hashMapLength=w.hash_map_length,
                            algorithm="compact", format="csr", indexDtype=torch.int32)

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    ms32, _, fill32_ms, _ = timed_steps(L, step32, args.steps)
    fb32 = fill_bytes(nq, P, 4)
    int32_line = {"ms_per_step": ms32, "pairs_per_s": P / (ms32 / 1e3), "fill_kernel_ms": fill32_ms,
                  "fill_GB/s": fb32 / (fill32_ms / 1e3) / 1e9, "fill_algorithmic_bytes": fb32}

    dk = kernels[dominant]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": dk["achieved"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["frac"], "traffic": dk["traffic"],
                "algorithmic_bytes": dk["algorithmic_bytes"], "avg_ms": dk["avg_ms"], "launches": dk["launches"],
                "end_to_end_frac": e2e_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "kernels": kernels}

    line = {
        "metric": METRIC,
        "value": pairs_per_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 positions, f64 periodic distance terms, int64 indices",
        "data": "synthetic (seeded jittered grid)",
        "config": {"workload": w.name, "points": nq, "dim": 3, "periodic": "xyz", "mode": w.mode,
                   "format": "csr", "hash_map_length": w.hash_map_length, "pairs": P},
        "roofline": roofline,
        "end_to_end": {"algorithmic_bytes": e2e_bytes, "GB/s": e2e_bytes / (ms / 1e3) / 1e9,
                       "frac": e2e_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "count_phase_ms": count_ms,
                       "fill_kernel_ms": fill_ms, "points_per_s": nq / (ms / 1e3)},
        "int32_indices": int32_line,
    }
    if not args.no_batched:
        line["batched"] = bench_batched(torch)
    if not args.no_served:
        line["served"] = bench_served(torch)
    if not args.no_adaptive:
        line["adaptive"] = bench_adaptive(torch)
    if not args.no_nonself:
        line["nonself"] = bench_nonself(torch)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, single=synthetic.periodic_3d(n=max(10, args.n // 2)))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
