"""GPU parity at the configured sizes of the batched, adaptive and served
workloads:

* batched: 256 clouds (log-uniform 1e3..2e5 points, half 2-D half 3-D,
  random periodicity) in one batched launch and in a captured-graph replay,
  every cloud's CSR against the oracle;
* adaptive: 2M points with k-NN supports (8x ratio), symmetric, through the
  multi-level search -- the pair set symmetric, rows consistent with the
  offsets, int32 and int64 outputs identical, and 64 sampled rows against a
  CPU brute force of the reference's pair arithmetic in the compact order
  (the single-level oracle needs hours on this cloud: its hMax cells hold
  thousands of points);
* served: the 64-run stream (1e3..1e6 points) through continuous batching,
  every request identical to the same request served alone, and requests of
  every size class against the oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import synthetic
from torchcompactradius_amd.search import CloudSpec, SearchPlan, run_mlm

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
DEV = "cuda"


def _spec(w):
    mask = sum(1 << d for d, p in enumerate(w.periodic) if p)
    return CloudSpec(qpos=torch.from_numpy(w.pos).to(DEV), qsup=torch.from_numpy(w.sup).to(DEV), rpos=None,
                     rsup=None, domain_min=[float(x) for x in w.domain_min],
                     domain_max=[float(x) for x in w.domain_max], periodic=mask, mode=w.mode,
                     hash_map_length=w.hash_map_length)


def test_batched_256_clouds_vs_oracle():
    ws = synthetic.batched_clouds()
    print(f"case: batched {len(ws)} clouds, {sum(w.pos.shape[0] for w in ws)} points", flush=True)
    plan = SearchPlan([_spec(w) for w in ws], want_rows=False)
    res = plan.run()
    torch.cuda.synchronize()
    offs = res.offsets.cpu().numpy()
    col = res.col.cpu().numpy()
    plan32 = SearchPlan([_spec(w) for w in ws], want_rows=False, index_dtype=torch.int32)
    plan32.capture(res.total + 1024)
    rep = plan32.replay()
    rep2 = plan32.replay()  # a second replay of the same graph
    torch.cuda.synchronize()
    assert rep.total == res.total and rep2.total == res.total
    assert torch.equal(rep2.col.to(torch.int64).cpu(), res.col.cpu())
    assert torch.equal(rep.offsets.cpu(), res.offsets.cpu())
    q0 = 0
    for k, w in enumerate(ws):
        n = w.pos.shape[0]
        o = oracle.radius_search(w.pos, w.sup, domain_min=w.domain_min, domain_max=w.domain_max,
                                 periodic=w.periodic, mode=w.mode, hash_map_length=w.hash_map_length,
                                 want_rows=False)
        ip = offs[q0:q0 + n + 1] - offs[q0]
        np.testing.assert_array_equal(ip, o.indptr.astype(np.int64), err_msg=f"cloud {k}")
        np.testing.assert_array_equal(col[offs[q0]:offs[q0 + n]], o.col, err_msg=f"cloud {k}")
        q0 += n
    plan32.release_graph()


def _brute_rows(pos, sup, rows, mode="symmetric"):
    """Neighbours of some rows by exhaustive search with the reference's
    arithmetic in a non-periodic domain: fp32 differences, each square added
    in fp64 to an fp32 running sum, fp32 sqrt, strict '<'
    (cpp/common.h:106-129, cpp/neighborhoodDynamic.h:63-70)."""
    out = []
    for i in rows:
        d = (pos[i] - pos).astype(np.float32)
        acc = np.zeros(pos.shape[0], np.float32)
        for k in range(pos.shape[1]):
            acc = (acc.astype(np.float64) + d[:, k].astype(np.float64) ** 2).astype(np.float32)
        dist = np.sqrt(acc)  # correctly rounded fp32 square root
        h = ((sup[i] + sup) / np.float32(2)).astype(np.float32) if mode == "symmetric" else np.maximum(sup[i], sup)
        out.append(np.nonzero(dist < h)[0])
    return out


def test_adaptive_2m_multilevel():
    w = synthetic.adaptive_3d(n=2_000_000)
    n = w.pos.shape[0]
    print(f"case: adaptive {n} points, multi-level search", flush=True)
    spec = CloudSpec(qpos=torch.from_numpy(w.pos).to(DEV), qsup=torch.from_numpy(w.sup).to(DEV), rpos=None,
                     rsup=None, domain_min=[-1.0] * 3, domain_max=[1.0] * 3, periodic=0, mode="symmetric",
                     hash_map_length=w.hash_map_length)
    r = run_mlm(spec, want_rows=True)
    r2 = run_mlm(spec, want_rows=False, index_dtype=torch.int32)
    torch.cuda.synchronize()
    assert r.total > 25 * n
    assert torch.equal(r2.col.to(torch.int64), r.col) and torch.equal(r2.offsets, r.offsets)
    del r2
    rows_of = torch.repeat_interleave(torch.arange(n, device=DEV), r.offsets[1:] - r.offsets[:-1])
    assert torch.equal(rows_of, r.row)
    del rows_of
    key = torch.sort(r.row * n + r.col).values
    assert torch.equal(key, torch.sort(r.col * n + r.row).values)  # the pair set is symmetric
    del key
    rng = np.random.default_rng(3)
    rows = rng.integers(0, n, 64)
    offs = r.offsets.cpu().numpy()
    from torchcompactradius_amd.util import compute_h
    hcell = np.float32(compute_h([-1.0] * 3, [1.0] * 3, float(w.sup.max())))
    cells = np.floor((w.pos + np.float32(1.0)) / hcell).astype(np.int64)
    for i, j in zip(rows, _brute_rows(w.pos, w.sup, rows)):
        got = r.col[int(offs[i]):int(offs[i + 1])].cpu().numpy()
        dc = cells[j] - cells[i]
        o = (dc[:, 0] + 1) * 9 + (dc[:, 1] + 1) * 3 + (dc[:, 2] + 1)
        want = j[np.lexsort((j, o))]  # stencil offset of the candidate's cell, then index
        np.testing.assert_array_equal(got, want, err_msg=f"row {i}")


def test_served_64_runs_full_stream():
    from torchcompactradius_amd.serve import GpuBackend, Scheduler, make_simulations, poisson_requests
    print("case: served 64 runs, 256 requests, 1e3..1e6 points", flush=True)
    results = {}
    for batched in (True, False):
        sims = make_simulations(torch, n_runs=64, seed=5)
        reqs = poisson_requests(256, 100000.0, [s.pos.shape[0] for s in sims], 5, [s.last_pairs for s in sims])
        be = GpuBackend(torch, sims, slots=2, batched=batched, keep_inputs=batched)
        sch = Scheduler(be, budget_bytes=200e9, max_batch=64 if batched else 1)
        done = sorted(sch.run(reqs), key=lambda q: q.rid)
        torch.cuda.synchronize()
        results[batched] = (sims, done, be, sch)
    sims, mix, be, sch = results[True]
    _, alone, be_alone, _ = results[False]
    assert max(sch.batches) > 1
    assert len(mix) == len(alone) == 256
    for a, b in zip(mix, alone):
        assert (a.rid, a.run, a.step) == (b.rid, b.run, b.step)
        ia, ca = be.result(a)
        ib, cb = be_alone.result(b)
        assert torch.equal(ia, ib) and torch.equal(ca, cb), f"request {a.rid}"
    # oracle checks: the largest request, the smallest, and a spread between
    by_size = sorted(mix, key=lambda q: q.n)
    picks = {by_size[0].rid: by_size[0], by_size[-1].rid: by_size[-1]}
    for q in by_size[:: max(1, len(by_size) // 6)]:
        picks[q.rid] = q
    for q in picks.values():
        s = sims[q.run]
        o = oracle.radius_search(q.inputs.cpu().numpy(), s.sup.cpu().numpy(), domain_min=s.domain_min,
                                 domain_max=s.domain_max, periodic=s.periodic, mode=s.mode,
                                 hash_map_length=s.hash_map_length, want_rows=False)
        ip, col = be.result(q)
        np.testing.assert_array_equal(ip.cpu().numpy(), o.indptr, err_msg=f"request {q.rid} ({q.n} points)")
        np.testing.assert_array_equal(col.cpu().numpy(), o.col, err_msg=f"request {q.rid}")
    assert by_size[-1].n > 500_000
