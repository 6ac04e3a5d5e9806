"""GPU tests of the serving loop: every request served in the continuous-
batching mix equals the oracle on its own positions and equals the same
request served alone (per-request plain launches)."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd.serve import GpuBackend, Scheduler, make_simulations, poisson_requests

pytestmark = pytest.mark.gpu


def _serve(batched, budget, rate, n_req=48, slots=2, max_batch=16, resident_graph=False):
    sims = make_simulations(torch, n_runs=8, n_min=1e3, n_max=2e4, seed=11, neighbors=30.0)
    reqs = poisson_requests(n_req, rate, [s.pos.shape[0] for s in sims], 3, [s.last_pairs for s in sims])
    be = GpuBackend(torch, sims, slots=slots, batched=batched, keep_inputs=True, resident_graph=resident_graph)
    sch = Scheduler(be, budget_bytes=budget, max_batch=max_batch if batched else 1)
    done = sch.run(reqs)
    torch.cuda.synchronize()
    return sims, sorted(done, key=lambda r: r.rid), sch, be


@pytest.mark.parametrize("budget", [4e9, 6e7])
def test_served_mix_matches_oracle_and_alone(budget):
    print(f"case: serve budget={budget}", flush=True)
    sims, mix, sch, be = _serve(True, budget, rate=50000.0)
    assert sch.in_use == 0 and sch.peak_in_use <= budget
    assert max(sch.batches) > 1
    _, alone, _, be_alone = _serve(False, budget, rate=50000.0)
    for a, b in zip(mix, alone):
        assert a.rid == b.rid and a.run == b.run and a.step == b.step
        ia, ca = be.result(a)
        ib, cb = be_alone.result(b)
        assert torch.equal(ia, ib) and torch.equal(ca.to(torch.int64), cb.to(torch.int64))
        assert be.digest(a) == be_alone.digest(b)
    for r in mix[::5]:
        s = sims[r.run]
        pos = r.inputs.cpu().numpy()
        o = oracle.radius_search(pos, s.sup.cpu().numpy(), domain_min=s.domain_min, domain_max=s.domain_max,
                                 periodic=s.periodic, mode=s.mode, hash_map_length=s.hash_map_length,
                                 want_rows=False)
        ip, col = be.result(r)
        np.testing.assert_array_equal(ip.cpu().numpy(), o.indptr)
        np.testing.assert_array_equal(col.cpu().numpy(), o.col)


@pytest.mark.parametrize("resident_graph", [False, True])
def test_served_alone_path_reuses_resident_plan(resident_graph):
    """At a low arrival rate requests are served alone through their
    simulation's resident plan (direct launch or captured graph replay)."""
    print(f"case: serve low rate graph={resident_graph}", flush=True)
    sims, done, sch, be = _serve(True, 4e9, rate=200.0, n_req=16, resident_graph=resident_graph)
    assert sum(1 for b in sch.batches if b == 1) >= len(sch.batches) // 2
    assert be.alone >= 1 and any(s.plan is not None for s in sims)
    for r in done:
        s = sims[r.run]
        o = oracle.radius_search(r.inputs.cpu().numpy(), s.sup.cpu().numpy(), domain_min=s.domain_min,
                                 domain_max=s.domain_max, periodic=s.periodic, mode=s.mode,
                                 hash_map_length=s.hash_map_length, want_rows=False)
        ip, col = be.result(r)
        np.testing.assert_array_equal(col.cpu().numpy(), o.col)
        np.testing.assert_array_equal(ip.cpu().numpy(), o.indptr)
