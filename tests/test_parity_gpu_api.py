"""GPU parity of the other entry points: batched multi-cloud searches through
one C-ABI call, radius() with batch ids, and the reference-layout native ops
(countNeighbors / buildNeighborList / computeHashIndices) -- all against the
CPU oracle, bit-exact on indices and order."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import (DomainDescription, PointCloud, coo_to_csr, native_ops, neighborSearchExisting,
                                    radius, radiusSearch, synthetic)
from torchcompactradius_amd.search import CloudSpec, run_search

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _cloud(rng, n, dim, per, h):
    pos = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
    sup = np.full(n, h, np.float32) * rng.uniform(0.8, 1.2, n).astype(np.float32)
    return pos, sup, per


@pytest.mark.parametrize("seed", [0, 1])
def test_batched_clouds_one_launch(seed):
    """Clouds of uneven size, dimension, mode and periodicity (self and
    non-self) in one count/fill call; each cloud's rows equal its own
    search."""
    rng = np.random.default_rng(300 + seed)
    modes = ["gather", "scatter", "symmetric", "superSymmetric"]
    specs, refs = [], []
    for k in range(9):
        dim = [2, 3, 1][k % 3]
        n = int(rng.integers(0 if k == 4 else 50, 3000))
        per = [bool(rng.integers(0, 2)) for _ in range(dim)]
        h = {1: 0.02, 2: 0.08, 3: 0.2}[dim]
        q, hq, _ = _cloud(rng, n, dim, per, h)
        selfs = k % 2 == 0
        r, hr = (None, None) if selfs else _cloud(rng, int(rng.integers(1, 2500)), dim, per, h)[:2]
        mode = modes[k % 4]
        L = int(rng.choice([1, 97, 4096]))
        mn, mx = [-1.0] * dim, [1.0] * dim
        pm = sum(1 << d for d in range(dim) if per[d])
        specs.append(CloudSpec(qpos=_t(q), qsup=_t(hq), rpos=_t(r), rsup=_t(hr), domain_min=mn, domain_max=mx,
                               periodic=pm, mode=mode, hash_map_length=L))
        refs.append(oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                                         hash_map_length=L))
    print(f"case: batched seed={seed}", flush=True)
    res = run_search(specs, want_rows=True)
    torch.cuda.synchronize()
    offs = res.offsets.cpu().numpy()
    row, col = res.row.cpu().numpy(), res.col.cpu().numpy()
    q0 = 0
    for s, o in zip(specs, refs):
        nq = s.qpos.shape[0]
        a, b = offs[q0], offs[q0 + nq]
        np.testing.assert_array_equal(row[a:b], o.row)
        np.testing.assert_array_equal(col[a:b], o.col)
        np.testing.assert_array_equal(np.diff(offs[q0:q0 + nq + 1]), o.counts)
        q0 += nq
    assert offs[-1] == sum(o.col.shape[0] for o in refs) == res.total


def test_batched_empty_reference_and_non_finite():
    """A batch mixing a query cloud whose reference set is empty, a
    self-search cloud with NaN / inf coordinates, and a regular cloud: each
    cloud's rows equal its own oracle search."""
    rng = np.random.default_rng(77)
    specs, refs = [], []
    cases = []
    q, hq, _ = _cloud(rng, 300, 2, [True, False], 0.08)
    cases.append((q, hq, np.zeros((0, 2), np.float32), np.zeros(0, np.float32), [True, False], "gather"))
    q, hq, _ = _cloud(rng, 2000, 3, [True, True, False], 0.2)
    q[7, 1] = np.nan
    q[100, 0] = np.inf
    q[200, :] = np.nan
    cases.append((q, hq, None, None, [True, True, False], "symmetric"))
    q, hq, _ = _cloud(rng, 1500, 2, [False, True], 0.08)
    cases.append((q, hq, None, None, [False, True], "scatter"))
    for q, hq, r, hr, per, mode in cases:
        dim = q.shape[1]
        mn, mx = [-1.0] * dim, [1.0] * dim
        pm = sum(1 << d for d in range(dim) if per[d])
        specs.append(CloudSpec(qpos=_t(q), qsup=_t(hq), rpos=_t(r), rsup=_t(hr), domain_min=mn, domain_max=mx,
                               periodic=pm, mode=mode, hash_map_length=4096))
        refs.append(oracle.radius_search(q, hq, r, hr, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                                         hash_map_length=4096))
    print("case: batched empty reference + non-finite", flush=True)
    res = run_search(specs, want_rows=True)
    torch.cuda.synchronize()
    offs = res.offsets.cpu().numpy()
    row, col = res.row.cpu().numpy(), res.col.cpu().numpy()
    q0 = 0
    for s, o in zip(specs, refs):
        nq = s.qpos.shape[0]
        a, b = offs[q0], offs[q0 + nq]
        np.testing.assert_array_equal(row[a:b], o.row)
        np.testing.assert_array_equal(col[a:b], o.col)
        q0 += nq
    assert offs[-1] == sum(o.col.shape[0] for o in refs) == res.total


def test_radius_with_batch_ids():
    rng = np.random.default_rng(77)
    n, m = 2000, 1500
    x = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    y = rng.uniform(0, 1, (m, 3)).astype(np.float32)
    bx = np.sort(rng.integers(0, 4, n))
    by = np.sort(rng.integers(0, 4, m))
    h = 0.12
    print("case: radius batch", flush=True)
    i, j = radius(_t(x), _t(y), h, batch_x=_t(bx), batch_y=_t(by))
    torch.cuda.synchronize()
    got = set(zip(i.cpu().tolist(), j.cpu().tolist()))
    want = set()
    for b in range(4):
        qi, ri = np.nonzero(bx == b)[0], np.nonzero(by == b)[0]
        o = oracle.radius_search(x[qi], np.full(len(qi), h, np.float32), y[ri], np.full(len(ri), h, np.float32),
                                 mode="gather")
        want |= set(zip(ri[o.col].tolist(), qi[o.row].tolist()))
    assert got == want


@pytest.mark.parametrize("algorithm", ["naive", "small", "mlm"])
def test_radius_algorithms(algorithm):
    """radius() with radiusSearch's other algorithms, with and without batch
    ids: the flipped pairs of each algorithm's own search (for 'naive', its
    '<=' and default-domain semantics)."""
    rng = np.random.default_rng(78 + len(algorithm))
    n, m = 900, 700
    x = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    y = rng.uniform(0, 1, (m, 2)).astype(np.float32)
    hx = rng.uniform(0.04, 0.07, n).astype(np.float32)
    hy = rng.uniform(0.04, 0.07, m).astype(np.float32)
    bx = np.sort(rng.integers(0, 3, n))
    by = np.sort(rng.integers(0, 3, m))
    mode = "symmetric"
    print(f"case: radius algorithm={algorithm}", flush=True)

    def want_pairs(qi, ri):
        if algorithm == "naive":
            ii, jj = oracle.radius_search_naive(x[qi], hx[qi], y[ri], hy[ri], mode=mode)
        elif algorithm == "small":
            ii, jj = oracle.radius_search_small(x[qi], hx[qi], y[ri], hy[ri], mode=mode)
        else:
            o = oracle.radius_search(x[qi], hx[qi], y[ri], hy[ri], mode=mode)
            ii, jj = o.row, o.col
        return set(zip(ri[jj].tolist(), qi[ii].tolist()))

    i, j = radius(_t(x), _t(y), (_t(hx), _t(hy)), mode=mode, algorithm=algorithm)
    assert set(zip(i.cpu().tolist(), j.cpu().tolist())) == want_pairs(np.arange(n), np.arange(m))
    i, j = radius(_t(x), _t(y), (_t(hx), _t(hy)), batch_x=_t(bx), batch_y=_t(by), mode=mode, algorithm=algorithm)
    want = set()
    for b in range(3):
        want |= want_pairs(np.nonzero(bx == b)[0], np.nonzero(by == b)[0])
    assert set(zip(i.cpu().tolist(), j.cpu().tolist())) == want


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
def test_native_ops_on_exported_structure(mode):
    """searchNeighbors_cpp's sequence (compactHashing/cppSearch.py:60-98) on
    the exported structure: countNeighbors -> int32 offsets ->
    buildNeighborList -> j = sortIndex[j]."""
    w = synthetic.periodic_3d(n=22)
    rng = np.random.default_rng(8)
    sup = (w.sup * rng.uniform(0.7, 1.3, w.sup.shape[0])).astype(np.float32)
    dom = DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max), torch.tensor(w.periodic), 3)
    print(f"case: native ops mode={mode}", flush=True)
    _, hm = radiusSearch(PointCloud(_t(w.pos), _t(sup)), None, mode=mode, domain=dom, hashMapLength=1009,
                         algorithm="compact", format="coo", returnStructure=True)
    q = _t(oracle.wrap(w.pos, w.domain_min, w.domain_max, w.periodic))
    qs = _t(sup)
    per = hm.periodicity
    counts = native_ops.countNeighbors(q, qs, 1, hm.sortedPositions, hm.sortedSupports, hm.hashTable,
                                       hm.hashMapLength, hm.numCells, hm.sortedCellTable, hm.qMin,
                                       float(hm.hCell.item()), hm.maxD, hm.minD, per, mode, False)
    offsets = torch.cat((torch.zeros(1, dtype=torch.int32, device=DEV),
                         torch.cumsum(counts, 0).to(torch.int32)))[:-1]
    total = int(offsets[-1].item() + counts[-1].item())
    i, j = native_ops.buildNeighborList(counts, offsets, total, q, qs, 1, hm.sortedPositions, hm.sortedSupports,
                                       hm.hashTable, hm.hashMapLength, hm.numCells, hm.sortedCellTable, hm.qMin,
                                       float(hm.hCell.item()), hm.maxD, hm.minD, per, mode, False)
    j = hm.sortIndex[j]
    torch.cuda.synchronize()
    o = oracle.radius_search(w.pos, sup, domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic,
                             mode=mode, hash_map_length=1009)
    np.testing.assert_array_equal(counts.cpu().numpy(), o.counts)
    np.testing.assert_array_equal(i.cpu().numpy(), o.row)
    np.testing.assert_array_equal(j.cpu().numpy(), o.col)


@pytest.mark.parametrize("mode", ["gather", "scatter", "symmetric", "superSymmetric"])
@pytest.mark.parametrize("self_search", [False, True])
def test_search_existing_matches_radius_search(mode, self_search):
    """neighborSearchExisting (neighborhood.py:28-132) on the structure of a
    radiusSearch gives that search's result for its (once-wrapped) queries,
    in COO and CSR, and the oracle's pairs."""
    w = synthetic.periodic_3d(n=18)
    rng = np.random.default_rng(31)
    sup = (w.sup * rng.uniform(0.7, 1.3, w.sup.shape[0])).astype(np.float32)
    dom = DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max), torch.tensor(w.periodic), 3)
    if self_search:
        qp, qs, ref = w.pos, sup, None
    else:
        qp = rng.uniform(-1.2, 1.2, (3000, 3)).astype(np.float32)   # some outside the domain: wrapped
        qs = (w.sup[0] * rng.uniform(0.6, 1.4, 3000)).astype(np.float32)
        ref = PointCloud(_t(w.pos), _t(sup))
    print(f"case: search existing mode={mode} self={self_search}", flush=True)
    out, hm = radiusSearch(PointCloud(_t(qp), _t(qs)), ref, mode=mode, domain=dom, hashMapLength=1009,
                           algorithm="compact", format="coo", returnStructure=True)
    qw = PointCloud(_t(oracle.wrap(qp, w.domain_min, w.domain_max, w.periodic)), _t(qs))
    got = neighborSearchExisting(qw, hm, mode=mode)
    torch.testing.assert_close(got.row, out.row, rtol=0, atol=0)
    torch.testing.assert_close(got.col, out.col, rtol=0, atol=0)
    assert (got.numRows, got.numCols) == (out.numRows, out.numCols)
    csr = neighborSearchExisting(qw, hm, mode=mode, format="csr")
    want = coo_to_csr(out, isSorted=True)
    for a, b in ((csr.indices, want.indices), (csr.indptr, want.indptr), (csr.rowEntries, want.rowEntries)):
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    o = oracle.radius_search(qp, qs, None if self_search else w.pos, None if self_search else sup,
                             domain_min=w.domain_min, domain_max=w.domain_max, periodic=w.periodic, mode=mode,
                             hash_map_length=1009)
    np.testing.assert_array_equal(got.row.cpu().numpy(), o.row)
    np.testing.assert_array_equal(got.col.cpu().numpy(), o.col)


def test_search_existing_fixed_support():
    """A float support (on the query cloud, or the structure's fixedSupport)
    acts as that support on every point: the structure of a supportOverride
    search then reproduces that search."""
    w = synthetic.periodic_3d(n=16)
    rng = np.random.default_rng(32)
    qp = rng.uniform(-1.0, 1.0, (2000, 3)).astype(np.float32)
    h = float(w.sup[0])
    dom = DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max), torch.tensor(w.periodic), 3)
    print("case: search existing fixed support", flush=True)
    out, hm = radiusSearch(PointCloud(_t(qp), None), PointCloud(_t(w.pos), None), supportOverride=h,
                           mode="symmetric", domain=dom, hashMapLength=4096, algorithm="compact",
                           returnStructure=True)
    qw = _t(oracle.wrap(qp, w.domain_min, w.domain_max, w.periodic))
    for got in (neighborSearchExisting(PointCloud(qw, h), hm),
                neighborSearchExisting(PointCloud(qw, None), hm._replace(fixedSupport=h))):
        torch.testing.assert_close(got.row, out.row, rtol=0, atol=0)
        torch.testing.assert_close(got.col, out.col, rtol=0, atol=0)


def test_search_existing_empty_query():
    w = synthetic.periodic_3d(n=10)
    dom = DomainDescription(torch.tensor(w.domain_min), torch.tensor(w.domain_max), torch.tensor(w.periodic), 3)
    _, hm = radiusSearch(PointCloud(_t(w.pos), _t(w.sup)), None, domain=dom, algorithm="compact",
                         returnStructure=True)
    got = neighborSearchExisting(PointCloud(torch.zeros((0, 3), device=DEV), torch.zeros(0, device=DEV)), hm,
                                 mode="gather")
    assert got.row.numel() == 0 and got.col.numel() == 0 and got.numRows == 0


def test_native_ops_errors():
    x = torch.zeros((4, 3), device=DEV)
    with pytest.raises(RuntimeError):
        native_ops.countNeighbors(x, None, 1, x, None, torch.zeros((8, 2), dtype=torch.int32, device=DEV), 8,
                                  torch.ones(3, dtype=torch.int32, device=DEV),
                                  torch.zeros((1, 3), dtype=torch.int32, device=DEV), None, 0.1,
                                  torch.ones(3, device=DEV), torch.zeros(3, device=DEV),
                                  torch.zeros(3, dtype=torch.bool, device=DEV), "gather", False)
    with pytest.raises(RuntimeError):
        native_ops.computeHashIndices(torch.zeros((3, 4), dtype=torch.int32, device=DEV), 16)


def test_compute_hash_indices():
    rng = np.random.default_rng(21)
    for dim in (1, 2, 3):
        cells = rng.integers(-2**31, 2**31 - 1, (500, dim), dtype=np.int64).astype(np.int32)
        cells[:100] = rng.integers(-5, 50, (100, dim)).astype(np.int32)
        for L in (1, 7, 4096, 2**20):
            print(f"case: hash dim={dim} L={L}", flush=True)
            got = native_ops.computeHashIndices(_t(cells), L).cpu().numpy()
            want = np.array([oracle.hash_index(c, L) for c in cells], dtype=np.int64)
            np.testing.assert_array_equal(got.astype(np.int64), want)


def test_speculative_capacity_paths():
    """Repeated searches of one shape launch count and fill back to back into
    outputs sized from the previous total; a denser cloud of the same shape
    overflows that guess and takes the separate fill, a sparser one fits."""
    rng = np.random.default_rng(31)
    q = rng.uniform(-1, 1, (3000, 3)).astype(np.float32)
    mn, mx = [-1.0] * 3, [1.0] * 3
    for h in (0.15, 0.15, 0.25, 0.1, 0.1):
        sup = np.full(3000, h, np.float32)
        print(f"case: speculative h={h}", flush=True)
        dom = DomainDescription(torch.tensor(mn), torch.tensor(mx), torch.tensor([True, False, True]), 3)
        g = radiusSearch(PointCloud(_t(q), _t(sup)), None, mode="gather", domain=dom, hashMapLength=4096,
                         algorithm="compact", format="coo")
        torch.cuda.synchronize()
        o = oracle.radius_search(q, sup, domain_min=mn, domain_max=mx, periodic=[True, False, True], mode="gather",
                                 hash_map_length=4096)
        np.testing.assert_array_equal(g.row.cpu().numpy(), o.row)
        np.testing.assert_array_equal(g.col.cpu().numpy(), o.col)


@pytest.mark.parametrize("capacity", [0, 10, None])
def test_csr_arrays_written_by_the_launch(capacity):
    """The CSR launch (indptr and rowEntries written in the fill phase)
    against the oracle on a batch of two clouds: with a capacity guess that
    fits, one that is too small (separate fill) and the shape's hint."""
    from torchcompactradius_amd.search import SearchPlan
    rng = np.random.default_rng(77)
    clouds = [(rng.uniform(-1, 1, (9000, 3)).astype(np.float32), np.full(9000, 0.12, np.float32),
               [True, True, True], "gather"),
              (rng.uniform(-1, 1, (700, 2)).astype(np.float32), rng.uniform(0.05, 0.2, 700).astype(np.float32),
               [False, True], "symmetric")]
    specs, refs = _specs_and_refs(rng, clouds)
    plan = SearchPlan(specs, want_rows=False)
    P = sum(int(o.col.shape[0]) for o in refs)
    if capacity is None:
        plan.run()  # leaves the pair count of this shape as the hint
    cap = None if capacity is None else (P + 4096 if capacity == 0 else capacity)
    print(f"case: csr launch capacity={cap}", flush=True)
    res = plan.run(capacity=cap, csr=True)
    torch.cuda.synchronize()
    assert res.total == P
    _check_batch(res, specs, refs)
    offs = res.offsets.cpu().numpy()
    assert res.indptr.dtype == torch.int32 and res.row_entries.dtype == torch.int64
    np.testing.assert_array_equal(res.indptr.cpu().numpy(), offs.astype(np.int32))
    counts = np.concatenate([o.counts for o in refs]).astype(np.int64)
    np.testing.assert_array_equal(res.row_entries.cpu().numpy(), counts)


def _specs_and_refs(rng, clouds):
    specs, refs = [], []
    for q, h, per, mode in clouds:
        dim = q.shape[1]
        mn, mx = [-1.0] * dim, [1.0] * dim
        pm = sum(1 << d for d in range(dim) if per[d])
        specs.append(CloudSpec(qpos=_t(q), qsup=_t(h), rpos=None, rsup=None, domain_min=mn, domain_max=mx,
                               periodic=pm, mode=mode, hash_map_length=4096))
        refs.append(oracle.radius_search(q, h, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                                         hash_map_length=4096, want_rows=False))
    return specs, refs


def _check_batch(res, specs, refs):
    offs = res.offsets.cpu().numpy()
    col = res.col.cpu().numpy().astype(np.int64)
    q0 = 0
    for s, o in zip(specs, refs):
        nq = s.qpos.shape[0]
        np.testing.assert_array_equal(col[offs[q0]:offs[q0 + nq]], o.col)
        np.testing.assert_array_equal(np.diff(offs[q0:q0 + nq + 1]), o.counts)
        q0 += nq


def test_graph_replay_int32_and_tuning():
    """A captured graph replays the whole chain; positions rewritten in place
    between replays are searched anew; int32 outputs and non-default launch
    shapes give the same pairs."""
    from torchcompactradius_amd.search import SearchPlan
    rng = np.random.default_rng(55)
    base = []
    for k in range(5):
        dim = 2 + k % 2
        n = int(rng.integers(500, 4000))
        q = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
        h = np.full(n, 0.12 if dim == 2 else 0.25, np.float32)
        base.append([q, h, [bool(rng.integers(0, 2)) for _ in range(dim)], ["gather", "symmetric"][k % 2]])
    specs, refs = _specs_and_refs(rng, base)
    for tuning in (None, (8, 64), (1024, 4096), (0, 0, 64, 4), (2048, 512, 128, 2)):
        for dt in (torch.int64, torch.int32):
            print(f"case: graph tuning={tuning} dtype={dt}", flush=True)
            plan = SearchPlan(specs, want_rows=False, index_dtype=dt)
            plan.tuning = tuning
            r = plan.run()
            assert r.col.dtype == dt
            _check_batch(r, specs, refs)
            plan.capture(r.total + 64)
            _check_batch(plan.replay(), specs, refs)
            # move every point a little, rewrite the plan's buffer in place, replay
            moved = []
            for q, h, per, mode in base:
                q2 = (q + rng.uniform(-0.01, 0.01, q.shape)).astype(np.float32)
                moved.append([q2, h, per, mode])
            specs2, refs2 = _specs_and_refs(rng, moved)
            plan.qpos.copy_(torch.cat([s.qpos.reshape(-1) for s in specs2]))
            _check_batch(plan.replay(), specs2, refs2)
            plan.release_graph()


def test_threads_with_different_launch_shapes():
    """Two host threads run differently tuned plans of the same batch at
    the same time (the library keeps one launch shape per thread): both give
    the oracle's pairs."""
    import threading
    from torchcompactradius_amd.search import SearchPlan
    rng = np.random.default_rng(56)
    base = []
    for k in range(4):
        dim = 2 + k % 2
        n = int(rng.integers(2000, 6000))
        q = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
        h = np.full(n, 0.1 if dim == 2 else 0.2, np.float32)
        base.append([q, h, [bool(rng.integers(0, 2)) for _ in range(dim)], "gather"])
    specs, refs = _specs_and_refs(rng, base)
    print("case: two threads, two launch shapes", flush=True)
    results, errors = {}, []

    def work(name, tuning):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                plan = SearchPlan(specs, want_rows=False)
                plan.tuning = tuning
                for _ in range(3):
                    r = plan.run()
                stream.synchronize()
                results[name] = r
        except Exception as e:  # reported by the main thread
            errors.append(e)

    ts = [threading.Thread(target=work, args=("a", (8, 64, 64, 4))),
          threading.Thread(target=work, args=("b", (2048, 0, 256, 1)))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for name in ("a", "b"):
        _check_batch(results[name], specs, refs)
