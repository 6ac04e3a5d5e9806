"""Pins oracle/mlm_ref.py, the restatement of the reference's own
algorithm='mlm' (py/multiLevelMemory/wrapper.py:145-308).

* Its arithmetic helpers against the library the reference calls (torch's
  fp32 log2/floor/pow and int32/int64 bit operations, written as the
  reference writes them) and against the pinned compact oracle (wrap,
  pair distance, hash).
* Whole rows against a closed form that bypasses the cell runs, the merged
  cell list and the hash map: for query i at level l (its support's), a
  sorted point s is visited once for each stencil offset o whose int32
  Morton code equals s's level-l code -- and whose decoded code hashes to
  the offset cell's bucket -- in ascending sorted position; the predicate
  then keeps it.  In 2-D, and in 1-D and 3-D on one level, that visit set
  holds every pair within support, so rows are also checked against an
  O(N^2) brute force of the predicate (the quirks reduce to these closed
  forms there); beyond level 0 in 1-D and 3-D pairs go missing.
* superSymmetric counts: #{j: d < h_i} + #{j: h_i < d <= h_j}, with the
  sync counters as the reference's atomics define them.
* The CSR quirk: indptr is the length-Nq offsets array.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle
from oracle import mlm_ref as M

F32 = np.float32


def test_morton_and_decode_match_torch():
    rng = np.random.default_rng(0)
    ci = rng.integers(-5, 1 << 17, (3000, 3)).astype(np.int32)
    t = torch.from_numpy(ci).to(torch.int64)

    def spread2(w):
        w = w & 0x00000000001fffff
        w = (w | w << 32) & 0x001f00000000ffff
        w = (w | w << 16) & 0x001f0000ff0000ff
        w = (w | w << 8) & 0x010f00f00f00f00f
        w = (w | w << 4) & 0x10c30c30c30c30c3
        return (w | w << 2) & 0x1249249249249249

    def spread1(x):
        x = x & 0x00000000ffffffff
        x = (x | (x << 16)) & 0x0000FFFF0000FFFF
        x = (x | (x << 8)) & 0x00FF00FF00FF00FF
        x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0F
        x = (x | (x << 2)) & 0x3333333333333333
        return (x | (x << 1)) & 0x5555555555555555

    m3 = (spread2(t[:, 0]) | (spread2(t[:, 1]) << 1) | (spread2(t[:, 2]) << 2)).numpy()
    m2 = (spread1(t[:, 0]) | (spread1(t[:, 1]) << 1)).numpy()
    np.testing.assert_array_equal(M.morton_encode(ci), m3)
    np.testing.assert_array_equal(M.morton_encode(ci[:, :2]), m2)
    for l in range(3):  # (morton >> 2 l).int()
        np.testing.assert_array_equal(M.as_i32(M.morton_encode(ci) >> (2 * l)),
                                      (torch.from_numpy(m3) >> (2 * l)).int().numpy())
    codes = rng.integers(-2 ** 31, 2 ** 31 - 1, 3000).astype(np.int32)
    c = torch.from_numpy(codes)

    def compact2(w):
        w = w & 0x1249249249249249
        w = (w ^ (w >> 2)) & 0x30c30c30c30c30c3
        w = (w ^ (w >> 4)) & 0xf00f00f00f00f00f
        w = (w ^ (w >> 8)) & 0x00ff0000ff0000ff
        w = (w ^ (w >> 16)) & 0x00ff00000000ffff
        return (w ^ (w >> 32)) & 0x00000000001fffff

    def compact1(x):
        x = x & 0x5555555555555555
        x = (x | (x >> 1)) & 0x3333333333333333
        x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0F
        x = (x | (x >> 4)) & 0x00FF00FF00FF00FF
        x = (x | (x >> 8)) & 0x0000FFFF0000FFFF
        return (x | (x >> 16)) & 0x00000000FFFFFFFF

    np.testing.assert_array_equal(M.morton_decode_i32(codes, 3),
                                  torch.stack([compact2(c), compact2(c >> 1), compact2(c >> 2)], 1).numpy())
    np.testing.assert_array_equal(M.morton_decode_i32(codes, 2), torch.stack([compact1(c), compact1(c >> 1)], 1).numpy())


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_resolution_levels_match_torch(dim):
    rng = np.random.default_rng(10 + dim)
    for _ in range(200):
        mn = rng.uniform(-2, 0, dim).astype(F32)
        mx = (mn + rng.uniform(0.5, 3, dim)).astype(F32)
        hmin = F32(rng.uniform(0.003, 0.2))
        if rng.random() < 0.3:  # extents that are exact multiples of hMin: log2 of a power of two
            hmin = F32((mx - mn).min() / 2 ** rng.integers(3, 9))
        hmax = F32(hmin * rng.uniform(1, 12))
        # cellOp.py:5-46 in torch ops
        ext = torch.from_numpy(mx) - torch.from_numpy(mn)
        q = (2 ** torch.floor(torch.log2(ext / float(hmin)))).int()
        hcell = (ext / q).max().item()
        _, qc, hc, hf, _ = M.grid_of_levels(mn, mx, float(hmin))
        assert np.array_equal(qc, q.numpy()) and F32(hcell) == hc
        assert hf == F32((ext / (2 ** 16)).max().item())
        want = []
        import math
        for i in range(int(math.ceil(math.log2(float(hmax) / hcell))) + 1):
            if (q / 2 ** i).min() <= 1:
                break
            want.append((q / 2 ** i).to(torch.int32).numpy())
        if not want:  # the reference's mergeCells fails on an empty list
            with pytest.raises(ValueError):
                M.resolution_levels(mn, mx, float(hmin), float(hmax))
            continue
        levels, qc, hc, hf, res = M.resolution_levels(mn, mx, float(hmin), float(hmax))
        assert levels == len(want) and np.array_equal(res, np.stack(want))


def test_helpers_match_compact_oracle():
    rng = np.random.default_rng(3)
    for dim in (1, 2, 3):
        mn, mx = np.full(dim, 0.1, F32), np.full(dim, 0.7, F32)
        per = [bool(x) for x in rng.random(dim) < 0.6]
        p = rng.uniform(-0.3, 1.1, (400, dim)).astype(F32)
        np.testing.assert_array_equal(M.wrap_once(p, mn, mx, per), oracle.wrap(p, mn, mx, per))
        d = M.mod_distance(p[0], p, mn, mx, per)
        np.testing.assert_array_equal(d, [oracle.pair_distance(p[0], q, mn, mx, per) for q in p])
        cells = rng.integers(-2 ** 31, 2 ** 31 - 1, (300, dim)).astype(np.int32)
        for L in (1, 97, 4096):
            np.testing.assert_array_equal(M.hash_index(cells, L), [oracle.hash_index(c, L) for c in cells])


def closed_form_rows(qpos, qsup, hsel, ypos, ysup, mn, mx, per, L, pred):
    """Rows by the closed form of the module docstring.  hsel[i]: the support
    choosing query i's level; pred(d, i, j) -> bool mask over sorted j."""
    st = M.Structure(M.wrap_once(ypos, mn, mx, per), ysup, mn, mx, L)
    dim = ypos.shape[1]
    rel = (st.spos - st.mn).astype(F32)
    mort = M.morton_encode(M.trunc_i32((rel / st.hcell).astype(F32)))
    rows = []
    for i in range(qpos.shape[0]):
        lg = M.log2f(F32(F32(hsel[i]) / st.hcell))
        li = int(np.clip(np.int64(M.trunc_i32(np.ceil(np.array([lg])))[0]) + 1, 1, st.levels))
        size = F32(st.hcell * F32(2 ** (li - 1)))
        cell = np.array([M.trunc_i32(np.floor((F32(qpos[i, d]) - st.mn[d]).astype(F32) / size).reshape(1))[0]
                         for d in range(dim)], np.int64)
        code = M.as_i32(mort >> (2 * (li - 1)))
        seq = []
        for o in itertools.product((-1, 0, 1), repeat=dim):
            cur = np.fmod(cell + np.array(o), st.res[li - 1].astype(np.int64)).astype(np.int32)[None, :]
            m = int(M.morton_encode_i32(cur)[0])
            if M.hash_index(M.morton_decode_i32(np.array([m], np.int32), dim), L)[0] != M.hash_index(cur, L)[0]:
                continue
            seq.extend(np.nonzero(code == m)[0].tolist())
        seq = np.asarray(seq, np.int64)
        if seq.size:
            d = M.mod_distance(qpos[i], st.spos[seq], mn, mx, per)
            seq = seq[pred(d, i, seq, st)]
        rows.append(st.si[seq] if seq.size else seq)
    return rows


def _cloud(rng, n, dim, spread=8.0):
    p = rng.uniform(0, 1, (n, dim)).astype(F32)
    h = (0.04 * np.exp2(rng.uniform(0, np.log2(spread), n))).astype(F32)
    return p, h


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [False, True])
def test_gather_rows_closed_form(dim, periodic):
    rng = np.random.default_rng(40 + dim + 5 * periodic)
    n = {1: 120, 2: 260, 3: 300}[dim]
    p, h = _cloud(rng, n, dim)
    mn, mx, per = np.zeros(dim, F32), np.ones(dim, F32), [periodic] * dim
    for mode in ("gather", "symmetric"):
        kind, row, col = M.search(p, h, None, None, domain_min=mn, domain_max=mx, periodic=per, mode=mode,
                                  hash_map_length=101)
        rows = closed_form_rows(p, h, h, p, h, mn, mx, per, 101, lambda d, i, j, st: d < h[i])
        want_col = np.concatenate(rows)
        want_row = np.concatenate([np.full(r.size, i) for i, r in enumerate(rows)])
        np.testing.assert_array_equal(row, want_row)
        np.testing.assert_array_equal(col, want_col)
        if dim == 2 and not periodic:
            # every pair within support, at least once (the C remainder steps
            # from the last cell of a dimension to cell 0, so with two cells
            # a neighbour cell can be visited twice): the O(N^2) brute force
            for i in range(n):
                d = M.mod_distance(p[i], p, mn, mx, per)
                assert sorted(set(col[row == i].tolist())) == np.nonzero(d < h[i])[0].tolist()


@pytest.mark.parametrize("dim", [1, 3])
def test_level_shift_quirk_drops_pairs_beyond_level_zero(dim):
    """Supports above hCell choose a coarser level, whose codes the reference
    shifts by 2 bits per level -- right for 2-D Morton codes, not for the
    1-D code (the cell itself) or the 3-D one: pairs go missing."""
    rng = np.random.default_rng(77 + dim)
    p, h = _cloud(rng, 300, dim, spread=8.0)
    mn, mx = np.zeros(dim, F32), np.ones(dim, F32)
    _, row, col = M.search(p, h, None, None, domain_min=mn, domain_max=mx, periodic=[False] * dim)
    brute = sum(int(np.sum(M.mod_distance(p[i], p, mn, mx, [False] * dim) < h[i])) for i in range(300))
    assert row.size < brute


@pytest.mark.parametrize("dim", [1, 3])
def test_single_level_is_complete(dim):
    rng = np.random.default_rng(78 + dim)
    n = 350
    p = rng.uniform(0, 1, (n, dim)).astype(F32)
    h = np.full(n, 0.11 if dim == 3 else 0.02, F32)
    mn, mx = np.zeros(dim, F32), np.ones(dim, F32)
    _, row, col = M.search(p, h, None, None, domain_min=mn, domain_max=mx, periodic=[False] * dim)
    for i in range(n):
        d = M.mod_distance(p[i], p, mn, mx, [False] * dim)
        assert sorted(set(col[row == i].tolist())) == np.nonzero(d < h[i])[0].tolist()


@pytest.mark.parametrize("dim", [2])
def test_super_symmetric_counts_and_rows(dim):
    rng = np.random.default_rng(90 + dim)
    n = 220
    p, h = _cloud(rng, n, dim, spread=6.0)
    mn, mx, per = np.zeros(dim, F32), np.ones(dim, F32), [False] * dim
    kind, idx, indptr, entries, nr, nc = M.search(p, h, None, None, domain_min=mn, domain_max=mx, periodic=per,
                                                  mode="superSymmetric", fmt="csr")
    assert indptr.shape[0] == n and indptr.dtype == np.int32          # the length-Nq indptr
    np.testing.assert_array_equal(indptr[1:], np.cumsum(entries)[:-1])
    for i in range(n):
        d = M.mod_distance(p[i], p, mn, mx, per)
        want = int(np.sum(d < h[i])) + int(np.sum((d > h[i]) & (d <= h)))
        assert entries[i] == want
        got = idx[indptr[i]:indptr[i] + entries[i]]
        assert sorted(set(got.tolist())) == np.nonzero((d < h[i]) | (d < h))[0].tolist()


def test_scatter_is_transposed_gather_and_non_self_rows_overflow():
    rng = np.random.default_rng(5)
    q, hq = _cloud(rng, 200, 2)
    mn, mx, per = np.zeros(2, F32), np.ones(2, F32), [False, False]
    _, row, col = M.search(q, hq, None, None, domain_min=mn, domain_max=mx, periodic=per, mode="scatter")
    _, grow, gcol = M.search(q, hq, None, None, domain_min=mn, domain_max=mx, periodic=per, mode="gather")
    order = np.argsort(gcol, kind="stable")
    np.testing.assert_array_equal(row, gcol[order])
    np.testing.assert_array_equal(col, grow[order])
    kind, idx, indptr, entries, nr, nc = M.search(q, hq, None, None, domain_min=mn, domain_max=mx, periodic=per,
                                                  mode="scatter", fmt="csr")
    assert indptr.shape[0] == 201 and idx.shape[0] == row.shape[0]  # scatter CSR: coo_to_csr, length Nq + 1
    # a separate query set: rows are written at offsets counted over the
    # reference set, and a query with more pairs writes past its row
    r, hr = _cloud(rng, 150, 2)
    with pytest.raises(ValueError):
        M.search(r, hr * 2, q, hq, domain_min=mn, domain_max=mx, periodic=per, mode="gather")
    with pytest.raises(ValueError):  # more queries than reference points
        M.search(q, hq, r[:50], hr[:50], domain_min=mn, domain_max=mx, periodic=per)
