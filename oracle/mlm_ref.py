"""The reference's own ``algorithm='mlm'`` search, restated in numpy (test
infrastructure only; the product never imports it).

torchCompactRadius 0.5.5 runs ``searchNeighbors_mlm``
(py/multiLevelMemory/wrapper.py:145-308) for ``radiusSearch(...,
algorithm='mlm')`` (py/neighborhood.py:255-262, 292-299), with the hash-map
cell table (``useDenseMLM=False``, the default).  Its behaviour differs from
the compact search in ways this restatement keeps, step for step:

* levels from ``computeResolutionLevels`` (py/multiLevelMemory/cellOp.py:5-46):
  qCells = 2^floor(log2(extent / hMin)) in fp32, hCell = max(extent / qCells),
  levels = ceil(log2(hMax / hCell)) + 1 resolutions while qCells / 2^i > 1;
* per-level codes ``(morton(cell at hCell) >> 2*level).int()``
  (py/multiLevelMemory/indexing.py:71-78) -- a 2-D shift also in 3-D;
* points sorted by the Morton code of a 2^16 grid (wrapper.py:68-77);
* per-level cells as runs of equal codes in that order, merged over the
  levels (cellData.py:44-62, cellOp.py:48-64), and a hash map over the
  decoded codes (hashData.py:21-39) -- the decode runs torch's int32
  arithmetic with its 64-bit masks truncated to 32 bits;
* counts over the REFERENCE set with unwrapped positions
  (wrapper.py:213-232); the traversal (cpp/multiLevelMemory/mlmUtil.h:
  234-282) picks the level of each point's support, steps to neighbour cells
  with the C remainder (no periodic wrap) and matches level and Morton code
  in the hash bucket; ``dist < h_i`` counts, and ``h_j < dist <= h_i``
  increments the reference point's sync counter and raises its synced
  support (countNeighborsHashmap.h:99-124, the CPU build's atomics);
* int32 offsets of length Nq from the counts (+ sync counts for
  superSymmetric, wrapper.py:255-270); the fill over the wrapped QUERY set
  (buildNeighborhoodHashmap.h:103-128): gather semantics for every mode but
  superSymmetric (``dist < h_i || dist < h_j`` with the synced support
  choosing the level); ``col = sortingIndices[col]``;
* CSR with the length-Nq offsets as indptr (wrapper.py:298-306); scatter as
  the transposed gather search (wrapper.py:191-192).

Orders the reference leaves to unstable sorts (``torch.argsort`` of the fine
Morton codes and of the hash indices, ``transposeCOO``) are taken stable
here, as the compact oracle does.  Where the reference writes past its
counts (queries that find more pairs than the reference set counted, or
more queries than reference points) its output is undefined; this
restatement raises ``ValueError`` there.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

F32, F64, I32, I64 = np.float32, np.float64, np.int32, np.int64
PRIMES = (73856093, 19349663, 83492791)  # cpp/hashing.h:27


def log2f(x: np.ndarray) -> np.ndarray:
    """fp32 log2: the exact log2 of the fp32 argument rounded to fp32."""
    x = np.asarray(x, F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.log2(x.astype(F64)).astype(F32)


def trunc_i32(x) -> np.ndarray:
    """(int32) of fp32 values as x86 converts them: toward zero, INT32_MIN
    for NaN and out-of-range values."""
    x = np.asarray(x, F64)
    ok = np.isfinite(x) & (x > -2147483649.0) & (x < 2147483648.0)
    out = np.full(x.shape, np.iinfo(I32).min, I64)
    out[ok] = np.trunc(x[ok]).astype(I64)
    return out.astype(I32)


def as_i32(v) -> np.ndarray:
    """Low 32 bits of int64 values, as signed int32 (torch's .int())."""
    v = np.asarray(v, I64)
    return (v & 0xFFFFFFFF).astype(np.uint32).view(I32)


def _spread1(x):  # indexing.py:29-35, int64
    x = x & 0x00000000FFFFFFFF
    x = (x | (x << 16)) & 0x0000FFFF0000FFFF
    x = (x | (x << 8)) & 0x00FF00FF00FF00FF
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0F
    x = (x | (x << 2)) & 0x3333333333333333
    x = (x | (x << 1)) & 0x5555555555555555
    return x


def _spread2(w):  # indexing.py:3-10, int64
    w = w & 0x00000000001FFFFF
    w = (w | w << 32) & 0x001F00000000FFFF
    w = (w | w << 16) & 0x001F0000FF0000FF
    w = (w | w << 8) & 0x010F00F00F00F00F
    w = (w | w << 4) & 0x10C30C30C30C30C3
    w = (w | w << 2) & 0x1249249249249249
    return w


def morton_encode(ci: np.ndarray) -> np.ndarray:
    """mortonEncode (indexing.py:60-68): int32 cells -> int64 codes (the
    1-D code is the int32 cell itself); int64 wrap-around as torch."""
    ci = np.asarray(ci, I32)
    dim = ci.shape[1]
    with np.errstate(over="ignore"):
        if dim == 1:
            return ci[:, 0].astype(I64)
        if dim == 2:
            return _spread1(ci[:, 0].astype(I64)) | (_spread1(ci[:, 1].astype(I64)) << 1)
        return _spread2(ci[:, 0].astype(I64)) | (_spread2(ci[:, 1].astype(I64)) << 1) | \
            (_spread2(ci[:, 2].astype(I64)) << 2)


def morton_encode_i32(cells: np.ndarray) -> np.ndarray:
    """mortonEncode<int32_t> of the C++ traversal (mlmUtil.h:139-148): the
    int64 code truncated to int32."""
    return as_i32(morton_encode(cells))


def _m32(c: int) -> np.int32:
    """A torch integer constant converted for an int32 tensor (low 32 bits)."""
    return np.uint32(c & 0xFFFFFFFF).view(I32)


def _compact1_i32(x):  # indexing.py:36-43 on an int32 tensor
    x = x & _m32(0x5555555555555555)
    x = (x | (x >> 1)) & _m32(0x3333333333333333)
    x = (x | (x >> 2)) & _m32(0x0F0F0F0F0F0F0F0F)
    x = (x | (x >> 4)) & _m32(0x00FF00FF00FF00FF)
    x = (x | (x >> 8)) & _m32(0x0000FFFF0000FFFF)
    x = (x | (x >> 16)) & _m32(0x00000000FFFFFFFF)
    return x


def _compact2_i32(w):  # indexing.py:11-18 on an int32 tensor
    w = w & _m32(0x1249249249249249)
    w = (w ^ (w >> 2)) & _m32(0x30C30C30C30C30C3)
    w = (w ^ (w >> 4)) & _m32(0xF00F00F00F00F00F)
    w = (w ^ (w >> 8)) & _m32(0x00FF0000FF0000FF)
    w = (w ^ (w >> 16)) & _m32(0x00FF00000000FFFF)
    w = (w ^ _sign_fill(w)) & _m32(0x00000000001FFFFF)  # torch: int32 >> 32 is the sign fill
    return w


def _sign_fill(w):
    """torch's int32 >> k for k >= 32: 0 or -1 by sign."""
    return np.where(w < 0, np.int32(-1), np.int32(0)).astype(I32)


def morton_decode_i32(code: np.ndarray, dim: int) -> np.ndarray:
    """mortonDecode (indexing.py:70-77) of int32 codes with torch's int32
    arithmetic (masks truncated to 32 bits): int32 [n, dim]."""
    code = np.asarray(code, I32)
    if dim == 1:
        return code[:, None]
    if dim == 2:
        return np.stack([_compact1_i32(code), _compact1_i32(code >> 1)], 1).astype(I32)
    return np.stack([_compact2_i32(code), _compact2_i32(code >> 1), _compact2_i32(code >> 2)], 1).astype(I32)


def hash_index(cells: np.ndarray, L: int) -> np.ndarray:
    """hashIndexing (cpp/hashing.h:20-35): uint32 wrap-around."""
    cells = np.asarray(cells, I32)
    u = cells.astype(np.uint32).astype(np.uint64)
    if cells.shape[1] == 1:
        return (u[:, 0] % np.uint64(L)).astype(I32)
    h = np.zeros(cells.shape[0], np.uint64)
    for d in range(cells.shape[1]):
        h = (h + u[:, d] * np.uint64(PRIMES[d])) & np.uint64(0xFFFFFFFF)
    return (h % np.uint64(L)).astype(I32)


def mod_distance(xi: np.ndarray, xj: np.ndarray, mn, mx, periodic) -> np.ndarray:
    """modDistance (cpp/common.h:106-129) of one point against many: fp32
    differences, fp64 moduloOp in periodic dims, fp64 squares added to an fp32
    running sum, fp32 sqrt."""
    xi = np.asarray(xi, F32)
    xj = np.asarray(xj, F32).reshape(-1, xi.shape[0])
    s = np.zeros(xj.shape[0], F32)
    for d in range(xi.shape[0]):
        diff32 = (xi[d] - xj[:, d]).astype(F32)
        if periodic[d]:
            h = F64(F32(F32(mx[d]) - F32(mn[d])))
            t = diff32.astype(F64) + h / 2.0
            diff = (t - np.floor(t / h) * h) - h / 2.0
        else:
            diff = diff32.astype(F64)
        s = (s.astype(F64) + diff * diff).astype(F32)
    return np.sqrt(s).astype(F32)


def wrap_once(pos, mn, mx, periodic):
    """getPeriodicPointCloud (py/util.py:375-387): torch.remainder per
    periodic dimension."""
    pos = np.asarray(pos, F32).copy()
    for d in range(pos.shape[1]):
        if periodic[d]:
            L = F32(F32(mx[d]) - F32(mn[d]))
            r = np.fmod((pos[:, d] - F32(mn[d])).astype(F32), L).astype(F32)
            r = np.where((r != 0) & ((r < 0) != (L < 0)), (r + L).astype(F32), r).astype(F32)
            pos[:, d] = (r + F32(mn[d])).astype(F32)
    return pos


def offsets(dim: int) -> np.ndarray:
    """getOffsets (mlmUtil.h:292-324): {-1,0,1}^dim, dim 0 slowest."""
    import itertools
    return np.array(list(itertools.product((-1, 0, 1), repeat=dim)), I32)


def grid_of_levels(mn, mx, hmin: float):
    """computeCellResolution (cellOp.py:5-18) and hFine (cellOp.py:24-25):
    (ext, qCells int32[dim], hCell, hFine, None)."""
    ext = (np.asarray(mx, F32) - np.asarray(mn, F32)).astype(F32)
    q = (ext / F32(hmin)).astype(F32)
    qcells = trunc_i32((F32(2.0) ** np.floor(log2f(q))).astype(F32))
    hcell = F32((ext / qcells.astype(F32)).astype(F32).max())
    hfine = F32((ext / F32(65536.0)).astype(F32).max())
    return ext, qcells, hcell, hfine, None


def resolution_levels(mn, mx, hmin: float, hmax: float):
    """computeResolutionLevels (cellOp.py:21-46): (levels, qCells int32[dim],
    hCell, hFine, cellResolutions int32[levels, dim])."""
    ext, qcells, hcell, hfine, _ = grid_of_levels(mn, mx, hmin)
    hcell = float(hcell)
    ratio = float(hmax) / hcell
    levels = int(math.ceil(math.log2(ratio)))
    res = []
    for i in range(levels + 1):
        if float((qcells.astype(F32) / F32(2 ** i)).astype(F32).min()) <= 1:
            break
        res.append(trunc_i32((qcells.astype(F32) / F32(2 ** i)).astype(F32)))
    if not res:
        raise ValueError("multi-level search: no resolution level (the reference fails to merge an empty cell list)")
    return len(res), qcells, F32(hcell), hfine, np.stack(res).astype(I32)


class Structure:
    """buildDataStructure (wrapper.py:46-116) over the wrapped reference set."""

    def __init__(self, ypos, ysup, mn, mx, L: int):
        ypos = np.asarray(ypos, F32)
        ysup = np.asarray(ysup, F32)
        n, dim = ypos.shape
        self.dim = dim
        self.mn = np.asarray(mn, F32)
        self.L = int(L)
        self.levels, self.qcells, self.hcell, self.hfine, self.res = resolution_levels(
            mn, mx, float(ysup.min()), float(ysup.max()))
        rel = (ypos - self.mn).astype(F32)
        ci = trunc_i32((rel / self.hcell).astype(F32))
        morton = morton_encode(ci)
        codes = [as_i32(morton >> (2 * l)) for l in range(self.levels)]
        fine = morton_encode(trunc_i32((rel / self.hfine).astype(F32)))
        self.si = np.argsort(fine, kind="stable").astype(I64)
        self.spos = ypos[self.si]
        self.ssup = ysup[self.si]
        begin, end, index, level = [], [], [], []
        for l in range(self.levels):
            c = codes[l][self.si]
            starts = np.nonzero(np.r_[True, c[1:] != c[:-1]])[0] if n else np.zeros(0, I64)
            ends = np.r_[starts[1:], n] if n else starts
            begin.append(starts)
            end.append(ends)
            index.append(c[starts])
            level.append(np.full(starts.shape[0], l, I32))
        self.cbegin = np.concatenate(begin).astype(I32)
        self.cend = np.concatenate(end).astype(I32)
        self.cindex = np.concatenate(index).astype(I32)
        self.clevel = np.concatenate(level).astype(I32)
        hashed = hash_index(morton_decode_i32(self.cindex, dim), self.L)
        self.sorted_cells = np.argsort(hashed, kind="stable").astype(I64)
        hs = hashed[self.sorted_cells]
        self.hoff = np.full(self.L, -1, I32)
        self.hocc = np.zeros(self.L, I32)
        if hs.size:
            u, first, cnt = np.unique(hs, return_index=True, return_counts=True)
            self.hoff[u] = first
            self.hocc[u] = cnt

    def traverse(self, pos, h):
        """iterateCellHashed (mlmUtil.h:234-282): the sorted slots visited for
        a point at `pos` with support `h`, in visiting order."""
        l = log2f(F32(F32(h) / self.hcell))
        with np.errstate(invalid="ignore"):
            cv = np.ceil(l)
        li = int(trunc_i32(np.array([cv]))[0])
        li = int(np.clip(np.int64(li) + 1, 1, self.levels))
        size = F32(self.hcell * F32(2 ** (li - 1)))
        cell = np.array([int(trunc_i32(np.floor((F32(pos[d]) - self.mn[d]).astype(F32) / size).reshape(1))[0])
                         for d in range(self.dim)], I64)
        res = self.res[li - 1].astype(I64)
        out = []
        for o in offsets(self.dim):
            cur = np.fmod(cell + o, res).astype(I32)[None, :]
            b = int(hash_index(cur, self.L)[0])
            mort = int(morton_encode_i32(cur)[0])
            off, occ = int(self.hoff[b]), int(self.hocc[b])
            for hh in range(off, off + occ):
                c = self.sorted_cells[hh]
                if self.clevel[c] != li - 1 or self.cindex[c] != mort:
                    continue
                out.extend(range(int(self.cbegin[c]), int(self.cend[c])))
        return np.asarray(out, I64)


def search(qpos, qsup, rpos, rsup, *, domain_min=None, domain_max=None, periodic=None, mode="gather",
           hash_map_length=4096, fmt="coo"):
    """searchNeighbors_mlm (wrapper.py:145-308) as radiusSearch calls it
    (neighborhood.py:292-299; referencePointCloud=None means the query
    cloud).  Returns ('coo', row int64, col int64) or ('csr', indices int64,
    indptr int32, rowEntries, numRows, numCols)."""
    qpos = np.asarray(qpos, F32)
    dim = qpos.shape[1]
    if rpos is None:
        rpos, rsup = qpos, qsup
    rpos = np.asarray(rpos, F32)
    qsup = np.asarray(qsup, F32)
    rsup = np.asarray(rsup, F32)
    if domain_min is None:
        mn, mx, per = np.zeros(dim, F32), np.ones(dim, F32), [False] * dim
    else:
        mn, mx = np.asarray(domain_min, F32), np.asarray(domain_max, F32)
        per = [bool(p) for p in periodic]
    if mode == "scatter":  # wrapper.py:191-192
        r = search(rpos, rsup, qpos, qsup, domain_min=mn, domain_max=mx, periodic=per, mode="gather",
                   hash_map_length=hash_map_length, fmt="coo")
        row, col = r[2], r[1]
        order = np.argsort(row, kind="stable")
        row, col = row[order], col[order]
        if fmt == "coo":
            return ("coo", row, col)
        nrows = qpos.shape[0]
        nj = np.bincount(row, minlength=nrows).astype(I64) if row.size else np.zeros(nrows, I64)
        indptr = np.zeros(nrows + 1, I64)
        np.cumsum(nj, out=indptr[1:])
        return ("csr", col, as_i32(indptr), nj, nrows, rpos.shape[0])
    nq, nr = qpos.shape[0], rpos.shape[0]
    xw = wrap_once(qpos, mn, mx, per)
    yw = wrap_once(rpos, mn, mx, per)
    st = Structure(yw, rsup, mn, mx, hash_map_length)
    inv = np.empty(nr, I64)
    inv[st.si] = np.arange(nr)
    counts = np.zeros(nr, I32)
    sync = np.zeros(nr, I32)
    nsup = st.ssup.copy()
    for i in range(nr):  # countNeighborsHashmap over the unwrapped reference set
        h_i = rsup[i]
        js = st.traverse(rpos[i], h_i)
        if js.size == 0:
            continue
        d = mod_distance(rpos[i], st.spos[js], mn, mx, per)
        counts[i] = int(np.sum(d < h_i))
        hit = (d > st.ssup[js]) & (d <= h_i)
        for j in js[hit]:
            sync[j] += 1
            if h_i > nsup[j]:
                nsup[j] = h_i
    sym = mode == "superSymmetric"
    if sym and nq != nr:
        raise ValueError("SuperSymmetric support mode only works for equal number of query and reference particles")
    sync_d = sync[inv]
    ssup_d = nsup[inv]
    tot = counts.astype(I64) + (sync_d if sym else 0)
    offs = np.zeros(nr, I64)
    if nr > 1:
        offs[1:] = np.cumsum(tot)[:-1]
    offs = as_i32(offs)
    length = int(offs[-1]) + int(tot[-1]) if nr else 0
    if nq > nr:
        raise ValueError("more query points than reference points: the reference reads offsets past its end")
    row = np.zeros(length, I64)
    col = np.zeros(length, I64)
    for i in range(nq):  # buildNeighborhoodHashmap over the wrapped query set
        hs = ssup_d[i] if sym else qsup[i]
        js = st.traverse(xw[i], hs)
        if js.size == 0:
            continue
        d = mod_distance(xw[i], st.spos[js], mn, mx, per)
        sel = (d < qsup[i]) | (d < st.ssup[js]) if sym else (d < qsup[i])
        js = js[sel]
        o = int(offs[i])
        end = int(offs[i + 1]) if i + 1 < nr else length
        if o + js.size > end:
            raise ValueError("a query finds more pairs than the reference set counted: the reference writes past "
                             "its row (undefined output)")
        row[o:o + js.size] = i
        col[o:o + js.size] = js
    col = st.si[col] if length else col
    if fmt == "coo":
        return ("coo", row, col)
    return ("csr", col, offs, (counts + sync_d).astype(I32) if sym else counts, nq, nr)
