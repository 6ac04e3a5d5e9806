"""Host driver of the compact search: validates tensors, fills ``tcr_cloud``
descriptors and runs build -> count -> scan -> fill through libtcr_hip.so on
the current HIP stream -- back to back into outputs sized from the previous
search of the same batch shape, or as count, one sync for the pair total,
then fill -- optionally replayed from a captured HIP graph.

Reference counterpart: neighborSearch_cpp + searchNeighbors_cpp
(src/torchCompactRadius/compactHashing/cppSearch.py:12-98,179-244).
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass
from typing import List, NamedTuple, Optional, Sequence, Tuple

import torch

from . import _lib
from ._lib import MODES, TcrCloud, check, ptr


class CompactHashMap(NamedTuple):
    """Structure of one search, the counterpart of the reference's record
    (src/torchCompactRadius/compactHashing/datastructure.py:8-31)."""
    sortedPositions: torch.Tensor
    referencePositions: torch.Tensor
    hashTable: torch.Tensor
    hashMapLength: int
    sortedCellTable: torch.Tensor
    numCells: torch.Tensor
    qMin: torch.Tensor
    qMax: torch.Tensor
    minD: torch.Tensor
    maxD: torch.Tensor
    sortIndex: torch.Tensor
    hCell: torch.Tensor
    periodicity: torch.Tensor
    searchRadius: int
    sortedSupports: Optional[torch.Tensor] = None
    referenceSupports: Optional[torch.Tensor] = None
    fixedSupport: Optional[float] = None


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _require_device(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} is not on the GPU: torchcompactradius_amd runs its search only on "
                           "HIP devices (no CPU path)")


def _f32(t: Optional[torch.Tensor], name: str) -> Optional[torch.Tensor]:
    if t is None:
        return None
    _require_device(t, name)
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32 (got {t.dtype})")
    return t.contiguous()


def periodic_mask(periodic, dim: int) -> int:
    if periodic is None:
        return 0
    if isinstance(periodic, torch.Tensor):
        periodic = periodic.tolist()
    if isinstance(periodic, bool):
        periodic = [periodic] * dim
    m = 0
    for d, p in enumerate(periodic):
        if bool(p):
            m |= 1 << d
    return m


def _host_floats(v, dim: int, name: str) -> List[float]:
    if isinstance(v, torch.Tensor):
        if v.device.type == "cpu" and v.dtype == torch.float32 and v.dim() == 1:
            v = v.tolist()
        else:
            v = v.detach().to("cpu", torch.float32).reshape(-1).tolist()
    v = [float(x) for x in v]
    if len(v) != dim:
        raise ValueError(f"{name} has {len(v)} entries, expected {dim}")
    return v


@dataclass
class CloudSpec:
    """One query/reference pair of a batched search."""
    qpos: torch.Tensor
    qsup: Optional[torch.Tensor]
    rpos: Optional[torch.Tensor]        # None: self-search
    rsup: Optional[torch.Tensor]
    domain_min: Sequence[float]
    domain_max: Sequence[float]
    periodic: int                       # bit mask
    mode: str = "gather"
    hash_map_length: int = 4096
    search_range: int = 1


class SearchResult(NamedTuple):
    offsets: torch.Tensor               # int64 [nq_total + 1]
    row: Optional[torch.Tensor]         # int64 [P] (local query index) or None
    col: torch.Tensor                   # int64 [P] (local reference index)
    total: int
    workspace: torch.Tensor
    clouds: ctypes.Array
    indptr: Optional[torch.Tensor] = None       # int32 [nq_total + 1] (csr=True)
    row_entries: Optional[torch.Tensor] = None  # int64 [nq_total] (csr=True)


def _cloud_array(specs: Sequence[CloudSpec], q_offs, r_offs) -> ctypes.Array:
    arr = (TcrCloud * len(specs))()
    for k, s in enumerate(specs):
        c = arr[k]
        dim = s.qpos.shape[1]
        c.dim = dim
        c.mode = MODES[s.mode]
        c.periodic_mask = s.periodic
        c.search_range = s.search_range
        for d in range(dim):
            c.domain_min[d] = s.domain_min[d]
            c.domain_max[d] = s.domain_max[d]
        c.hash_map_length = s.hash_map_length
        c.self_search = 1 if s.rpos is None else 0
        c.nq = s.qpos.shape[0]
        c.q_off, c.q_foff = q_offs[k]
        if s.rpos is None:
            c.nr, c.r_off, c.r_foff = c.nq, 0, 0
        else:
            c.nr = s.rpos.shape[0]
            c.r_off, c.r_foff = r_offs[k]
    return arr


# Pair total of the last search per batch shape (clouds, query points,
# reference points): the capacity guess of the next search of that shape.
_pair_hint: dict = {}


def _remember(key, total: int) -> None:
    if len(_pair_hint) > 4096:
        _pair_hint.clear()
    _pair_hint[key] = total


def _concat(specs: Sequence[CloudSpec], device):
    """Flat query/reference buffers of a batch and each cloud's offsets."""
    q_offs, r_offs = [], []
    qo = qf = ro = rf = 0
    for s in specs:
        q_offs.append((qo, qf))
        qo += s.qpos.shape[0]
        qf += s.qpos.numel()
        if s.rpos is not None:
            r_offs.append((ro, rf))
            ro += s.rpos.shape[0]
            rf += s.rpos.numel()
        else:
            r_offs.append((0, 0))
    if len(specs) == 1:
        s = specs[0]
        cat = (s.qpos.reshape(-1), s.qsup, None if s.rpos is None else s.rpos.reshape(-1), s.rsup)
    else:
        qpos_cat = torch.cat([s.qpos.reshape(-1) for s in specs])
        qs = [s.qsup for s in specs]
        qsup_cat = torch.cat([q if q is not None else torch.zeros(s.qpos.shape[0], device=device)
                              for q, s in zip(qs, specs)]) if any(q is not None for q in qs) else None
        rl = [s for s in specs if s.rpos is not None]
        rpos_cat = torch.cat([s.rpos.reshape(-1) for s in rl]) if rl else None
        rsl = [s.rsup for s in rl]
        rsup_cat = torch.cat([r if r is not None else torch.zeros(s.rpos.shape[0], device=device)
                              for r, s in zip(rsl, rl)]) if any(r is not None for r in rsl) else None
        cat = (qpos_cat, qsup_cat, rpos_cat, rsup_cat)
    return cat, q_offs, r_offs, qo, ro


class SearchPlan:
    """A batch of clouds bound to flat device buffers and a workspace: the
    unit that one search launch, a captured graph and the serving loop run.

    The buffers (``qpos``, ``qsup``, ``rpos``, ``rsup``) are concatenations of
    the clouds' tensors in order (or the caller's own tensors for one cloud);
    rewriting them in place and running the plan again searches the new
    values.  ``index_dtype`` is torch.int64 (the reference's) or torch.int32.
    """

    def __init__(self, specs: Sequence[CloudSpec], want_rows: bool, index_dtype=torch.int64):
        if len(specs) == 0:
            raise ValueError("empty batch")
        if index_dtype not in (torch.int64, torch.int32):
            raise ValueError("index_dtype must be torch.int64 or torch.int32")
        self.L = _lib.load()
        self.specs = list(specs)
        self.want_rows = want_rows
        self.index_dtype = index_dtype
        self.index_bytes = 8 if index_dtype == torch.int64 else 4
        self.device = specs[0].qpos.device
        (self.qpos, self.qsup, self.rpos, self.rsup), q_offs, r_offs, self.nq, self.nr = _concat(specs, self.device)
        self.clouds = _cloud_array(specs, q_offs, r_offs)
        self.n = len(specs)
        self.ws_bytes = self.L.tcr_search_workspace_size(self.clouds, self.n)
        if self.ws_bytes == 0:
            check(_lib.TCR_ERR_ARG)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.offsets = torch.empty(self.nq + 1, dtype=torch.int64, device=self.device)
        self.key = (self.n, self.nq, self.nr, tuple((c.dim, c.mode) for c in self.clouds[:8]))
        self._graph = None
        self._graph_out = None
        self.tuning = None          # launch shape (see autotune.py) or None: automatic

    def _install_tuning(self):
        from .autotune import install
        install(self.tuning)

    def _args(self):
        return (self.clouds, self.n, ptr(self.qpos), ptr(self.qsup), ptr(self.rpos), ptr(self.rsup),
                ptr(self.ws), self.ws_bytes)

    def _alloc(self, n: int):
        col = torch.empty(n, dtype=self.index_dtype, device=self.device)
        row = torch.empty(n, dtype=self.index_dtype, device=self.device) if self.want_rows else None
        return row, col

    def _finish(self, stream) -> int:
        total = ctypes.c_int64(0)
        check(self.L.tcr_search_finish(self.clouds, self.n, ptr(self.ws), self.ws_bytes, ptr(self.offsets),
                                       ctypes.byref(total), stream))
        return int(total.value)

    def _fill_exact(self, P: int, stream):
        row, col = self._alloc(P)
        if P > 0:  # nothing to write otherwise (and an empty tensor has no data pointer)
            check(self.L.tcr_search_fill(*self._args(), ptr(self.offsets), ptr(row), ptr(col), self.index_bytes,
                                         stream))
        return row, col

    def _csr(self, stream):
        indptr = torch.empty(self.nq + 1, dtype=torch.int32, device=self.device)
        row_entries = torch.empty(self.nq, dtype=torch.int64, device=self.device)
        check(self.L.tcr_csr_arrays(ptr(self.offsets), self.nq, ptr(indptr), ptr(row_entries), stream))
        return indptr, row_entries

    def run(self, capacity: Optional[int] = None, csr: bool = False, stream: Optional[int] = None) -> "SearchResult":
        """One search.  With a capacity guess (default: from the last search
        of this batch shape) count, scan and fill run back to back; a guess
        that is too small costs one extra fill launch.  ``csr`` also returns
        the int32 indptr and int64 rowEntries arrays, computed on the stream
        before the host waits.  ``stream``: the HIP stream handle (default:
        torch's current stream)."""
        if stream is None:
            stream = _stream()
        self._install_tuning()
        ip = re = None
        if capacity is None:
            hint = _pair_hint.get(self.key)
            capacity = None if hint is None else hint + hint // 32 + 256
        if capacity is not None:
            row_buf, col_buf = self._alloc(capacity)
            if csr and row_buf is None:
                # the CSR arrays are written in the fill phase, before the host waits
                ip = torch.empty(self.nq + 1, dtype=torch.int32, device=self.device)
                re = torch.empty(self.nq, dtype=torch.int64, device=self.device)
                check(self.L.tcr_search_launch_csr(*self._args(), ptr(self.offsets), ptr(col_buf), self.index_bytes,
                                                   capacity, ptr(ip), ptr(re), stream))
            else:
                check(self.L.tcr_search_launch(*self._args(), ptr(self.offsets), ptr(row_buf), ptr(col_buf),
                                               self.index_bytes, capacity, stream))
                if csr:
                    ip, re = self._csr(stream)
            P = self._finish(stream)
            _remember(self.key, P)
            if P <= capacity:
                return SearchResult(self.offsets, None if row_buf is None else row_buf[:P], col_buf[:P], P,
                                    self.ws, self.clouds, ip, re)
        else:
            total = ctypes.c_int64(0)
            check(self.L.tcr_search_count(*self._args(), ptr(self.offsets), ctypes.byref(total), stream))
            P = int(total.value)
            _remember(self.key, P)
            if csr:
                ip, re = self._csr(stream)
        row, col = self._fill_exact(P, stream)
        return SearchResult(self.offsets, row, col, P, self.ws, self.clouds, ip, re)

    def capture(self, capacity: int) -> None:
        """Capture the whole search into a HIP graph with outputs of
        ``capacity`` pairs (replayed by ``replay``)."""
        self.release_graph()
        row, col = self._alloc(capacity)
        self._install_tuning()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        handle = ctypes.c_void_p(0)
        check(self.L.tcr_graph_capture(*self._args(), ptr(self.offsets), ptr(row), ptr(col), self.index_bytes,
                                       capacity, ctypes.byref(handle), side.cuda_stream))
        torch.cuda.current_stream(self.device).wait_stream(side)
        self._graph = handle
        self._graph_out = (row, col, capacity)

    def replay(self) -> "SearchResult":
        """Replay the captured graph on the current stream.  The outputs are
        the captured buffers (views of their first P entries): consume or copy
        them before the next replay."""
        if self._graph is None:
            raise RuntimeError("no captured graph: call capture() first")
        stream = _stream()
        check(self.L.tcr_graph_launch(self._graph, stream))
        P = self._finish(stream)
        row, col, cap = self._graph_out
        if P <= cap:
            return SearchResult(self.offsets, None if row is None else row[:P], col[:P], P, self.ws, self.clouds)
        row, col = self._fill_exact(P, stream)
        return SearchResult(self.offsets, row, col, P, self.ws, self.clouds)

    def release_graph(self) -> None:
        if self._graph is not None:
            self.L.tcr_graph_destroy(self._graph)
            self._graph = None
            self._graph_out = None

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass


# Plans of recent single-cloud searches, keyed by the input buffers, the
# cloud description, the output kind and the stream: a repeated search of the
# same tensors (the usual simulation loop, positions updated in place) reuses
# the plan's descriptor array, workspace and offsets instead of rebuilding
# them.  A plan keeps its input tensors alive, so a key's data pointers
# cannot be taken by other tensors while it is cached.  Nothing a search
# returns to the caller aliases the plan's buffers.
_plan_cache: "OrderedDict[tuple, SearchPlan]" = OrderedDict()
_PLAN_CACHE_MAX = 4


def _tensor_key(t: Optional[torch.Tensor]):
    return None if t is None else (t.data_ptr(), tuple(t.shape), t.stride(), t.dtype)


def run_search(specs: Sequence[CloudSpec], want_rows: bool, index_dtype=torch.int64,
               csr: bool = False) -> "SearchResult":
    """Search a batch of clouds in one launch chain (see SearchPlan)."""
    if len(specs) != 1:
        return SearchPlan(specs, want_rows, index_dtype).run(csr=csr)
    s = specs[0]
    stream = _stream()
    key = (_tensor_key(s.qpos), _tensor_key(s.qsup), _tensor_key(s.rpos), _tensor_key(s.rsup),
           tuple(s.domain_min), tuple(s.domain_max), s.periodic, s.mode, s.hash_map_length, s.search_range,
           want_rows, index_dtype, s.qpos.device, stream)
    plan = _plan_cache.get(key)
    if plan is None:
        plan = SearchPlan(specs, want_rows, index_dtype)
        _plan_cache[key] = plan
        if len(_plan_cache) > _PLAN_CACHE_MAX:
            _plan_cache.popitem(last=False)
    else:
        _plan_cache.move_to_end(key)
    return plan.run(csr=csr, stream=stream)


# Keys per query the multi-level count pass may keep (its single-traversal
# mode), and the bound on the slot buffer that allows.
MLM_SLOT_CAP = 128
MLM_SLOT_BYTES_MAX = 16 << 30


def run_mlm(spec: CloudSpec, want_rows: bool, index_dtype=torch.int64, max_levels: int = 4) -> "SearchResult":
    """Multi-level search of one cloud (same output as ``run_search([spec])``).
    Raises ``_lib.Unsupported`` for inputs it does not take.  The count pass
    keeps each query's keys in a slot of MLM_SLOT_CAP; when every row fits,
    the fill only sorts them (one traversal instead of two)."""
    if index_dtype not in (torch.int64, torch.int32):
        raise ValueError("index_dtype must be torch.int64 or torch.int32")
    L = _lib.load()
    device = spec.qpos.device
    (qpos, qsup, rpos, rsup), q_offs, r_offs, nq, nr = _concat([spec], device)
    clouds = _cloud_array([spec], q_offs, r_offs)
    ws_bytes = L.tcr_mlm_workspace_size(clouds)
    if ws_bytes == 0:
        check(_lib.TCR_ERR_ARG)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    offsets = torch.empty(nq + 1, dtype=torch.int64, device=device)
    # 32-bit slot keys: (stencil offset << 27 | reference index)
    cap = MLM_SLOT_CAP if 0 < nq * MLM_SLOT_CAP * 4 <= MLM_SLOT_BYTES_MAX and nr < (1 << 27) else 0
    slots = torch.empty(nq * cap, dtype=torch.int32, device=device) if cap else None
    total = ctypes.c_int64(0)
    levels = ctypes.c_int32(0)
    complete = ctypes.c_int32(0)
    stream = _stream()
    check(L.tcr_mlm_count(clouds, int(max_levels), ptr(qpos), ptr(qsup), ptr(rpos), ptr(rsup), ptr(ws), ws_bytes,
                          ptr(offsets), ctypes.byref(total), ctypes.byref(levels), ptr(slots), cap,
                          ctypes.byref(complete), stream))
    P = int(total.value)
    col = torch.empty(P, dtype=index_dtype, device=device)
    row = torch.empty(P, dtype=index_dtype, device=device) if want_rows else None
    if P > 0:
        ib = 8 if index_dtype == torch.int64 else 4
        if complete.value:
            check(L.tcr_mlm_fill(clouds, levels.value, ptr(qpos), ptr(qsup), ptr(rpos), ptr(rsup), ptr(ws), ws_bytes,
                                 ptr(offsets), ptr(row), ptr(col), ib, None, ptr(slots), cap, stream))
        else:
            scratch = col if index_dtype == torch.int64 else torch.empty(P, dtype=torch.int64, device=device)
            check(L.tcr_mlm_fill(clouds, levels.value, ptr(qpos), ptr(qsup), ptr(rpos), ptr(rsup), ptr(ws), ws_bytes,
                                 ptr(offsets), ptr(row), ptr(col), ib, ptr(scratch), None, 0, stream))
    return SearchResult(offsets, row, col, P, ws, clouds)


def run_mlm_reference(qpos: torch.Tensor, qsup: torch.Tensor, rpos: torch.Tensor, rsup: torch.Tensor,
                      domain_min: Sequence[float], domain_max: Sequence[float], pmask: int, mode: str,
                      hash_map_length: int, fmt: str):
    """algorithm='mlm' with the reference's own semantics: searchNeighbors_mlm
    (py/multiLevelMemory/wrapper.py:145-308, hash-map cells) on
    libtcr_hip.so's tcr_mlmref_* entry points.  Returns the reference's
    result tuple: SparseCOO(row, col, nq, nr) or -- its CSR quirk --
    SparseCSR(col, int32 offsets of length nr, int32 row lengths, nq, nr);
    scatter is the transposed gather search (CSR through coo_to_csr)."""
    from .util import SparseCOO, SparseCSR
    L = _lib.load()
    stream = _stream()
    dev = qpos.device
    nq, dim = qpos.shape
    nr = rpos.shape[0]
    if mode == "scatter":  # wrapper.py:191-192
        inner = run_mlm_reference(rpos, rsup, qpos, qsup, domain_min, domain_max, pmask, "gather",
                                  hash_map_length, "coo")
        P = int(inner.row.numel())
        wsb = L.tcr_mlmref_transpose_workspace_size(P, nq)
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
        out_row = torch.empty(P, dtype=torch.int64, device=dev)
        out_col = torch.empty(P, dtype=torch.int64, device=dev)
        ip = re = None
        if fmt == "csr":
            ip = torch.empty(nq + 1, dtype=torch.int32, device=dev)
            re = torch.empty(nq, dtype=torch.int64, device=dev)
        check(L.tcr_mlmref_transpose(ptr(inner.row) if P else None, ptr(inner.col) if P else None, P, nq, ptr(ws),
                                     wsb, ptr(out_row) if P else None, ptr(out_col) if P else None, ptr(ip), ptr(re),
                                     stream))
        if fmt == "coo":
            return SparseCOO(out_row, out_col, nq, nr)
        return SparseCSR(out_col, ip, re, nq, nr)
    mn = (ctypes.c_float * dim)(*[float(x) for x in domain_min[:dim]])
    mx = (ctypes.c_float * dim)(*[float(x) for x in domain_max[:dim]])
    scratch = torch.empty(64, dtype=torch.uint8, device=dev)
    grid = _lib.TcrMlmrefGrid()
    wsb = ctypes.c_size_t(0)
    check(L.tcr_mlmref_plan(ptr(rsup), nr, nq, dim, mn, mx, ptr(scratch), ctypes.byref(grid), ctypes.byref(wsb),
                            stream))
    ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
    counts = torch.empty(nr, dtype=torch.int32, device=dev)
    offsets = torch.empty(nr, dtype=torch.int32, device=dev)
    total = ctypes.c_int64(0)
    args = (ctypes.byref(grid), ptr(qpos), ptr(qsup), nq, ptr(rpos), ptr(rsup), nr, mn, mx, pmask, MODES[mode],
            int(hash_map_length), ptr(ws), wsb.value)
    check(L.tcr_mlmref_count(*args, ptr(counts), ptr(offsets), ctypes.byref(total), stream))
    if mode == "superSymmetric" and nq != nr:  # wrapper.py:245-246
        raise ValueError("SuperSymmetric support mode only works for equal number of query and reference particles")
    P = int(total.value)
    row = torch.empty(P, dtype=torch.int64, device=dev)
    col = torch.empty(P, dtype=torch.int64, device=dev)
    check(L.tcr_mlmref_fill(*args, ptr(offsets), P, ptr(row) if P else None, ptr(col) if P else None, stream))
    if fmt == "coo":
        return SparseCOO(row, col, nq, nr)
    return SparseCSR(col, offsets, counts, nq, nr)


def run_naive(qpos: torch.Tensor, qsup: Optional[torch.Tensor], rpos: torch.Tensor, rsup: Optional[torch.Tensor],
              domain_min: Sequence[float], domain_max: Sequence[float], periodic: int, wraps: int, mode: str,
              fixed_support: Optional[float] = None, index_dtype=torch.int64, small: bool = False):
    """Brute-force search on libtcr_hip.so: (ii, jj) query and reference
    indices.  small=False: algorithm='naive' (searchAlgorithms/radiusNaive.py:
    8-56), ordered by reference index, then query index; small=True:
    algorithm='small' (cpp/neighborhoodSmall_cpu.cpp, commented out of the
    reference build), ordered by query index, then reference index."""
    if index_dtype not in (torch.int64, torch.int32):
        raise ValueError("index_dtype must be torch.int64 or torch.int32")
    L = _lib.load()
    for t, name in ((qpos, "queryPointCloud.positions"), (rpos, "referencePointCloud.positions"),
                    (qsup, "queryPointCloud.supports"), (rsup, "referencePointCloud.supports")):
        if t is not None:
            _require_device(t, name)
    device = qpos.device
    nx, ny, dim = int(qpos.shape[0]), int(rpos.shape[0]), int(qpos.shape[1])
    mn = (ctypes.c_float * 3)(*[float(v) for v in domain_min[:dim]])
    mx = (ctypes.c_float * 3)(*[float(v) for v in domain_max[:dim]])
    fixed = ctypes.byref(ctypes.c_double(float(fixed_support))) if fixed_support is not None else None
    ws_bytes = L.tcr_naive_workspace_size(nx, ny, dim)
    if ws_bytes == 0:
        check(_lib.TCR_ERR_ARG)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    offsets = torch.empty((nx if small else ny) + 1, dtype=torch.int64, device=device)
    stream = _stream()
    args = (ptr(qpos), ptr(qsup), nx, ptr(rpos), ptr(rsup), ny, dim, mn, mx, int(periodic), int(wraps),
            MODES[mode], fixed, ptr(ws), ws_bytes)
    count, fill = (L.tcr_small_count, L.tcr_small_fill) if small else (L.tcr_naive_count, L.tcr_naive_fill)
    check(count(*args, ptr(offsets), stream))
    total = int(offsets[-1].item())
    ii = torch.empty(total, dtype=index_dtype, device=device)
    jj = torch.empty(total, dtype=index_dtype, device=device)
    if total > 0:
        check(fill(*args, ptr(offsets), ptr(ii), ptr(jj), 8 if index_dtype == torch.int64 else 4, stream))
    return ii, jj


def export_structure(res: SearchResult, cloud: int, spec: CloudSpec) -> CompactHashMap:
    """Copy one cloud's structure out of the workspace (returnStructure)."""
    L = _lib.load()
    c = res.clouds[cloud]
    dim, nr, Lh = c.dim, c.nr, c.hash_map_length
    dev = res.workspace.device
    sort_index = torch.empty(nr, dtype=torch.int64, device=dev)
    sorted_pos = torch.empty((nr, dim), dtype=torch.float32, device=dev)
    sorted_sup = torch.empty(nr, dtype=torch.float32, device=dev)
    hash_table = torch.empty((Lh, 2), dtype=torch.int32, device=dev)
    num_cells = torch.empty(dim, dtype=torch.int32, device=dev)
    hcell = torch.empty((), dtype=torch.float32, device=dev)
    ncells = torch.empty(1, dtype=torch.int64, device=dev)
    cell_table = torch.empty((max(nr, 1), 3), dtype=torch.int32, device=dev)
    check(L.tcr_search_export(res.clouds, len(res.clouds), cloud, ptr(res.workspace), res.workspace.numel(),
                              ptr(sort_index), ptr(sorted_pos), ptr(sorted_sup), ptr(cell_table), ptr(hash_table),
                              ptr(num_cells), ptr(hcell), ptr(ncells), _stream()))
    nc = int(ncells.item())
    mn = torch.tensor(list(spec.domain_min), dtype=torch.float32, device=dev)
    mx = torch.tensor(list(spec.domain_max), dtype=torch.float32, device=dev)
    per = torch.tensor([bool((spec.periodic >> d) & 1) for d in range(dim)], dtype=torch.bool, device=dev)
    has_sup = (spec.rsup if spec.rpos is not None else spec.qsup) is not None
    return CompactHashMap(
        sortedPositions=sorted_pos, referencePositions=spec.rpos if spec.rpos is not None else spec.qpos,
        hashTable=hash_table, hashMapLength=Lh, sortedCellTable=cell_table[:nc], numCells=num_cells,
        qMin=mn, qMax=mx, minD=mn, maxD=mx, sortIndex=sort_index, hCell=hcell, periodicity=per,
        searchRadius=spec.search_range, sortedSupports=sorted_sup if has_sup else None,
        referenceSupports=spec.rsup if spec.rpos is not None else spec.qsup)
