"""Public search API: ``radiusSearch``, ``neighborSearchExisting`` and ``radius``.

Mirrors torchCompactRadius 0.5.5 ``src/torchCompactRadius/neighborhood.py``:
``radiusSearch`` (:136-309) with the same arguments, assertions and result
types; ``algorithm='compact'`` (:283) runs on libtcr_hip.so.
``neighborSearchExisting`` (:28-132) searches query points against an
exported structure.  ``radius`` (:313-365) implements the
torch_cluster-compatible behaviour the reference documents (README.md:45) --
flipped ``(i, j)`` order, per-batch search with index offsets -- on top of
the batched search.
"""
from __future__ import annotations

from typing import List, Optional, Tuple, Union

import torch

from ._lib import Unsupported
from .search import CloudSpec, CompactHashMap, export_structure, periodic_mask, run_mlm, run_search, _host_floats
from .util import DomainDescription, PointCloud, SparseCOO, SparseCSR, coo_to_csr

ALGORITHMS = ("naive", "small", "compact", "mlm")


def _domain(domain: Optional[DomainDescription], dim: int):
    """Default domain [0,1]^d, non-periodic (neighborhood.py:190-197)."""
    if domain is None:
        return [0.0] * dim, [1.0] * dim, 0
    return (_host_floats(domain.min, dim, "domain.min"), _host_floats(domain.max, dim, "domain.max"),
            periodic_mask(domain.periodic, dim))


def radiusSearch(
        queryPointCloud: PointCloud,
        referencePointCloud: Optional[PointCloud] = None,
        supportOverride: Optional[float] = None,
        mode: str = "gather",
        domain: Optional[DomainDescription] = None,
        hashMapLength=4096,
        algorithm: str = "naive",
        verbose: bool = False,
        format: str = "coo",
        returnStructure: bool = False,
        useDenseMLM: bool = False,
        hashMapLengthAlgorithm: str = "primes",
        *,
        indexDtype: torch.dtype = torch.int64,
        mlmSemantics: str = "compact",
) -> Union[SparseCOO, SparseCSR, Tuple[Union[SparseCOO, SparseCSR], CompactHashMap]]:
    """Fixed-radius neighbour search (neighborhood.py:136-309).

    Returns the (query, reference) pairs with ``dist < support`` for the
    support mode: ``gather`` h_i, ``scatter`` h_j, ``symmetric``
    (h_i + h_j) / 2, ``superSymmetric`` max(h_i, h_j).  Rows are query
    indices, columns reference indices, row-sorted; within a row the order is
    the reference compact path's (cell stencil order, then original index).

    Every algorithm runs on the GPU.  ``'naive'`` (the default) keeps the
    reference's brute-force semantics (``<=``, its own periodic ``mod``, pairs
    ordered by reference index); ``'small'`` is the brute-force op the
    reference keeps commented out (compact distance and predicate, pairs in
    (query, reference) order).  ``'cluster'`` is not accepted, as in the
    reference without torch_cluster.  ``'mlm'`` -- the multi-level search for
    variable supports -- returns exactly the compact result (pairs and order);
    for inputs it does not take (gather mode, points outside the grid, ...)
    and with ``returnStructure`` it runs the compact search.  For 'compact'
    and 'mlm', ``supportOverride`` maps a scalar support onto every point (the
    reference's fixed-support compact variant is not callable).
    ``indexDtype`` (keyword only, an extension) selects int64 row/col
    indices -- the reference's -- or int32 ones.  ``mlmSemantics``
    (keyword only): ``'compact'`` (default) as above; ``'reference'`` runs
    'mlm' with the reference's own multi-level semantics, quirks included
    (searchNeighbors_mlm, multiLevelMemory/wrapper.py:145-308): counts over
    the reference set, symmetric as gather, superSymmetric support
    synchronisation, no periodic cell stepping, level codes shifted by two
    bits per level (pairs go missing beyond level 0 in 1-D and 3-D), a CSR
    whose indptr is the length-Nr offsets array; ``returnStructure`` gives
    ``(result, None)`` and ``useDenseMLM`` is not built (ValueError).
    """
    with torch.no_grad():
        assert algorithm in ALGORITHMS, f"algorithm = {algorithm} not supported"
        assert format in ["coo", "csr"], f"format = {format} not supported"
        assert mode in ["symmetric", "scatter", "gather", "superSymmetric"], f"mode = {mode} not supported"
        assert queryPointCloud is not None, f"referencePointCloud = {queryPointCloud} is None"
        if supportOverride is not None and not isinstance(supportOverride, float):
            raise ValueError(f"supportOverride = {supportOverride} must be a float")
        if mlmSemantics not in ("compact", "reference"):
            raise ValueError(f"mlmSemantics = {mlmSemantics} must be 'compact' or 'reference'")
        if algorithm == "mlm" and mlmSemantics == "reference" and useDenseMLM:
            raise ValueError("useDenseMLM: the dense multi-level cell table is not built; "
                             "useDenseMLM=False runs the hash-map table")
        dim = queryPointCloud.positions.shape[1]
        if referencePointCloud is not None:
            assert referencePointCloud.positions.shape[1] == dim, \
                f"referencePositions.shape[1] = {referencePointCloud.positions.shape[1]} != dimensionality = {dim}"
        if domain is not None:
            assert domain.min.shape[0] == dim, f"domainMin.shape[0] = {domain.min.shape[0]} != dimensionality = {dim}"
            assert domain.max.shape[0] == dim, f"domainMax.shape[0] = {domain.max.shape[0]} != dimensionality = {dim}"
        if queryPointCloud.supports is not None:
            assert queryPointCloud.supports.shape[0] == queryPointCloud.positions.shape[0], \
                "support.shape[0] != queryPositions.shape[0]"
        if referencePointCloud is not None and referencePointCloud.supports is not None:
            assert referencePointCloud.supports.shape[0] == referencePointCloud.positions.shape[0], \
                "support.shape[0] != referencePositions.shape[0]"
        if algorithm in ("naive", "small"):
            return _brute(algorithm == "small", queryPointCloud, referencePointCloud, supportOverride, mode, domain,
                          verbose, format, returnStructure, indexDtype)

        qpos = queryPointCloud.positions
        qsup = queryPointCloud.supports
        self_search = referencePointCloud is None
        rpos = None if self_search else referencePointCloud.positions
        rsup = None if self_search else referencePointCloud.supports
        if supportOverride is not None:
            qsup = torch.full((qpos.shape[0],), supportOverride, dtype=torch.float32, device=qpos.device)
            if not self_search:
                rsup = torch.full((rpos.shape[0],), supportOverride, dtype=torch.float32, device=rpos.device)
        from .search import _f32
        qpos, qsup = _f32(qpos, "queryPointCloud.positions"), _f32(qsup, "queryPointCloud.supports")
        rpos, rsup = _f32(rpos, "referencePointCloud.positions"), _f32(rsup, "referencePointCloud.supports")
        mn, mx, pmask = _domain(domain, dim)
        if verbose:
            print(f"radiusSearch(compact): nq={qpos.shape[0]} nr={qpos.shape[0] if rpos is None else rpos.shape[0]} "
                  f"dim={dim} mode={mode} periodic_mask={pmask} hashMapLength={hashMapLength} format={format}")
        if algorithm == "mlm" and mlmSemantics == "reference":
            if qsup is None or (not self_search and rsup is None):
                raise ValueError("the multi-level search reads the supports of both point clouds")
            from .search import run_mlm_reference
            res = run_mlm_reference(qpos, qsup, qpos if self_search else rpos, qsup if self_search else rsup,
                                    mn, mx, pmask, mode, int(hashMapLength), format)
            return (res, None) if returnStructure else res
        spec = CloudSpec(qpos=qpos, qsup=qsup, rpos=rpos, rsup=rsup, domain_min=mn, domain_max=mx,
                         periodic=pmask, mode=mode, hash_map_length=int(hashMapLength))
        res = None
        if algorithm == "mlm" and not returnStructure:
            try:
                res = run_mlm(spec, want_rows=(format == "coo"), index_dtype=indexDtype)
            except Unsupported:
                res = None      # the single-level search gives the same output
        if res is None:
            res = run_search([spec], want_rows=(format == "coo"), index_dtype=indexDtype, csr=(format == "csr"))
        nq = qpos.shape[0]
        nr = nq if rpos is None else rpos.shape[0]
        if format == "coo":
            out: Union[SparseCOO, SparseCSR] = SparseCOO(res.row, res.col, nq, nr)
        elif res.indptr is not None:
            out = SparseCSR(res.col, res.indptr, res.row_entries, nq, nr)
        else:
            out = SparseCSR(res.col, res.offsets.int(), res.offsets[1:] - res.offsets[:-1], nq, nr)
        if returnStructure:
            return out, export_structure(res, 0, spec)
        return out


def _brute(small: bool, queryPointCloud: PointCloud, referencePointCloud: Optional[PointCloud],
           supportOverride: Optional[float], mode: str, domain: Optional[DomainDescription], verbose: bool,
           format: str, returnStructure: bool, indexDtype: torch.dtype):
    """radiusSearch's brute-force branches (neighborhood.py:186-279, 298-309);
    without a domain the default [0,1]^d is not periodic, and with one the
    periodic coordinates are wrapped once here (getPeriodicPointCloud).

    'naive' (radiusNaive.py:8-56): radiusNaive wraps them once more; pairs in
    its order -- by reference index, then query index -- and the CSR form is
    coo_to_csr(isSorted=True) of them, as the reference builds it (its indices
    in that order too).  supportOverride selects radiusNaiveFixed (squared
    distance <= support^2), which the reference dispatches to (:226-229) but
    cannot run (it passes PointCloud records where the scripted function takes
    tensors).

    'small' (cpp/neighborhoodSmall_cpu.cpp:45-302, commented out of the
    reference build, whose binding is None -- the reference raises there): the
    compact path's distance and strict predicate over all pairs, in (query,
    reference) order; supportOverride selects its fixed variant (modDistance2 <
    support^2 in fp32)."""
    from .search import _f32, run_naive
    qpos = _f32(queryPointCloud.positions, "queryPointCloud.positions")
    qsup = _f32(queryPointCloud.supports, "queryPointCloud.supports")
    if referencePointCloud is None:
        rpos, rsup = qpos, qsup
    else:
        rpos = _f32(referencePointCloud.positions, "referencePointCloud.positions")
        rsup = _f32(referencePointCloud.supports, "referencePointCloud.supports")
    dim = qpos.shape[1]
    mn, mx, pmask = _domain(domain, dim)
    wraps = 0 if domain is None else (1 if small else 2)
    if verbose:
        name = "neighborSearchSmall" if small else "radiusNaive"
        print(f"Calling {name}{'Fixed' if supportOverride is not None else ''}")
    ii, jj = run_naive(qpos, qsup, rpos, rsup, mn, mx, pmask, wraps, mode, supportOverride, indexDtype, small)
    out: Union[SparseCOO, SparseCSR] = SparseCOO(ii, jj, qpos.shape[0], rpos.shape[0])
    if format == "csr":
        out = coo_to_csr(out, isSorted=True)
    if returnStructure:
        return out, None
    return out


def neighborSearchExisting(
        queryPointCloud: PointCloud,
        hashMap: CompactHashMap, mode: str = "symmetric", searchRadius: int = 1, variant: str = "cpp",
        verbose: bool = False, format: str = "coo") -> Union[SparseCOO, SparseCSR]:
    """Search query points against an existing structure (neighborhood.py:28-132),
    e.g. the ``CompactHashMap`` of ``radiusSearch(..., returnStructure=True)``.

    Runs searchNeighbors_cpp's sequence (compactHashing/cppSearch.py:60-98)
    on the structure's tensors through the reference-layout ops of
    libtcr_hip.so: countNeighbors, int32 exclusive offsets,
    buildNeighborList, ``j = sortIndex[j]``; rows are query indices, columns
    reference indices.  Query positions are used as given: ``radiusSearch``
    wraps its queries into a periodic domain once before searching, so wrapped
    positions reproduce its result.

    The reference function reads ``hashMap.periodic`` (neighborhood.py:81), a
    field its record does not have (compactHashing/datastructure.py:26 names it
    ``periodicity``), so there every call raises; this is the search that code
    describes.  A float support -- ``queryPointCloud.supports`` given as a
    float, or ``hashMap.fixedSupport`` when the query has no supports -- is
    applied as that support to every query and reference point (the
    reference's fixed-support ops are commented out of its build).  The
    per-point Python variants ('python', 'pythonDynamic') are not provided.
    """
    with torch.no_grad():
        if variant in ("python", "pythonDynamic"):
            raise NotImplementedError(f"variant='{variant}' is not provided by torchcompactradius_amd; use 'cpp'")
        if variant != "cpp":
            raise ValueError("variant must be either cpp, python or pythonDynamic")
        assert format in ["coo", "csr"], f"format = {format} not supported"
        from . import native_ops
        qpos = queryPointCloud.positions
        sup = queryPointCloud.supports
        fixed: Optional[float] = None
        if sup is None:
            qsup = None
            fixed = hashMap.fixedSupport
        elif isinstance(sup, torch.Tensor):
            qsup = sup
        else:
            qsup = None
            fixed = float(sup)
        sorted_pos = hashMap.sortedPositions
        sorted_sup = hashMap.sortedSupports
        nq, nr = int(qpos.shape[0]), int(sorted_pos.shape[0])
        if fixed is not None:
            qsup = torch.full((nq,), fixed, dtype=torch.float32, device=qpos.device)
            sorted_sup = torch.full((nr,), fixed, dtype=torch.float32, device=sorted_pos.device)
        hcell = float(hashMap.hCell.item()) if isinstance(hashMap.hCell, torch.Tensor) else float(hashMap.hCell)
        args = (qpos, qsup, searchRadius, sorted_pos, sorted_sup, hashMap.hashTable, hashMap.hashMapLength,
                hashMap.numCells, hashMap.sortedCellTable, hashMap.qMin, hcell, hashMap.maxD, hashMap.minD,
                hashMap.periodicity, mode, verbose)
        if verbose:
            print(f"neighborSearchExisting: nq={nq} nr={nr} mode={mode} searchRadius={searchRadius} "
                  f"fixedSupport={fixed} format={format}")
        counts = native_ops.countNeighbors(*args)
        offsets = torch.zeros(nq, dtype=torch.int32, device=qpos.device)
        if nq > 1:
            torch.cumsum(counts[:-1], 0, dtype=torch.int32, out=offsets[1:])
        total = int(offsets[-1].item() + counts[-1].item()) if nq else 0
        i, j = native_ops.buildNeighborList(counts, offsets, total, *args)
        j = hashMap.sortIndex[j]
        sparse = SparseCOO(i, j, nq, nr)
        if format == "csr":
            return coo_to_csr(sparse, isSorted=True)
        return sparse


def radius(queryPositions: torch.Tensor,
           referencePositions: Optional[torch.Tensor],
           support: Union[float, torch.Tensor, Tuple[torch.Tensor, torch.Tensor]],
           batch_x: Optional[torch.Tensor] = None, batch_y: Optional[torch.Tensor] = None,
           mode: str = "gather",
           domain: Optional[DomainDescription] = None,
           hashMapLength=4096,
           algorithm: str = "compact",
           verbose: bool = False,
           returnStructure: bool = False):
    """torch_cluster-style radius (neighborhood.py:313-365, README.md:45):
    returns ``(i, j)`` with ``i`` indexing the reference set and ``j`` the
    query set (flipped relative to radiusSearch).  With ``batch_x`` /
    ``batch_y``, each batch id is searched separately and indices are offset
    into the concatenated sets: for 'compact' all ids in one batched launch,
    for the other algorithms one radiusSearch per id.  ``algorithm`` is
    radiusSearch's (the reference's own radius passes its arguments to
    radiusSearch out of order and cannot run, so its default 'naive' defines
    no result; 'compact' is the default here)."""
    with torch.no_grad():
        if isinstance(support, tuple):
            qsup, rsup = support
        elif isinstance(support, torch.Tensor):
            qsup, rsup = support, support if referencePositions is None else None
        else:
            qsup = torch.full((queryPositions.shape[0],), float(support), device=queryPositions.device)
            rsup = None if referencePositions is None else torch.full(
                (referencePositions.shape[0],), float(support), device=referencePositions.device)
        dim = queryPositions.shape[1]
        mn, mx, pmask = _domain(domain, dim)
        if batch_x is None and batch_y is None:
            ref = None if referencePositions is None else PointCloud(referencePositions, rsup)
            res = radiusSearch(PointCloud(queryPositions, qsup), ref, mode=mode, domain=domain,
                               hashMapLength=hashMapLength, algorithm=algorithm, format="coo",
                               returnStructure=returnStructure)
            coo = res[0] if returnStructure else res
            if returnStructure:
                return coo.col, coo.row, res[1]
            return coo.col, coo.row
        ref_pos = queryPositions if referencePositions is None else referencePositions
        ref_sup = qsup if referencePositions is None else rsup
        bx = batch_x if batch_x is not None else torch.zeros(queryPositions.shape[0], dtype=torch.long,
                                                             device=queryPositions.device)
        by = batch_y if batch_y is not None else (bx if referencePositions is None else torch.zeros(
            ref_pos.shape[0], dtype=torch.long, device=ref_pos.device))
        ids = torch.unique(torch.cat([bx, by])).tolist()
        if algorithm != "compact":
            assert algorithm in ALGORITHMS, f"algorithm = {algorithm} not supported"
            cols, rows = [], []
            for b in ids:
                qi = torch.nonzero(bx == b).reshape(-1)
                ri = torch.nonzero(by == b).reshape(-1)
                out = radiusSearch(PointCloud(queryPositions[qi].contiguous(),
                                              None if qsup is None else qsup[qi].contiguous()),
                                   PointCloud(ref_pos[ri].contiguous(),
                                              None if ref_sup is None else ref_sup[ri].contiguous()),
                                   mode=mode, domain=domain, hashMapLength=hashMapLength, algorithm=algorithm,
                                   format="coo")
                rows.append(qi[out.row])
                cols.append(ri[out.col])
            empty = torch.empty(0, dtype=torch.long, device=queryPositions.device)
            cols_global = torch.cat(cols) if cols else empty
            rows_global = torch.cat(rows) if rows else empty
            if returnStructure:
                return cols_global, rows_global, None
            return cols_global, rows_global
        specs, qidx, ridx = [], [], []
        for b in ids:
            qi = torch.nonzero(bx == b).reshape(-1)
            ri = torch.nonzero(by == b).reshape(-1)
            qidx.append(qi)
            ridx.append(ri)
            specs.append(CloudSpec(qpos=queryPositions[qi].float().contiguous(),
                                   qsup=None if qsup is None else qsup[qi].float().contiguous(),
                                   rpos=ref_pos[ri].float().contiguous(),
                                   rsup=None if ref_sup is None else ref_sup[ri].float().contiguous(),
                                   domain_min=mn, domain_max=mx, periodic=pmask, mode=mode,
                                   hash_map_length=int(hashMapLength)))
        res = run_search(specs, want_rows=True)
        # rows/cols are local to each cloud; map them back to global indices
        counts = res.offsets[1:] - res.offsets[:-1]
        qglob = torch.cat(qidx)
        rows_global = qglob[torch.repeat_interleave(torch.arange(qglob.shape[0], device=qglob.device), counts)]
        cloud_of_pair = torch.repeat_interleave(
            torch.arange(len(specs), device=qglob.device),
            torch.tensor([int(s.qpos.shape[0]) for s in specs], device=qglob.device))
        cloud_of_pair = torch.repeat_interleave(cloud_of_pair, counts)
        r_base = torch.tensor([0] + [int(r.shape[0]) for r in ridx], device=qglob.device).cumsum(0)[:-1]
        rglob_all = torch.cat(ridx)
        cols_global = rglob_all[res.col + r_base[cloud_of_pair]]
        if returnStructure:
            return cols_global, rows_global, None
        return cols_global, rows_global
