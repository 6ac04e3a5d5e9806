"""The reference's native op layer on the reference's own structure layout.

Same names, argument lists and meaning as the torch ops the reference
registers (cpp/cppWrapper.cpp:33-59, signatures at cpp/neighborhood.h:102-124
and cpp/hashing.h:38), and as ``countNeighbors_cpp`` / ``buildNeighborList_cpp``
/ ``hashCells_cpp`` of compactHashing/cppWrapper.py:17-21, so the reference's
``searchNeighbors_cpp`` (compactHashing/cppSearch.py:12-98) runs unchanged on
top of them.  Compute goes through libtcr_hip.so on torch's current stream;
inputs must be contiguous HIP tensors, and invalid inputs raise
``RuntimeError`` as the reference's ``std::runtime_error`` does
(cpp/neighborhoodDynamic_cpu.cpp:18-40,86-96).
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _lib
from ._lib import MODES, ptr


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _need(t: Optional[torch.Tensor], name: str, dtype, rank: Optional[int] = None, optional: bool = False):
    if t is None:
        if optional:
            return None
        raise RuntimeError(f"{name} is undefined")
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be on a HIP device (this library has no CPU path)")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype} (got {t.dtype})")
    if rank is not None and t.dim() != rank:
        raise RuntimeError(f"{name} must have rank {rank} (got {t.dim()})")
    return t


def _mode(mode: str) -> int:
    if mode not in MODES:
        raise RuntimeError(f"Invalid support mode: {mode}")
    return MODES[mode]


def _check(rc: int) -> None:
    if rc != _lib.TCR_OK:
        raise RuntimeError(_lib.load().tcr_last_error().decode(errors="replace"))


def _common(queryPositions, querySupport, sortedPositions, sortedSupport, hashTable, numCells, cellTable,
            qMin, maxDomain, minDomain, periodicity, mode):
    _need(queryPositions, "queryPositions", torch.float32, 2)
    _need(sortedPositions, "sortedPositions", torch.float32, 2)
    dim = queryPositions.shape[1]
    if dim not in (1, 2, 3) or sortedPositions.shape[1] != dim:
        raise RuntimeError(f"Unsupported dimension: {dim}")
    m = _mode(mode)
    _need(querySupport, "querySupport", torch.float32, 1, optional=m == MODES["scatter"])
    _need(sortedSupport, "sortedSupport", torch.float32, 1, optional=m == MODES["gather"])
    _need(hashTable, "hashTable", torch.int32, 2)
    _need(numCells, "numCells", torch.int32, 1)
    _need(cellTable, "cellTable", torch.int32, 2)
    _need(maxDomain, "maxDomain", torch.float32, 1)
    _need(minDomain, "minDomain", torch.float32, 1)
    _need(periodicity, "periodicity", torch.bool, 1)
    if qMin is not None:
        _need(qMin, "qMin", torch.float32, 1)
    return dim, m


def countNeighbors(queryPositions: torch.Tensor, querySupport: Optional[torch.Tensor], searchRange: int,
                   sortedPositions: torch.Tensor, sortedSupport: Optional[torch.Tensor],
                   hashTable: torch.Tensor, hashMapLength: int, numCells: torch.Tensor, cellTable: torch.Tensor,
                   qMin: torch.Tensor, hCell: float, maxDomain: torch.Tensor, minDomain: torch.Tensor,
                   periodicity: torch.Tensor, mode: str, verbose: bool = False) -> torch.Tensor:
    """countNeighbors (cpp/neighborhood.h:102-108): int32 [Nq] neighbour
    counts of each query against the sorted reference structure."""
    dim, m = _common(queryPositions, querySupport, sortedPositions, sortedSupport, hashTable, numCells, cellTable,
                     qMin, maxDomain, minDomain, periodicity, mode)
    nq = queryPositions.shape[0]
    counts = torch.zeros(nq, dtype=torch.int32, device=queryPositions.device)
    L = _lib.load()
    _check(L.tcr_count_neighbors(ptr(queryPositions), ptr(querySupport), nq, int(searchRange),
                                 ptr(sortedPositions), ptr(sortedSupport), sortedPositions.shape[0],
                                 ptr(hashTable), int(hashMapLength), ptr(numCells), ptr(cellTable),
                                 cellTable.shape[0], ptr(qMin), float(hCell), ptr(maxDomain), ptr(minDomain),
                                 ptr(periodicity), dim, m, ptr(counts), _stream()))
    return counts


def buildNeighborList(neighborCounter: torch.Tensor, neighborOffsets: torch.Tensor, neighborListLength: int,
                      queryPositions: torch.Tensor, querySupport: Optional[torch.Tensor], searchRange: int,
                      sortedPositions: torch.Tensor, sortedSupport: Optional[torch.Tensor],
                      hashTable: torch.Tensor, hashMapLength: int, numCells: torch.Tensor, cellTable: torch.Tensor,
                      qMin: torch.Tensor, hCell: float, maxDomain: torch.Tensor, minDomain: torch.Tensor,
                      periodicity: torch.Tensor, mode: str, verbose: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """buildNeighborList (cpp/neighborhood.h:117-124): (i, j_sorted) int64
    [neighborListLength], query i's pairs at neighborOffsets[i] in the
    traversal order."""
    dim, m = _common(queryPositions, querySupport, sortedPositions, sortedSupport, hashTable, numCells, cellTable,
                     qMin, maxDomain, minDomain, periodicity, mode)
    _need(neighborCounter, "neighborCounter", torch.int32, 1)
    _need(neighborOffsets, "neighborOffsets", torch.int32, 1)
    n = int(neighborListLength)
    if n < 0:
        raise RuntimeError("neighborListLength must be >= 0")
    dev = queryPositions.device
    li = torch.zeros(n, dtype=torch.int64, device=dev)
    lj = torch.zeros(n, dtype=torch.int64, device=dev)
    L = _lib.load()
    _check(L.tcr_build_neighbor_list(ptr(neighborCounter), ptr(neighborOffsets), n, ptr(queryPositions),
                                     ptr(querySupport), queryPositions.shape[0], int(searchRange),
                                     ptr(sortedPositions), ptr(sortedSupport), sortedPositions.shape[0],
                                     ptr(hashTable), int(hashMapLength), ptr(numCells), ptr(cellTable),
                                     cellTable.shape[0], ptr(qMin), float(hCell), ptr(maxDomain), ptr(minDomain),
                                     ptr(periodicity), dim, m, ptr(li), ptr(lj), _stream()))
    return li, lj


def computeHashIndices(cellIndices: torch.Tensor, hashMapLength: int) -> torch.Tensor:
    """computeHashIndices (cpp/hashing.h:38): int32 [C] bucket of each int32
    cell coordinate row."""
    _need(cellIndices, "cellIndices", torch.int32, 2)
    n, dim = cellIndices.shape
    out = torch.zeros(n, dtype=torch.int32, device=cellIndices.device)
    L = _lib.load()
    _check(L.tcr_compute_hash_indices(ptr(cellIndices), n, dim, int(hashMapLength), ptr(out), _stream()))
    return out
