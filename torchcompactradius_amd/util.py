"""Public value types and helpers, mirroring torchCompactRadius 0.5.5.

References (paths relative to the reference repository):
``src/torchCompactRadius/util.py`` -- volumeToSupport (:6-26), the four
dataclasses (:280-333), coo_to_csr (:335-358), csr_to_coo (:360-370).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch


def volumeToSupport(volume: float, targetNeighbors: int, dim: int):
    """Support radius h for which a particle of ``volume`` sees
    ``targetNeighbors`` neighbours (util.py:6-26): 1-D ``N v / 2``, 2-D
    ``sqrt(N v / pi)``, 3-D ``cbrt(3 N v / (4 pi))``."""
    # the reference is TorchScript-compiled: torch.sqrt / torch.pow on Python
    # floats evaluate in double and return floats
    if dim == 1:
        return targetNeighbors * volume / 2
    elif dim == 2:
        return math.sqrt(targetNeighbors * volume / np.pi)
    else:
        return math.pow(targetNeighbors * volume / np.pi * 3 / 4, 1 / 3)


@dataclass
class DomainDescription:
    """Domain bounds per dimension and the per-dimension periodic flags
    (util.py:280-293)."""
    min: torch.Tensor
    max: torch.Tensor
    periodic: torch.Tensor
    dim: int


@dataclass
class PointCloud:
    """Positions ``[N, d]`` and optional supports ``[N]`` (util.py:295-306)."""
    positions: torch.Tensor
    supports: Optional[torch.Tensor] = None


@dataclass
class SparseCOO:
    """Neighbour list in coordinate format (util.py:308-318)."""
    row: torch.Tensor
    col: torch.Tensor
    numRows: int
    numCols: int


@dataclass
class SparseCSR:
    """Neighbour list in CSR format (util.py:321-333): ``indptr`` int32
    ``[numRows + 1]``, ``rowEntries`` int64 ``[numRows]``."""
    indices: torch.Tensor
    indptr: torch.Tensor
    rowEntries: torch.Tensor
    numRows: int
    numCols: int


def coo_to_csr(coo: SparseCOO, isSorted: bool = False) -> SparseCSR:
    """COO -> CSR (util.py:335-358): rows sorted (stable) unless isSorted,
    rowEntries int64, indptr = int64 cumsum cast to int32."""
    if not isSorted:
        order = torch.argsort(coo.row, stable=True)
        row, col = coo.row[order], coo.col[order]
    else:
        row, col = coo.row, coo.col
    nj = torch.bincount(row, minlength=coo.numRows).to(coo.row.dtype) if row.numel() else \
        torch.zeros(coo.numRows, dtype=coo.row.dtype, device=coo.row.device)
    indptr = torch.zeros(coo.numRows + 1, dtype=torch.int64, device=coo.row.device)
    indptr[1:] = torch.cumsum(nj, 0)
    return SparseCSR(col, indptr.int(), nj, coo.numRows, coo.numCols)


def csr_to_coo(csr: SparseCSR) -> SparseCOO:
    """CSR -> COO (util.py:360-370)."""
    counts = csr.rowEntries.to(torch.int64)
    row = torch.repeat_interleave(torch.arange(csr.numRows, device=csr.indices.device), counts)
    return SparseCOO(row, csr.indices, csr.numRows, csr.numCols)


def compute_h(qMin, qMax, referenceSupport) -> float:
    """Cell size of the compact grid for support hMax = referenceSupport
    (util.py:28-79), in fp32: numCells = floor(extent / hMax) per dimension,
    one cell fewer everywhere if any extent / (extent / numCells) overshoots,
    a 1.0001 stretch if any extent / h is not integral, the largest h."""
    mn = np.asarray(qMin, dtype=np.float32).reshape(-1)
    mx = np.asarray(qMax, dtype=np.float32).reshape(-1)
    hmax = np.float32(referenceSupport)
    ext = (mx - mn).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        nc = np.floor(ext / hmax).astype(np.float32)
        h = (ext / nc).astype(np.float32)
        if np.any((ext / h).astype(np.float32) - nc > np.float32(0)):
            nc = (nc - np.float32(1)).astype(np.float32)
            h = (ext / nc).astype(np.float32)
        if np.any(np.ceil((ext / h).astype(np.float32)) > (ext / h).astype(np.float32)):
            h = (h * np.float32(1.0001)).astype(np.float32)
    out = h[0]
    for v in h[1:]:
        if v > out or np.isnan(v):
            out = v
    return float(np.float32(out))
