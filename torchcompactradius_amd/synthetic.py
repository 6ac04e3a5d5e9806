"""Seeded synthetic point clouds for the benchmark and parity workloads.

All clouds are generated on the host in float64 with
``numpy.random.default_rng(seed)`` and cast to float32, so the GPU search and
the CPU oracle receive identical bits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Workload:
    name: str
    pos: np.ndarray                 # float32 [N, d]
    sup: np.ndarray                 # float32 [N]
    domain_min: np.ndarray          # float32 [d]
    domain_max: np.ndarray          # float32 [d]
    periodic: List[bool]
    mode: str
    hash_map_length: int
    format: str
    rpos: Optional[np.ndarray] = None   # None: self-search
    rsup: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)


def jittered_grid(n_per_dim: int, dim: int, lo: float, hi: float, jitter_frac: float,
                  rng: np.random.Generator) -> np.ndarray:
    """Cell centres of an n^dim grid over [lo, hi]^dim, each jittered by
    U(+-jitter_frac * dx); index order is row-major with dim 0 slowest."""
    dx = (hi - lo) / n_per_dim
    g = lo + (np.arange(n_per_dim) + 0.5) * dx
    mesh = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), axis=-1).reshape(-1, dim)
    mesh = mesh + rng.uniform(-jitter_frac * dx, jitter_frac * dx, size=mesh.shape)
    return mesh.astype(np.float32)


def support_for(neighbors: float, dx: float, dim: int) -> float:
    if dim == 1:
        return neighbors * dx / 2
    if dim == 2:
        return math.sqrt(neighbors * dx * dx / math.pi)
    return (3.0 * neighbors * dx ** 3 / (4.0 * math.pi)) ** (1.0 / 3.0)


def plain_2d_1024(seed: int = 1) -> Workload:
    """2-D 32x32 jittered grid in [-1,1]^2 (1024 points), ~32 neighbours,
    non-periodic, gather, hashMapLength 4096, COO."""
    rng = np.random.default_rng(seed)
    n, dim = 32, 2
    dx = 2.0 / n
    pos = jittered_grid(n, dim, -1.0, 1.0, 0.05, rng)
    h = np.float32(math.sqrt(32 * dx * dx / math.pi))
    return Workload("plain-2d-1024", pos, np.full(pos.shape[0], h, np.float32),
                    np.full(dim, -1, np.float32), np.full(dim, 1, np.float32), [False] * dim,
                    "gather", 4096, "coo", meta={"dx": dx, "h": float(h)})


def periodic_3d(n: int = 100, seed: int = 2, neighbors: float = 50.0) -> Workload:
    """3-D n^3 jittered grid in [-1,1]^3 (n=100: 1M points), jitter
    +-0.1 dx, uniform h for ~50 neighbours, periodic in every dim, gather,
    hashMapLength 2^20, CSR."""
    rng = np.random.default_rng(seed)
    dim = 3
    dx = 2.0 / n
    pos = jittered_grid(n, dim, -1.0, 1.0, 0.1, rng)
    h = np.float32(support_for(neighbors, dx, dim))
    return Workload(f"periodic-3d-{n**3}", pos, np.full(pos.shape[0], h, np.float32),
                    np.full(dim, -1, np.float32), np.full(dim, 1, np.float32), [True] * dim,
                    "gather", 1 << 20, "csr", meta={"dx": dx, "h": float(h), "n": n})


def batched_clouds(seed: int = 4, n_clouds: int = 256, n_min: float = 1e3, n_max: float = 2e5,
                   scale: float = 1.0) -> List[Workload]:
    """Independent clouds for one batched launch: sizes log-uniform in
    [n_min, n_max] (times ``scale``), alternating 2-D / 3-D jittered grids in
    [-1,1]^d (jitter +-0.1 dx), per-cloud support for U{20..60} neighbours,
    per-dimension periodicity ~ Bernoulli(0.5), gather, CSR."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_clouds):
        dim = 2 if k % 2 == 0 else 3
        N = math.exp(rng.uniform(math.log(n_min), math.log(n_max))) * scale
        n = max(2, int(round(N ** (1.0 / dim))))
        dx = 2.0 / n
        pos = jittered_grid(n, dim, -1.0, 1.0, 0.1, rng)
        neighbors = int(rng.integers(20, 61))
        h = np.float32(support_for(neighbors, dx, dim))
        periodic = [bool(x) for x in (rng.random(dim) < 0.5)]
        out.append(Workload(f"batched-{k}-{dim}d-{pos.shape[0]}", pos, np.full(pos.shape[0], h, np.float32),
                            np.full(dim, -1, np.float32), np.full(dim, 1, np.float32), periodic,
                            "gather", 1 << 16, "csr", meta={"dx": dx, "h": float(h), "neighbors": neighbors}))
    return out


def adaptive_3d(n: int = 2_000_000, seed: int = 3, neighbors: int = 50, ratio: float = 8.0,
                clusters: int = 8) -> Workload:
    """Adaptive-support 3-D cloud: 70% of the points from `clusters` isotropic
    Gaussians (sigma log-uniform in [0.02, 0.2], centres U[-0.7, 0.7]^3,
    rejection-sampled into [-1, 1]^3) and 30% uniform in [-1, 1]^3; each
    point's support is its distance to the `neighbors`-th nearest other point
    (k-NN density estimate), clipped to [q, `ratio` q] with q the 1st
    percentile of those distances.  Symmetric mode, non-periodic, COO,
    self-search."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    n_g = int(0.7 * n)
    centres = rng.uniform(-0.7, 0.7, (clusters, 3))
    sig = np.exp(rng.uniform(math.log(0.02), math.log(0.2), clusters))
    parts = []
    for c in range(clusters):
        m = n_g // clusters + (1 if c < n_g % clusters else 0)
        got = np.empty((0, 3))
        while got.shape[0] < m:
            x = centres[c] + sig[c] * rng.standard_normal((2 * (m - got.shape[0]) + 16, 3))
            x = x[np.all(np.abs(x) <= 1.0, axis=1)]
            got = np.concatenate([got, x])
        parts.append(got[:m])
    parts.append(rng.uniform(-1.0, 1.0, (n - n_g, 3)))
    pos = np.concatenate(parts).astype(np.float32)
    d, _ = cKDTree(pos.astype(np.float64)).query(pos.astype(np.float64), k=neighbors + 1, workers=-1)
    h = d[:, -1]
    q = np.percentile(h, 1.0)
    h = np.clip(h, q, ratio * q).astype(np.float32)
    return Workload(f"adaptive-3d-{n}", pos, h, np.full(3, -1, np.float32), np.full(3, 1, np.float32),
                    [False] * 3, "symmetric", 1 << 20, "coo",
                    meta={"h_min": float(h.min()), "h_max": float(h.max()), "neighbors": neighbors})
