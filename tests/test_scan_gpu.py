"""GPU parity of the offsets scan at its tile boundaries: query counts around
multiples of the 8192-count scan tile (one tile, a full tile, a tile and one
query, several tiles), and a search with more tiles than the GPU holds at
once (their prefixes chain through tiles that have already finished).  The
CSR indptr/rowEntries written by the scan and the pairs are checked against
the oracle."""
import numpy as np
import pytest
import torch

import oracle
from torchcompactradius_amd import DomainDescription, PointCloud, radiusSearch

pytestmark = pytest.mark.gpu

DEV = "cuda"
TILE = 8192


def _search_csr(pos, sup, mn, mx, periodic):
    dom = DomainDescription(torch.tensor(mn, dtype=torch.float32), torch.tensor(mx, dtype=torch.float32),
                            torch.tensor(periodic, dtype=torch.bool), len(mn))
    out = radiusSearch(PointCloud(torch.from_numpy(pos).to(DEV), torch.from_numpy(sup).to(DEV)), None,
                       mode="gather", domain=dom, hashMapLength=4096, algorithm="compact", format="csr")
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("n", [1, 2, TILE - 1, TILE, TILE + 1, 3 * TILE + 17])
def test_offsets_at_tile_boundaries(n):
    rng = np.random.default_rng(30 + n % 97)
    mn, mx, periodic = [0.0, 0.0], [1.0, 1.0], [True, False]
    pos = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    sup = np.full(n, np.float32(min(0.3, 3.0 / np.sqrt(max(n, 1)))), dtype=np.float32)
    print(f"case: offsets scan n={n}", flush=True)
    out = _search_csr(pos, sup, mn, mx, periodic)
    o = oracle.radius_search(pos, sup, domain_min=mn, domain_max=mx, periodic=periodic, mode="gather",
                             hash_map_length=4096)
    np.testing.assert_array_equal(out.indptr.cpu().numpy(), o.indptr.astype(np.int32))
    np.testing.assert_array_equal(out.rowEntries.cpu().numpy(), np.diff(o.indptr).astype(np.int64))
    np.testing.assert_array_equal(out.indices.cpu().numpy(), o.col)


def test_offsets_more_tiles_than_resident():
    """6M queries (733 scan tiles; two 64 KB-LDS tiles fit a compute unit, so
    the later tiles start only after earlier ones have finished)."""
    rng = np.random.default_rng(41)
    side = 2450   # 6.0M points on a jittered 2-D grid
    g = (np.arange(side, dtype=np.float32) + np.float32(0.5)) / np.float32(side)
    x, y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel()], 1).astype(np.float32)
    pos += rng.uniform(-0.2, 0.2, pos.shape).astype(np.float32) / np.float32(side)
    n = pos.shape[0]
    sup = np.full(n, np.float32(1.6 / side), dtype=np.float32)   # a handful of neighbours each
    mn, mx, periodic = [0.0, 0.0], [1.0, 1.0], [True, True]
    print(f"case: offsets scan n={n}", flush=True)
    out = _search_csr(pos, sup, mn, mx, periodic)
    o = oracle.radius_search(pos, sup, domain_min=mn, domain_max=mx, periodic=periodic, mode="gather",
                             hash_map_length=4096)
    assert o.col.shape[0] > 2 * n
    np.testing.assert_array_equal(out.indptr.cpu().numpy(), o.indptr.astype(np.int32))
    np.testing.assert_array_equal(out.indices.cpu().numpy(), o.col)
