"""CPU checks of the host-side helpers that mirror the reference's util
module."""
import numpy as np

import oracle
from torchcompactradius_amd.util import compute_h


def test_compute_h_matches_oracle_on_random_domains():
    rng = np.random.default_rng(0)
    for _ in range(5000):
        dim = int(rng.integers(1, 4))
        mn = rng.uniform(-3, 1, dim).astype(np.float32)
        mx = (mn + rng.uniform(0.1, 5, dim)).astype(np.float32)
        h = np.float32(rng.uniform(0.001, 1.0))
        a, b = np.float32(compute_h(mn, mx, h)), np.float32(oracle.compute_h(mn, mx, h))
        assert a == b or (np.isnan(a) and np.isnan(b))


def test_compute_h_workload_values():
    # 2-D 1024-point workload: hCell 0.2 (10 x 10 cells); 1M-point 3-D: 0.046512
    assert np.float32(compute_h([-1, -1], [1, 1], np.float32(np.sqrt(32 * (2 / 32) ** 2 / np.pi)))) == np.float32(0.2)
    h = np.float32((3 * 50 * 0.02 ** 3 / (4 * np.pi)) ** (1 / 3))
    assert abs(compute_h([-1] * 3, [1] * 3, h) - 0.046512) < 1e-6


def test_search_existing_variant_errors():
    """neighborSearchExisting's variant dispatch (neighborhood.py:78-108):
    unknown variants raise ValueError before any device work; the per-point
    Python variants are not provided."""
    import pytest
    import torch
    from torchcompactradius_amd import PointCloud, neighborSearchExisting
    q = PointCloud(torch.zeros((4, 3)), torch.ones(4))
    with pytest.raises(ValueError):
        neighborSearchExisting(q, None, variant="fortran")
    for v in ("python", "pythonDynamic"):
        with pytest.raises(NotImplementedError):
            neighborSearchExisting(q, None, variant=v)


def test_mlm_semantics_flag_validation():
    """radiusSearch's mlmSemantics switch: unknown values and the dense
    multi-level table (not built) raise before any device work."""
    import pytest
    import torch
    from torchcompactradius_amd import PointCloud, radiusSearch
    p = torch.zeros((4, 2))
    h = torch.ones(4)
    with pytest.raises(ValueError):
        radiusSearch(PointCloud(p, h), algorithm="mlm", mlmSemantics="bogus")
    with pytest.raises(ValueError):
        radiusSearch(PointCloud(p, h), algorithm="mlm", mlmSemantics="reference", useDenseMLM=True)
