"""torchcompactradius_amd -- MI355X-native compact-hashing radius search.

A drop-in for torchCompactRadius's ``radiusSearch(..., algorithm='compact')``:
same API types and results, computed by hand-written gfx950 HIP kernels in
``libtcr_hip.so`` (C ABI in ``include/tcr.h``).
"""
from .neighborhood import neighborSearchExisting, radius, radiusSearch
from .search import CompactHashMap
from .util import (DomainDescription, PointCloud, SparseCOO, SparseCSR, coo_to_csr, csr_to_coo,
                   volumeToSupport)

__version__ = "0.1.0"

__all__ = ["radiusSearch", "radius", "neighborSearchExisting", "DomainDescription", "PointCloud", "SparseCOO", "SparseCSR",
           "coo_to_csr", "csr_to_coo", "volumeToSupport", "CompactHashMap"]
