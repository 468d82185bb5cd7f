"""mojovec_amd: MI355X-native (gfx950) batched vector-scan path for MojoVec.

Exact Flat and SQ8 search with top-k, IVF coarse assignment, PQ LUT build
with inverted-list ADC scan and the Lloyd k-means that trains them, as HIP
kernels behind the C ABI in include/mojovec_gpu.h. The Python layer mirrors
MojoVec's IVF-PQ collection facade and its index classes.
"""
from ._lib import MojoVecError, load as _load_native
from ._ivfpq import IVFPQCollection, IVFPQStats, QueryResult
from .index import (
    METRIC_INNER_PRODUCT,
    METRIC_L2,
    IndexFlat,
    IndexFlatSQ8,
    IndexIVFFlat,
    IndexIVFPQ,
    KMeans,
    ProductQuantizer,
)

from .serving import Server

__version__ = "0.1.0"

_load_native()

__all__ = [
    "IVFPQCollection",
    "IVFPQStats",
    "QueryResult",
    "IndexFlat",
    "IndexFlatSQ8",
    "IndexIVFFlat",
    "IndexIVFPQ",
    "KMeans",
    "ProductQuantizer",
    "METRIC_L2",
    "METRIC_INNER_PRODUCT",
    "MojoVecError",
    "Server",
]
