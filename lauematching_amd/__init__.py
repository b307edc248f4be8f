"""MI355X-native LaueMatching indexing hot path (forward simulation + compare).

The compute path is liblauematch.so (HIP, gfx950) behind the C ABI in
include/lauematch.h; this package is its Python host interface.
"""
from .api import (LM_FWD_FP32_GUARD, LM_FWD_FP64, Context, Geometry, Hits, OrientationDB,
                  quantize_u8, read_params)

__all__ = ["Context", "Geometry", "Hits", "OrientationDB", "quantize_u8", "read_params",
           "LM_FWD_FP64", "LM_FWD_FP32_GUARD"]
