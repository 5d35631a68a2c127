"""partconv_amd: multi-level non-uniform partitioned convolution on AMD MI355X.

Python surface over libpartconv_amd.so (C ABI in include/partconv_amd.h).
"""
from ._lib import last_error, lib  # noqa: F401
from .partconv import (PartConvFilter, PartConvMax, PartConvMulti, PartConvMultiRelaxed,  # noqa: F401
                       doubleArray, greedy_partition, intArray, readImpulseResponse)

__all__ = ["PartConvFilter", "PartConvMax", "PartConvMulti", "PartConvMultiRelaxed", "intArray", "doubleArray",
           "greedy_partition", "readImpulseResponse", "last_error", "lib"]
