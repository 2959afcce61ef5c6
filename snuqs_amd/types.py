"""Execution options (src/braket/snuqs/types.py of the reference).

AcceleratorType.CUDA keeps the reference's name for the GPU and means the HIP
device here; HIP is an alias.  CPU and HYBRID, and the CPU / STORAGE offload
tiers, are parsed so existing callers get a clear error rather than a
different code path: the engine has one execution path, on the GPU.
"""
from enum import Enum


class AcceleratorType(Enum):
    CPU = 0
    CUDA = 1
    HYBRID = 2
    HIP = 3


class PrefetchType(Enum):
    NONE = 0
    CUDA = 1
    STORAGE = 2


class OffloadType(Enum):
    NONE = 0
    CPU = 1
    STORAGE = 2
