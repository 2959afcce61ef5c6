"""Subcircuit: the transpiler's output (src/braket/snuqs/subcircuit.py:8-17)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

from .types import AcceleratorType, OffloadType, PrefetchType


@dataclass
class Subcircuit:
    qubit_count: int
    qubit_count_cpu: Optional[int] = None
    qubit_count_cuda: Optional[int] = None
    qubit_count_slice: Optional[int] = None
    accelerator: AcceleratorType = AcceleratorType.CUDA
    prefetch: PrefetchType = PrefetchType.NONE
    offload: OffloadType = OffloadType.NONE
    operations: list = field(default_factory=list)
