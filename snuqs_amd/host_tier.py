"""States larger than HBM: the state lives in pinned host memory and every
pass streams it through device buffers (the single-GPU counterpart of the
reference's hybrid and CPU-offload tiers, src/braket/snuqs/simulation.py:
286-351, whose Swap-based localisation, transpile.py:209-298, is replaced
here by choosing which host bits each pass gathers into a slice; no data is
permuted).

``HostStateSimulation(n, local_qubits)`` keeps an n-qubit state in pinned host
memory and evolves it pass by pass: ``plan_host_passes`` groups the gates,
each pass's gates are rewritten onto the device bits of its slice, fused into
<= 5-qubit blocks and run as one ``Program`` on every slice
(``Program.run_host_pass``, sq_host_tier_pass), with copies in, compute and
copies out overlapped on three streams over a ring of device buffers.
"""
from __future__ import annotations

import numpy as np

from ._C import DeviceType
from ._C import functionals as _fn
from ._C.operation import gate_operations as g
from ._C.result_types import StateVector
from .transpile import fuse_operations, plan_host_passes


def _to_device(op, local_index: dict, n: int, m: int):
    """The gate on device qubits: host qubit q is host bit n-1-q, device bit
    local_index[n-1-q], device qubit m-1-(that bit)."""
    t = list(op.targets)
    if not t:
        return op
    return g.Unitary([m - 1 - local_index[n - 1 - q] for q in t], np.array(op.matrix()))


class HostStateSimulation:
    """An n-qubit state in pinned host memory, evolved through device slices
    of ``local_qubits`` qubits (``buffers`` device buffers of 2^local_qubits
    amplitudes).  ``contiguous_bits`` (default local_qubits - 8) is the size
    of the contiguous chunks a slice is gathered from: larger chunks copy
    faster, more free bits mean fewer passes."""

    def __init__(self, qubit_count: int, local_qubits: int, *, dtype: str = "complex64",
                 contiguous_bits: int | None = None, buffers: int = 3, fuse: int = 5,
                 threads: int = 16):
        if not 1 <= local_qubits <= qubit_count:
            raise ValueError("local_qubits must be in [1, qubit_count]")
        self.qubit_count = qubit_count
        self.local_qubits = local_qubits
        self.dtype = dtype
        k = max(1, local_qubits - 8) if contiguous_bits is None else contiguous_bits
        self.contiguous_bits = min(k, local_qubits)
        self._fuse = fuse
        self._state = StateVector(qubit_count, True, dtype=dtype)  # pinned host memory
        _fn.host_initialize_basis_z(self._state, threads=threads)
        self._zero = True  # every slice but the first is zero
        self._buffers = [StateVector(DeviceType.CUDA, local_qubits, dtype=dtype)
                         for _ in range(max(1, buffers))]
        self.passes_run = 0

    def plan(self, operations):
        return plan_host_passes(operations, self.qubit_count, self.local_qubits,
                                self.contiguous_bits)

    def evolve(self, operations) -> None:
        ops = list(operations)
        n, m = self.qubit_count, self.local_qubits
        for local_bits, idx in self.plan(ops):
            where = {b: j for j, b in enumerate(local_bits)}
            dev_ops = [_to_device(ops[i], where, n, m) for i in idx]
            if self._fuse and self._fuse >= 2:
                dev_ops = fuse_operations(dev_ops, self._fuse)
            prog = _fn.Program(m, dev_ops, dtype=self.dtype)
            prog.run_host_pass(self._state, local_bits, self._buffers, zero_slices=self._zero)
            self._zero = False
            self.passes_run += 1

    @property
    def host_state(self) -> StateVector:
        return self._state

    @property
    def state_vector(self) -> np.ndarray:
        """The host state as a numpy array (a view, no copy)."""
        return np.asarray(self._state)

    @property
    def probabilities(self) -> np.ndarray:
        return np.abs(self.state_vector.astype(np.complex128)) ** 2
