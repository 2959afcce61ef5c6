"""State-vector simulation driven by the HIP engine.

Mirrors src/braket/snuqs/simulation.py of the reference (Simulation ABC,
StateVectorSimulation with state_vector / probabilities / retrieve_samples /
evolve) for the single-GPU tier.  Differences, by design:

* the state lives in HBM (``DeviceType.CUDA`` = the HIP device) in complex64
  or complex128; the reference's CPU / hybrid / offload tiers
  (simulation.py:113-194, 286-542) collapse into one GPU's memory;
* ``evolve`` uploads the whole gate list once (``_C.functionals.Program``) and
  runs it in one call instead of one pybind call per gate
  (simulation.py:274-284), after optional fusion into <= 5-qubit blocks;
* ``retrieve_samples`` draws the uniforms from numpy's global RandomState as
  np.random.choice does (simulation.py:220) and searches the cumulative
  distribution on the GPU: with ``exact_sampling`` the indices are identical
  to np.random.choice(len(sv), p=probabilities, size=shots);
* ``probabilities`` is computed on the GPU in float64 (for complex64 divided
  by the FP64 norm^2, so it sums to 1 within FP64 rounding).
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np

from ._C import DeviceType
from ._C.functionals import Program, initialize_basis_z
from ._C.result_types import StateVector
from . import _C
from .subcircuit import Subcircuit
from .transpile import fuse_operations
from .types import AcceleratorType, OffloadType


class Simulation(ABC):
    """Tracks the evolution of a quantum system with ``qubit_count`` qubits."""

    def __init__(self, qubit_count: int, shots: int):
        self._qubit_count = qubit_count
        self._shots = shots

    @property
    def qubit_count(self) -> int:
        return self._qubit_count

    @property
    def shots(self) -> int:
        return self._shots

    @abstractmethod
    def marginal_probabilities(self, targets) -> np.ndarray:
        """Distribution over `targets` (targets[0] the most significant bit of
        the outcome index)."""

    @abstractmethod
    def retrieve_samples(self) -> np.ndarray:
        """Sampled basis-state indices (one per shot)."""

    @property
    @abstractmethod
    def probabilities(self) -> np.ndarray:
        """Probability of each computational basis state."""

    @abstractmethod
    def evolve(self, subcircuit: Subcircuit) -> None:
        """Apply the subcircuit's gates to the state."""


class StateVectorSimulation(Simulation):
    def __init__(self, qubit_count: int, shots: int, subcircuit: Subcircuit,
                 *, precision: str = "complex128", fuse: int = 5,
                 exact_sampling: bool = True):
        super().__init__(qubit_count=qubit_count, shots=shots)
        if subcircuit.offload not in (OffloadType.NONE, None):
            raise TypeError("offload tiers are not part of the single-GPU engine; "
                            "use offload=None")
        if subcircuit.accelerator not in (AcceleratorType.CUDA, AcceleratorType.HIP):
            raise TypeError(f"accelerator {subcircuit.accelerator} is not supported: the "
                            "engine runs on the GPU (accelerator='cuda' or 'hip')")
        self._precision = precision
        self._fuse = fuse
        self._exact = exact_sampling
        self._state_vector = StateVector(DeviceType.CUDA, qubit_count, dtype=precision)
        initialize_basis_z(self._state_vector)

    @property
    def precision(self) -> str:
        return self._precision

    @property
    def device_state(self) -> StateVector:
        """The device-resident state (for the engine's functionals)."""
        return self._state_vector

    def evolve(self, subcircuit: Subcircuit) -> None:
        ops = list(subcircuit.operations)
        if not ops:
            return
        if self._fuse and self._fuse >= 2:
            ops = fuse_operations(ops, self._fuse)
        Program(self._qubit_count, ops, dtype=self._precision).run(self._state_vector)

    @property
    def state_vector(self) -> np.ndarray:
        """Host copy of the state (complex128 for a complex128 simulation)."""
        return np.array(self._state_vector.cpu(), copy=True)

    @property
    def _normalize(self) -> bool:
        # FP32 rounding drifts a complex64 state's norm by ~1e-7 over deep
        # circuits, beyond the sqrt(eps) = 1.5e-8 that numpy's choice accepts
        # in sum(p); the FP32 state cannot be rescaled that finely (the
        # correction is below half an ulp), so its probabilities are divided
        # by the FP64 norm instead
        return self._precision == "complex64"

    @property
    def probabilities(self) -> np.ndarray:
        return _C.functionals.probabilities(self._state_vector, normalize=self._normalize)

    def marginal_probabilities(self, targets) -> np.ndarray:
        """Distribution over `targets` (targets[0] the most significant bit of
        the outcome index), reduced on the GPU in FP64 (1..20 targets); a
        permutation of all qubits reorders the full distribution."""
        targets = [int(t) for t in targets]
        n = self._qubit_count
        if sorted(targets) == list(range(n)):
            p = self.probabilities
            if targets == list(range(n)):
                return p
            # outcome r reads qubit targets[j] at bit n-1-j: index of r in p
            r = np.arange(p.size, dtype=np.int64)
            src = np.zeros_like(r)
            for j, q in enumerate(targets):
                src |= ((r >> (n - 1 - j)) & 1) << (n - 1 - q)
            return p[src]
        m = _C.functionals.marginal_probabilities(self._state_vector, targets)
        return m / m.sum() if self._normalize else m

    def retrieve_samples(self) -> np.ndarray:
        if self._shots == 0:
            return np.array([], dtype=np.int64)
        u = np.random.random_sample(self._shots)
        return _C.functionals.sample(self._state_vector, u, exact=self._exact,
                                     normalize=self._normalize)

    def amplitudes(self, indices) -> np.ndarray:
        return _C.functionals.amplitudes(self._state_vector, [int(i) for i in indices])

    def diagonalized_moments(self, rotations, targets, eigenvalues) -> tuple:
        """<O> and <O^2> of an observable O = D^dagger diag(eigenvalues) D:
        the diagonalizing gates D act on a copy of the state, whose marginal
        distribution over `targets` (targets[0] the most significant outcome
        bit, at most 20) weights the eigenvalues (observables.py:245-305:
        eigendecomposition and diagonalizing gate of a Hermitian)."""
        work = _C.result_types.StateVector(_C.DeviceType.CUDA, self._qubit_count,
                                           dtype=self._precision)
        work.copy(self._state_vector)
        for op in rotations:
            _C.functionals.apply(work, op, self._qubit_count, list(op.targets))
        p = np.asarray(_C.functionals.marginal_probabilities(work, list(targets)), np.float64)
        lam = np.asarray(eigenvalues, dtype=np.float64)
        return float(p @ lam), float(p @ (lam * lam))

    def expectation(self, paulis, qubits, coeffs) -> float:
        return _C.functionals.expectation(self._state_vector, list(paulis),
                                          [list(q) for q in qubits], list(coeffs))[0]


class HostStateVectorSimulation(Simulation):
    """The state in pinned host memory, evolved through device slices (the
    reference's hybrid accelerator and CPU-offload tiers,
    simulation.py:286-351): for states larger than HBM.  ``local_qubits``
    (the reference's qubit_count_cuda) is the device slice size;
    ``contiguous_bits`` (the reference's qubit_count_slice plays this part)
    the size of the chunks slices are gathered from.  Results come from the
    host state: state_vector, amplitudes, probabilities (FP64, divided by the
    FP64 norm^2 for complex64) and samples drawn by numpy's choice exactly as
    simulation.py:217-220 does; expectations of Pauli and Hermitian
    observables need the state on the device and are not offered."""

    def __init__(self, qubit_count: int, shots: int, subcircuit: Subcircuit, *,
                 precision: str = "complex128", fuse: int = 5, local_qubits: int | None = None,
                 contiguous_bits: int | None = None, buffers: int = 3):
        super().__init__(qubit_count=qubit_count, shots=shots)
        from .host_tier import HostStateSimulation

        item = 8 if precision == "complex64" else 16
        if local_qubits is None:
            # three device buffers within half of the free HBM
            free = _C.core.cuda.mem_info()[0]
            local_qubits = int(np.floor(np.log2(max(free // 2 // (3 * item), 2))))
        m = max(1, min(int(local_qubits), qubit_count))
        self._precision = precision
        self._host = HostStateSimulation(qubit_count, m, dtype=precision,
                                         contiguous_bits=contiguous_bits, buffers=buffers,
                                         fuse=fuse)

    @property
    def precision(self) -> str:
        return self._precision

    def evolve(self, subcircuit: Subcircuit) -> None:
        self._host.evolve(list(subcircuit.operations))

    @property
    def state_vector(self) -> np.ndarray:
        return self._host.state_vector.astype(np.complex128 if self._precision == "complex128"
                                              else np.complex64)

    def amplitudes(self, indices) -> np.ndarray:
        a = self._host.state_vector
        return np.array([a[int(i)] for i in indices], dtype=np.complex128)

    @property
    def probabilities(self) -> np.ndarray:
        p = self._host.probabilities
        return p / p.sum() if self._precision == "complex64" else p

    def marginal_probabilities(self, targets) -> np.ndarray:
        targets = [int(t) for t in targets]
        n = self._qubit_count
        p = self.probabilities.reshape([2] * n)
        rest = tuple(q for q in range(n) if q not in targets)
        m = p.sum(axis=rest) if rest else p
        # axes left in ascending qubit order; reorder to `targets`
        order = sorted(targets)
        return np.transpose(m, [order.index(q) for q in targets]).reshape(-1)

    def retrieve_samples(self) -> np.ndarray:
        if self._shots == 0:
            return np.array([], dtype=np.int64)
        p = self.probabilities
        return np.random.choice(p.size, p=p, size=self._shots)

    def expectation(self, paulis, qubits, coeffs) -> float:
        raise NotImplementedError("expectation values need the device-resident state: run "
                                  "without offload='cpu' / accelerator='hybrid'")

    def diagonalized_moments(self, rotations, targets, eigenvalues) -> tuple:
        raise NotImplementedError("observables need the device-resident state: run without "
                                  "offload='cpu' / accelerator='hybrid'")
