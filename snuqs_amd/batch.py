"""Parameter batches: one circuit structure, many parameter sets.

The reference runs one circuit per call and has no batching (its
optimize_operations stub, src/braket/snuqs/transpile.py:233-234, is where
one would go).  Here K parameter sets of a parametrised circuit run together:
the K states sit in one device buffer (``StateBatch``), every gate is one
launch over all K states with each state's own matrix (``BatchProgram``,
sq_batch_*), and the energies of all K states come from one batched Pauli
reduction (``expectation_batch``).  Fusion groups gates by their targets
only, so every set fuses into the same blocks (``fuse_operations_batch``).
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy as np

from ._C import functionals as _fn
from ._C import result_types as _rt
from .transpile import fuse_with_plan, fusion_plan

BatchProgram = _fn.BatchProgram
StateBatch = _rt.StateBatch
expectation_batch = _fn.expectation_batch


def fuse_operations_batch(parameter_sets: Sequence[Sequence], max_qubits: int = 5):
    """Fuse every set with the grouping of set 0 (the sets must share their
    gates' targets), so the fused circuits share their structure."""
    plan = fusion_plan(parameter_sets[0], max_qubits)
    return [fuse_with_plan(ops, plan) for ops in parameter_sets]


def _ops(built):
    return built[0] if isinstance(built, tuple) else built


class ParameterBatch:
    """K parameter sets of ``build(theta) -> gate list`` (or ``(gates,
    param_map)``) on ``num_qubits`` qubits, evolved from |0...0> together.

    ``thetas`` has shape (K, P).  ``fuse`` (0 to disable) fuses each set into
    blocks of at most that many qubits with one grouping for all sets.
    """

    def __init__(self, num_qubits: int, build: Callable, thetas, *, dtype: str = "complex64",
                 fuse: int = 5):
        thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
        self.num_qubits = num_qubits
        self.dtype = dtype
        self._build = build
        self._fuse = fuse
        sets = [_ops(build(t)) for t in thetas]
        self._plan = fusion_plan(sets[0], fuse) if fuse and fuse >= 2 else None
        self.program = BatchProgram(num_qubits, self._fused(sets), dtype=dtype)
        self.states = StateBatch(len(sets), num_qubits, dtype=dtype)

    def _fused(self, sets):
        return [fuse_with_plan(s, self._plan) for s in sets] if self._plan else sets

    @property
    def num_sets(self) -> int:
        return self.states.num_sets

    def set_parameters(self, thetas) -> None:
        """New parameter sets (same K) for the same circuit structure."""
        thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
        self.program.set_gates(self._fused([_ops(self._build(t)) for t in thetas]))

    def run(self, *, use_graph: bool = False) -> None:
        """|0...0> in every state, then the circuit on all of them."""
        self.states.initialize_basis_z()
        self.program.run(self.states, use_graph=use_graph)

    def energies(self, paulis, qubits, coeffs, *, use_graph: bool = False) -> np.ndarray:
        """<psi_k| sum_t c_t P_t |psi_k> for every set k (FP64 reductions)."""
        self.run(use_graph=use_graph)
        e, _ = expectation_batch(self.states, paulis, qubits, coeffs)
        return e


def parameter_shift_gradient(num_qubits: int, build: Callable, theta, paulis, qubits, coeffs, *,
                             dtype: str = "complex64", fuse: int = 5, shift: float = np.pi / 2,
                             max_sets: int = 1024) -> tuple[float, np.ndarray]:
    """Energy and gradient of <H> at ``theta`` by the two-term shift rule,
    dE/dtheta_p = (E(theta + s e_p) - E(theta - s e_p)) / (2 sin s), with the
    2P shifted circuits (plus the unshifted one) evaluated as parameter
    batches of at most ``max_sets`` states.  Exact when every parameter
    enters one gate of the form exp(-i theta/2 G) with G^2 = I (RotX/Y/Z, XX,
    YY, ZZ, ...); parameters shared by several gates need per-gate shifts."""
    theta = np.asarray(theta, dtype=np.float64)
    npar = theta.size
    sets = [theta]
    for p in range(npar):
        for sg in (1.0, -1.0):
            t = theta.copy()
            t[p] += sg * shift
            sets.append(t)
    sets = np.array(sets)
    e = np.empty(len(sets))
    batch = None
    for s0 in range(0, len(sets), max_sets):
        chunk = sets[s0:s0 + max_sets]
        if batch is None or batch.num_sets != len(chunk):
            batch = ParameterBatch(num_qubits, build, chunk, dtype=dtype, fuse=fuse)
        else:
            batch.set_parameters(chunk)
        e[s0:s0 + len(chunk)] = batch.energies(paulis, qubits, coeffs)
    grad = (e[1::2] - e[2::2]) / (2.0 * np.sin(shift))
    return float(e[0]), grad
