"""Circuits of uneven qubit counts as one launch sequence.

No reference counterpart: the reference simulates one circuit per run and has
no batching (``transpile.py:233-234`` is where it would go).  For workloads of
many circuits, each too small to fill the GPU alone (ensembles, sweeps,
training batches of different circuits), ``circuit_batch`` compiles every
circuit (by default fused into blocks of at most 5 qubits) and the engine
(``_C.functionals.CircuitBatch``, ``sq_cbatch_*``) runs them together: launch
s applies block s of every circuit that has one, so the batch takes as many
launches as its longest circuit has blocks, captured in one hipGraph.

    batch = circuit_batch([(14, ops_a), (20, ops_b), ...], dtype="complex64")
    batch.run()
    psi_b = batch.state(1)          # numpy copy of circuit 1's state
"""
from __future__ import annotations

import numpy as np

from . import _C
from .transpile import fuse_operations


class CircuitBatchHandle:
    """The engine's batch plus the qubit counts (host side)."""

    def __init__(self, engine, num_qubits, dtype):
        self.engine = engine
        self.num_qubits = list(num_qubits)
        self.dtype = dtype

    def run(self, use_graph=True, stream=None):
        """Every state <- |0...0>, then every circuit's blocks (asynchronous)."""
        self.engine.run(use_graph=use_graph, stream=stream)

    def copy_state(self, j, sv, stream=None):
        """Circuit j's state into a device StateVector of its size and dtype."""
        self.engine.copy_state(j, sv, stream=stream)

    def state(self, j):
        """Host (numpy) copy of circuit j's state (synchronises)."""
        sv = _C.result_types.StateVector(_C.DeviceType.CUDA, self.num_qubits[j], dtype=self.dtype)
        self.engine.copy_state(j, sv)
        return np.array(sv.cpu(), copy=True)

    def stats(self):
        return self.engine.stats()


def circuit_batch(circuits, dtype="complex64", fuse=True, max_fused_qubits=5):
    """circuits: [(num_qubits, operations), ...] with gate objects as for
    ``Program``; each circuit is fused into blocks of at most
    ``max_fused_qubits`` (<= 5) qubits unless ``fuse`` is false (then every
    gate must act on at most 5 qubits)."""
    if not 1 <= max_fused_qubits <= 5:
        raise ValueError("max_fused_qubits must be in 1..5")
    compiled = []
    for n, ops in circuits:
        ops = list(ops)
        if fuse and ops:
            ops = fuse_operations(ops, max_fused_qubits)
        compiled.append((int(n), ops))
    engine = _C.functionals.CircuitBatch(compiled, dtype=dtype)
    return CircuitBatchHandle(engine, [n for n, _ in compiled], dtype)
