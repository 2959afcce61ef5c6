"""Circuit batches: argument checks that run before any device allocation
(CPU only)."""
import numpy as np
import pytest


def _g():
    from snuqs_amd import _C

    return _C.operation.gate_operations


@pytest.mark.parametrize("circuits", [
    [],                                    # no circuits
    [(0, [])],                             # qubit count outside 1..30
    [(31, [])],
    [(3, [("CX", [0, 3])])],               # target out of range
    [(3, [("CX", [1, 1])])],               # repeated target
    [(2, [("Unitary3", [0, 1, 2])])],      # more targets than qubits
    [(8, [("Unitary6", [0, 1, 2, 3, 4, 5])])],  # more than 5 targets (unfused)
])
def test_circuit_batch_rejects_bad_circuits(circuits):
    from snuqs_amd.cbatch import circuit_batch

    g = _g()

    def make(name, t):
        if name.startswith("Unitary"):
            return g.Unitary(t, np.eye(1 << len(t)))
        return getattr(g, name)(t)

    built = [(n, [make(name, t) for name, t in ops]) for n, ops in circuits]
    with pytest.raises((ValueError, RuntimeError)):
        circuit_batch(built, fuse=False)


def test_circuit_batch_fusion_argument():
    from snuqs_amd.cbatch import circuit_batch

    with pytest.raises(ValueError):
        circuit_batch([(3, [])], max_fused_qubits=6)
