"""Scheduler (CPU): pseudo_sort reproduces the reference order; fusion
preserves the circuit's unitary action (checked with the CPU oracle)."""
import numpy as np
import pytest

from helpers import run_oracle
from oracle import oracle as O


def test_pseudo_sort_matches_reference_order():
    from snuqs_amd.circuits import random_gate_list
    from snuqs_amd.transpile import pseudo_sort_operations_descending

    rng = np.random.default_rng(0)
    for n in (3, 7, 12):
        ops = random_gate_list(n, 150, rng, with_unitary=True)
        got = pseudo_sort_operations_descending(ops)
        want = O.pseudo_sort_operations_descending(ops)
        assert all(a is b for a, b in zip(got, want))
        # reordering changes evaluation order only, not the state
        assert np.abs(run_oracle(n, got) - run_oracle(n, ops)).max() < 1e-13


@pytest.mark.parametrize("maxq", [2, 3, 4, 5])
def test_fusion_preserves_state(maxq):
    from snuqs_amd.circuits import random_gate_list
    from snuqs_amd.transpile import fuse_operations

    rng = np.random.default_rng(maxq)
    n = 9
    ops = random_gate_list(n, 250, rng, with_unitary=True)
    fused = fuse_operations(ops, maxq)
    assert len(fused) < len(ops)
    assert all(len(op.targets) <= max(maxq, 3) for op in fused)
    assert np.abs(run_oracle(n, fused) - run_oracle(n, ops)).max() < 1e-12


def test_fusion_of_layered_circuit_uses_five_qubit_blocks():
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    n = 10
    ops = random_layered(n, 6, seed=3)
    fused = fuse_operations(ops, 5)
    assert max(len(op.targets) for op in fused) == 5
    assert np.abs(run_oracle(n, fused) - run_oracle(n, ops)).max() < 1e-12


def test_pseudo_sort_hand_worked_orders_product():
    """The product's reorder on the hand-worked cases that pin the oracle."""
    from test_oracle import _PSEUDO_SORT_CASES, _Op
    from snuqs_amd.transpile import pseudo_sort_operations_descending

    for name, ops, want in _PSEUDO_SORT_CASES:
        got = pseudo_sort_operations_descending([_Op(nm, t) for nm, t in ops])
        assert [op.name for op in got] == want, name


def test_fusion_plan_is_shared_by_parameter_sets():
    """Sets that differ only in angles fuse into the same blocks (same
    targets, gate by gate), each with its own matrices: equal to fusing the
    set alone."""
    from snuqs_amd.batch import fuse_operations_batch
    from snuqs_amd.circuits import hea_ansatz
    from snuqs_amd.transpile import fuse_operations

    n, layers = 9, 3
    rng = np.random.default_rng(1)
    sets = [hea_ansatz(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers))[0] for _ in range(4)]
    fused = fuse_operations_batch(sets, 5)
    for s, f in zip(sets, fused):
        alone = fuse_operations(s, 5)
        assert [list(op.targets) for op in f] == [list(op.targets) for op in fused[0]]
        assert [list(op.targets) for op in f] == [list(op.targets) for op in alone]
        for a, b in zip(f, alone):
            assert np.array_equal(np.array(a.matrix()), np.array(b.matrix()))
