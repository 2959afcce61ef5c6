"""Circuits of uneven qubit counts as one launch sequence (sq_cbatch_*):
every circuit's state against the FP64 CPU oracle and against the same
circuit run alone as a Program; graph replays against eager runs; the
largest offsets (a 26-qubit circuit beside small ones); copy_state's checks."""
import numpy as np
import pytest

from helpers import device_state, run_oracle, to_np

pytestmark = pytest.mark.gpu


def _mixed_circuit(C, n, rng, depth):
    """Gates of 1-3 targets of every kind (dense, diagonal, permutation,
    controlled) plus 4-5 target blocks once fused."""
    from snuqs_amd.circuits import random_layered

    g = C.operation.gate_operations
    ops = [g.Hadamard([q]) for q in range(n)]
    if n >= 2:
        ops += random_layered(n, depth, seed=int(rng.integers(1 << 30)))
        ops += [g.CX([0, n - 1]), g.RotZ([n - 1], 0.3), g.Swap([0, 1])]
    if n >= 3:
        ops += [g.RotX([2, 0, 1], 0.7, ctrl_modifiers=[1, 0]), g.CCNot([1, 2, 0])]
    ops += [g.RotY([q], float(rng.uniform(0, 6.28))) for q in range(n)]
    return ops


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 2e-5)])
@pytest.mark.parametrize("fuse", [False, True])
def test_cbatch_vs_oracle_and_programs(C, dtype, tol, fuse):
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.transpile import fuse_operations

    rng = np.random.default_rng(5)
    sizes = [3, 10, 1, 12, 7, 9, 2, 11]
    circs = [(n, _mixed_circuit(C, n, rng, 3)) for n in sizes]
    batch = circuit_batch(circs, dtype=dtype, fuse=fuse)
    batch.run(use_graph=False)
    st = batch.stats()
    nblocks = [len(fuse_operations(ops, 5)) if fuse else len(ops) for _, ops in circs]
    assert max(nblocks) + 2 <= st["launches"] <= 2 * max(nblocks) + 2
    for j, (n, ops) in enumerate(circs):
        got = batch.state(j).astype(np.complex128)
        ref = run_oracle(n, ops)
        assert np.abs(got - ref).max() < tol, (j, n)
        # the same blocks run alone through a Program
        prog_ops = fuse_operations(ops, 5) if fuse else ops
        sv = device_state(C, n, dtype)
        C.functionals.Program(n, prog_ops, dtype=dtype).run(sv)
        assert np.abs(got - to_np(C, sv)).max() < tol, (j, n)


def test_cbatch_graph_replays_match_eager(C):
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.circuits import random_layered

    circs = [(n, random_layered(n, 4, seed=n)) for n in (14, 6, 18, 13, 16)]
    batch = circuit_batch(circs, dtype="complex64")
    batch.run(use_graph=False)
    eager = [batch.state(j) for j in range(len(circs))]
    for r in range(3):
        batch.run(use_graph=True)  # captured on the first call, replayed after
        C.core.cuda.device_synchronize()
        for j in range(len(circs)):
            assert np.array_equal(batch.state(j), eager[j]), (r, j)


def test_cbatch_uneven_gate_counts_and_large_offsets(C):
    """A 26-qubit circuit (its state past 2^28 bytes of offsets for the ones
    after it) beside circuits of 0, 1 and many blocks; complex128 against
    Programs run alone (the oracle for the small ones)."""
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    g = C.operation.gate_operations
    circs = [(5, []), (4, [g.Hadamard([2])]), (26, random_layered(26, 3, seed=2)),
             (8, random_layered(8, 12, seed=3)), (20, random_layered(20, 2, seed=4))]
    batch = circuit_batch(circs, dtype="complex128")
    batch.run()
    for j, (n, ops) in enumerate(circs):
        got = batch.state(j)
        if n <= 12:
            ref = run_oracle(n, ops)
        else:
            sv = device_state(C, n, "complex128")
            C.functionals.Program(n, fuse_operations(ops, 5), dtype="complex128").run(sv)
            ref = to_np(C, sv)
        assert np.abs(got - ref).max() < 1e-12, (j, n)
        assert abs(np.vdot(got, got).real - 1.0) < 1e-12, (j, n)


def test_cbatch_copy_state_checks(C):
    from snuqs_amd.cbatch import circuit_batch

    g = C.operation.gate_operations
    batch = circuit_batch([(3, [g.Hadamard([0])]), (5, [g.CX([0, 4])])], dtype="complex64")
    batch.run()
    with pytest.raises(ValueError):  # wrong size
        batch.copy_state(0, C.result_types.StateVector(C.DeviceType.CUDA, 5, dtype="complex64"))
    with pytest.raises(ValueError):  # wrong dtype
        batch.copy_state(1, C.result_types.StateVector(C.DeviceType.CUDA, 5, dtype="complex128"))
    with pytest.raises(IndexError):
        batch.copy_state(2, C.result_types.StateVector(C.DeviceType.CUDA, 5, dtype="complex64"))
    sv = C.result_types.StateVector(C.DeviceType.CUDA, 5, dtype="complex64")
    batch.copy_state(1, sv)
    a = np.array(sv.cpu())
    assert a[0] == 1 and np.count_nonzero(a) == 1


@pytest.mark.parametrize("dtype,tol", [("complex64", 2e-5), ("complex128", 1e-12)])
def test_cbatch_many_circuits(C, dtype, tol):
    """300 circuits (more entries per launch than a workgroup stages in LDS,
    so entries are found by the global search), 2-11 qubits, fused (5-target
    blocks on states with full tiles take the MFMA entry table) against the
    oracle; a second run of the batch gives the same bits."""
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.circuits import random_layered

    rng = np.random.default_rng(300)
    sizes = [int(rng.integers(2, 12)) for _ in range(300)]
    circs = [(n, random_layered(n, 3, seed=5000 + j)) for j, n in enumerate(sizes)]
    batch = circuit_batch(circs, dtype=dtype)
    batch.run()
    picks = [int(j) for j in rng.choice(len(circs), 40, replace=False)]
    first = {j: batch.state(j) for j in picks}
    for j in picks:
        n, ops = circs[j]
        got = first[j].astype(np.complex128)
        assert np.abs(got - run_oracle(n, ops)).max() < tol, (j, n)
    batch.run()
    for j in picks:
        assert np.array_equal(batch.state(j), first[j]), j


def test_cbatch_tiles_per_wave_bitwise(C):
    """Steps whose 5-target entries hold enough tiles for 4 tiles per wave
    (40 circuits of 20 qubits) give the same bits as 1 tile per wave, and
    agree with Programs run alone."""
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    circs = [(20, random_layered(20, 3, seed=700 + j)) for j in range(40)]
    saved = C.core.tune_get("mfma_tiles_per_wave")
    states = {}
    try:
        for tiles in (1, 4):
            C.core.tune_set("mfma_tiles_per_wave", tiles)
            batch = circuit_batch(circs, dtype="complex64")
            batch.run()
            states[tiles] = [batch.state(j) for j in (0, 13, 39)]
            del batch
    finally:
        C.core.tune_set("mfma_tiles_per_wave", saved)
    for j, a, b in zip((0, 13, 39), states[1], states[4]):
        assert np.array_equal(a, b), j
        sv = device_state(C, 20, "complex64")
        C.functionals.Program(20, fuse_operations(circs[j][1], 5), dtype="complex64").run(sv)
        assert np.abs(b.astype(np.complex128) - to_np(C, sv)).max() < 2e-5, j
