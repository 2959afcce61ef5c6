"""Parameter batches (sq_batch_*): K parameter sets of one circuit structure,
one launch per gate for all K states.  Checked bitwise against K separate
programs (the batched kernels do the same arithmetic), against the CPU oracle,
and through the parameter-shift gradient against the adjoint sweep."""
import numpy as np
import pytest

from helpers import run_oracle
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _structure(n, ngates, rng, kinds=None):
    """A random gate structure: (kind, targets) or ('unitary', targets)."""
    from snuqs_amd.circuits import GATE_KINDS

    kinds = kinds if kinds is not None else GATE_KINDS
    out = []
    for _ in range(ngates):
        i = int(rng.integers(0, len(kinds) + 1))
        if i == len(kinds):
            k = int(rng.integers(1, 4))
            out.append(("unitary", [int(x) for x in rng.permutation(n)[:k]]))
        else:
            out.append((kinds[i], [int(x) for x in rng.permutation(n)[:kinds[i][1]]]))
    return out


def _instantiate(struct, rng, zero_angle=False):
    from snuqs_amd.circuits import haar_unitary, make_gate
    from snuqs_amd._C.operation import gate_operations as g

    ops = []
    for kind, t in struct:
        if kind == "unitary":
            ops.append(g.Unitary(t, haar_unitary(1 << len(t), rng)))
        else:
            a = np.zeros(3) if zero_angle else rng.random(3) * 2 * np.pi
            ops.append(make_gate(kind, t, a))
    return ops


def _separate(C, n, sets, dtype):
    out = []
    for ops in sets:
        sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
        C.functionals.initialize_basis_z(sv)
        C.functionals.Program(n, ops, dtype=dtype).run(sv)
        out.append(np.array(sv.cpu(), copy=True))
    return np.array(out)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,nsets", [(10, 5), (14, 8), (17, 3)])
def test_batch_equals_separate_programs(C, dtype, n, nsets):
    """Every gate class (phase, diagonal, permutation, dense 1-3 on high and
    low bits, 4-5 target MFMA blocks from fusion, global phase) with its own
    angles per set: bitwise equal to running each set as its own program."""
    from snuqs_amd.batch import fuse_operations_batch
    from snuqs_amd.circuits import GATE_KINDS
    from snuqs_amd._C.operation import gate_operations as g

    rng = np.random.default_rng(100 * n + nsets)
    struct = _structure(n, 60, rng)
    sets = [[g.Hadamard([q]) for q in range(n)] + _instantiate(struct, rng)
            for _ in range(nsets)]
    for fuse in (False, True):
        ss = fuse_operations_batch(sets, 5) if fuse else sets
        assert all(len(s) == len(ss[0]) for s in ss)
        prog = C.functionals.BatchProgram(n, ss, dtype=dtype)
        b = C.result_types.StateBatch(nsets, n, dtype=dtype)
        b.initialize_basis_z()
        prog.run(b)
        got = b.cpu()
        want = _separate(C, n, ss, dtype)
        assert got.shape == want.shape
        assert np.array_equal(got, want), (fuse, np.abs(got - want).max())
        if fuse:
            assert any(len(op.targets) == 5 for op in ss[0])  # MFMA blocks ran
    # and the oracle, set by set
    tol = 1e-12 if dtype == "complex128" else 1e-5
    for y in range(nsets):
        ref = run_oracle(n, sets[y])
        assert np.abs(got[y].astype(np.complex128) - ref).max() <= tol, y


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_batch_mixed_classes_fall_back_to_dense(C, dtype):
    """A gate whose sets fall into different classes (angle 0 makes RotY the
    identity, CPhaseShift(0) a no-op, PSwap(0) a plain swap) runs as a dense
    gate in every set: checked against the oracle."""
    from snuqs_amd._C.operation import gate_operations as g

    n, rng = 12, np.random.default_rng(5)
    kinds = [(g.RotY, 1, 1), (g.CPhaseShift, 2, 1), (g.PSwap, 2, 1), (g.RotZ, 1, 1),
             (g.ZZ, 2, 1), (g.U, 1, 3), (g.XX, 2, 1), (g.GPhase, 0, 1)]
    struct = _structure(n, 40, rng, kinds)
    sets = [[g.Hadamard([q]) for q in range(n)] + _instantiate(struct, rng, zero_angle=(y == 1))
            for y in range(4)]
    prog = C.functionals.BatchProgram(n, sets, dtype=dtype)
    b = C.result_types.StateBatch(4, n, dtype=dtype)
    b.initialize_basis_z()
    prog.run(b)
    got = b.cpu()
    tol = 1e-12 if dtype == "complex128" else 1e-5
    for y in range(4):
        assert np.abs(got[y].astype(np.complex128) - run_oracle(n, sets[y])).max() <= tol, y


def test_batch_graph_set_gates_and_views(C):
    """hipGraph replays equal eager runs; set_gates equals a new batch; a
    state view shares the batch's memory."""
    from snuqs_amd.batch import fuse_operations_batch
    from snuqs_amd.circuits import hea_ansatz

    n, K = 16, 6
    rng = np.random.default_rng(9)
    th = rng.uniform(0, 2 * np.pi, (K, 2 * n * 3))
    sets = fuse_operations_batch([hea_ansatz(n, 3, t)[0] for t in th], 5)
    prog = C.functionals.BatchProgram(n, sets, dtype="complex64")
    b = C.result_types.StateBatch(K, n, dtype="complex64")
    b.initialize_basis_z()
    prog.run(b)
    eager = b.cpu()
    for _ in range(3):
        b.initialize_basis_z()
        prog.run(b, use_graph=True)
        C.functionals.norm2(b.state(0))  # eager work between replays
        assert np.array_equal(b.cpu(), eager)
    th2 = rng.uniform(0, 2 * np.pi, (K, 2 * n * 3))
    sets2 = fuse_operations_batch([hea_ansatz(n, 3, t)[0] for t in th2], 5)
    prog.set_gates(sets2)
    b.initialize_basis_z()
    prog.run(b, use_graph=True)
    fresh = C.functionals.BatchProgram(n, sets2, dtype="complex64")
    b2 = C.result_types.StateBatch(K, n, dtype="complex64")
    b2.initialize_basis_z()
    fresh.run(b2)
    assert np.array_equal(b.cpu(), b2.cpu())
    v = b.state(K - 1)
    assert np.array_equal(np.array(v.cpu()), b2.cpu()[K - 1])
    assert prog.stats()["launches"] == len(sets2[0])


def test_batch_expectation_equals_per_state(C):
    from snuqs_amd.circuits import random_pauli_sum

    n, K = 13, 7
    rng = np.random.default_rng(4)
    b = C.result_types.StateBatch(K, n, dtype="complex128")
    states = []
    for y in range(K):
        v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        v /= np.linalg.norm(v)
        states.append(v)
        h = C.result_types.StateVector(n, dtype="complex128")
        np.array(h, copy=False)[:] = v
        b.state(y).copy(h)
    obs = random_pauli_sum(n, 90, 3, rng)  # two term launches
    e, terms = C.functionals.expectation_batch(b, *obs)
    assert e.shape == (K,) and terms.shape == (K, 90)
    for y in range(K):
        e1, t1 = C.functionals.expectation(b.state(y), *obs)
        assert np.array_equal(terms[y], t1) and e[y] == e1
        ref = sum(c * O.expect_pauli(states[y], p, q, n).real for p, q, c in zip(*obs))
        assert abs(e[y] - ref) <= 1e-12


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-9), ("complex64", 1e-5)])
def test_parameter_shift_batch_vs_adjoint(C, dtype, tol):
    """The 2P + 1 shifted circuits of a 12-qubit ansatz as one batch: the
    energy and shift-rule gradient against the adjoint sweep (complex128)."""
    from snuqs_amd.batch import parameter_shift_gradient
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    n, layers = 12, 3
    rng = np.random.default_rng(12)
    th = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    obs = random_pauli_sum(n, 16, 3, rng)
    e, grad = parameter_shift_gradient(n, lambda t: hea_ansatz(n, layers, t)[0], th, *obs,
                                       dtype=dtype, max_sets=40)
    ops, pmap = hea_ansatz(n, layers, th)
    prog = C.functionals.Program(n, ops, dtype="complex128")
    st = [C.result_types.StateVector(C.DeviceType.CUDA, n, dtype="complex128") for _ in range(3)]
    e_adj, g_adj = prog.adjoint(st[0], st[1], st[2], [], pmap, th.size, *obs)
    assert abs(e - e_adj) <= tol
    assert np.abs(grad - g_adj).max() <= tol


def test_batch_argument_errors(C):
    from snuqs_amd._C.operation import gate_operations as g

    with pytest.raises(ValueError):
        C.functionals.BatchProgram(4, [[g.RotY([0], 0.1)], [g.RotY([1], 0.1)]])
    with pytest.raises(ValueError):
        C.functionals.BatchProgram(4, [[g.RotY([0], 0.1)], []])
    prog = C.functionals.BatchProgram(4, [[g.RotY([0], 0.1)], [g.RotY([0], 0.2)]])
    with pytest.raises(ValueError):
        prog.run(C.result_types.StateBatch(3, 4))
    with pytest.raises(ValueError):
        prog.run(C.result_types.StateBatch(2, 5))
    with pytest.raises(ValueError):
        prog.run(C.result_types.StateBatch(2, 4, dtype="complex64"))


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batch_tiles_per_wave_bitwise(C, dtype):
    """32 sets of 20 qubits: enough tiles for the 5-target blocks to walk 4
    tiles per wave (mfma_tiles_per_wave 4) as well as 1; both grids give the same
    bits, equal to separate programs."""
    from snuqs_amd.batch import fuse_operations_batch
    from snuqs_amd.circuits import hea_ansatz

    n, layers, nsets = 20, 2, 32
    rng = np.random.default_rng(2020)
    thetas = rng.uniform(0, 2 * np.pi, (nsets, 2 * n * layers))
    ss = fuse_operations_batch([hea_ansatz(n, layers, t)[0] for t in thetas], 5)
    assert any(len(op.targets) == 5 for op in ss[0])
    saved = C.core.tune_get("mfma_tiles_per_wave")
    got = {}
    try:
        for tiles in (1, 4):
            C.core.tune_set("mfma_tiles_per_wave", tiles)
            prog = C.functionals.BatchProgram(n, ss, dtype=dtype)
            b = C.result_types.StateBatch(nsets, n, dtype=dtype)
            b.initialize_basis_z()
            prog.run(b)
            got[tiles] = b.cpu()
            del prog, b
    finally:
        C.core.tune_set("mfma_tiles_per_wave", saved)
    assert np.array_equal(got[1], got[4])
    want = _separate(C, n, [ss[0], ss[17], ss[31]], dtype)
    assert np.array_equal(got[4][[0, 17, 31]], want)
