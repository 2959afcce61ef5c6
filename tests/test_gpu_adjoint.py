"""Adjoint gradients vs central finite differences and parameter shift, both
computed with the FP64 CPU oracle."""
import numpy as np
import pytest

from helpers import device_state
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_energy(n, ops, paulis, qubits, coeffs):
    st = O.basis_state(n)
    for op in ops:
        O.apply(st, np.array(op.matrix()), op.targets, n)
    e = 0.0
    for p, qs, c in zip(paulis, qubits, coeffs):
        e += c * O.expect_pauli(st, p, qs, n).real
    return e


def rebuild(ops_fn, thetas):
    return ops_fn(thetas)


def _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints=0):
    prog = C.functionals.Program(n, ops, dtype=dtype)
    sv = device_state(C, n, dtype)
    w0 = device_state(C, n, dtype)
    w1 = device_state(C, n, dtype)
    cps = [device_state(C, n, dtype) for _ in range(checkpoints)]
    return prog.adjoint(sv, w0, w1, cps, pmap, nparams, *obs)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-8), ("complex64", 1e-5)])
@pytest.mark.parametrize("checkpoints", [0, 3])
def test_hea_adjoint_vs_finite_differences(C, dtype, tol, checkpoints):
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    n, layers = 8, 3
    rng = np.random.default_rng(24)
    thetas = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    obs = random_pauli_sum(n, 12, 3, rng)
    ops, pmap = hea_ansatz(n, layers, thetas)
    e, grads = _adjoint(C, n, ops, pmap, thetas.size, obs, dtype, checkpoints)
    e_ref = oracle_energy(n, ops, *obs)
    assert abs(e - e_ref) < (1e-12 if dtype == "complex128" else 1e-5)
    h = 1e-5
    fd = np.empty_like(thetas)
    for j in range(thetas.size):
        tp, tm = thetas.copy(), thetas.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (oracle_energy(n, hea_ansatz(n, layers, tp)[0], *obs)
                 - oracle_energy(n, hea_ansatz(n, layers, tm)[0], *obs)) / (2 * h)
    # parameter shift (exact for exp(-i theta/2 P) gates)
    ps = np.empty_like(thetas)
    for j in range(thetas.size):
        tp, tm = thetas.copy(), thetas.copy()
        tp[j] += np.pi / 2
        tm[j] -= np.pi / 2
        ps[j] = 0.5 * (oracle_energy(n, hea_ansatz(n, layers, tp)[0], *obs)
                       - oracle_energy(n, hea_ansatz(n, layers, tm)[0], *obs))
    assert np.abs(ps - fd).max() < 1e-7  # the two oracles agree
    assert np.abs(grads - ps).max() < tol


def test_adjoint_all_parametric_gate_kinds(C):
    """Every parametric gate (including U's three angles and shared params)."""
    g = C.operation.gate_operations
    n = 5
    rng = np.random.default_rng(6)
    th = rng.uniform(0, 2 * np.pi, 7)

    def build(t):
        ops = [g.Hadamard([q]) for q in range(n)]
        pm = [[] for _ in range(n)]
        spec = [
            (g.RotX([0], t[0]), [(0, 0)]), (g.PhaseShift([1], t[1]), [(0, 1)]),
            (g.CPhaseShift([1, 2], t[2]), [(0, 2)]), (g.CPhaseShift01([3, 2], t[2]), [(0, 2)]),
            (g.PSwap([0, 4], t[3]), [(0, 3)]), (g.XY([2, 4], t[4]), [(0, 4)]),
            (g.XX([1, 3], t[0]), [(0, 0)]), (g.YY([0, 2], t[1]), [(0, 1)]),
            (g.ZZ([3, 4], t[4]), [(0, 4)]), (g.U([2], t[5], t[6], t[3]), [(0, 5), (1, 6), (2, 3)]),
            (g.CPhaseShift10([4, 0], t[6]), [(0, 6)]), (g.CPhaseShift00([1, 4], t[5]), [(0, 5)]),
            (g.RotY([3], t[2]), [(0, 2)]), (g.RotZ([1], t[1]), [(0, 1)]),
        ]
        for op, m in spec:
            ops.append(op)
            pm.append(m)
        return ops, pm

    obs = (["XZ", "YY", "Z", "XYZ"], [[0, 1], [2, 3], [4], [0, 2, 4]], [0.4, -0.7, 1.3, 0.2])
    ops, pm = build(th)
    e, grads = _adjoint(C, n, ops, pm, th.size, obs, "complex128")
    h = 1e-6
    fd = np.empty_like(th)
    for j in range(th.size):
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (oracle_energy(n, build(tp)[0], *obs) - oracle_energy(n, build(tm)[0], *obs)) / (2 * h)
    assert np.abs(grads - fd).max() < 1e-7


@pytest.mark.parametrize("n", [16, 23])
@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-9), ("complex64", 2e-4)])
def test_hea_adjoint_vs_gpu_parameter_shift(C, dtype, tol, n):
    """16 and 23 qubits (targets on every bit position, including the ones
    above the lowest eight; at 23 qubits each workgroup of a diagonal run
    covers two 1024-amplitude chunks): adjoint gradients against parameter
    shift evaluated with the engine itself in complex128 (exact for the
    exp(-i theta/2 P) rotations)."""
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    layers = 2
    rng = np.random.default_rng(16)
    thetas = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    obs = random_pauli_sum(n, 10, 4, rng)
    ops, pmap = hea_ansatz(n, layers, thetas)
    e, grads = _adjoint(C, n, ops, pmap, thetas.size, obs, dtype, checkpoints=2)
    sv = device_state(C, n, "complex128")

    def energy(th):
        C.functionals.initialize_basis_z(sv)
        C.functionals.Program(n, hea_ansatz(n, layers, th)[0], dtype="complex128").run(sv)
        return C.functionals.expectation(sv, *obs)[0]

    assert abs(e - energy(thetas)) < (1e-12 if dtype == "complex128" else 1e-5)
    ps = np.empty_like(thetas)
    for j in range(thetas.size):
        tp, tm = thetas.copy(), thetas.copy()
        tp[j] += np.pi / 2
        tm[j] -= np.pi / 2
        ps[j] = 0.5 * (energy(tp) - energy(tm))
    assert np.abs(grads - ps).max() < tol


def _planner_circuit(n, thetas):
    """Gates that exercise the adjoint's pass planning: a diagonal run of 80
    gates (RotZ on every qubit, ZZ / CPhaseShift / PhaseShift pairs, CZ) with
    shared parameters and more than 32 derivatives, layers of one-target
    rotations on distinct qubits (low address bits included), and dense
    two-qubit gates between them."""
    from snuqs_amd import _C as C

    g = C.operation.gate_operations
    ops, pmap = [], []

    def add(op, params=()):
        ops.append(op)
        pmap.append([(a, p) for a, p in params])

    for q in range(n):
        add(g.Hadamard([q]))
    for rep in range(2):
        for q in range(n):                                   # RotY layer, own params
            add(g.RotY([q], thetas[q]), [(0, q)])
        for q in range(n):                                   # diagonal run starts
            add(g.RotZ([q], thetas[n + q % 3]), [(0, n + q % 3)])  # shared params
        for q in range(n - 1):
            add(g.ZZ([q, q + 1], thetas[n + 3 + q % 5]), [(0, n + 3 + q % 5)])
            add(g.CZ([q, (q + 3) % n]))
            add(g.CPhaseShift([(q + 2) % n, q], thetas[n + 8]), [(0, n + 8)])
            add(g.PhaseShift([q], thetas[n + 9 + q % 2]), [(0, n + 9 + q % 2)])
        for q in range(n):                                   # the run goes past 64 gates
            add(g.S([q]))
            add(g.T([(q + 1) % n]))
        for q in range(0, n, 2):                             # shared RotX params
            add(g.RotX([q], thetas[n + 11]), [(0, n + 11)])
        add(g.XX([0, n - 1], thetas[n + 12]), [(0, n + 12)])
    return ops, pmap


@pytest.mark.parametrize("dtype,etol,gtol", [("complex128", 1e-10, 1e-7),
                                             ("complex64", 1e-5, 1e-4)])
@pytest.mark.parametrize("n", [6, 9, 12])
@pytest.mark.parametrize("checkpoints", [0, 2, 5])
def test_adjoint_pass_planning_vs_finite_differences(C, n, checkpoints, dtype, etol, gtol):
    """Diagonal runs (split at 64 gates and 32 derivatives, shared parameters
    finalised in order), one-target groups on high bits (a lane per group) and
    on bits 0-3 (pairs across lanes; at 6 qubits the state is one wave),
    checkpoint gates that cut runs and groups, and 6- and 9-qubit states too
    small for runs: gradients against central differences of the CPU oracle
    (complex128 oracle for both precisions; the complex64 bar is looser by
    the FP32 rounding of the states)."""
    from snuqs_amd.circuits import random_pauli_sum

    rng = np.random.default_rng(n + checkpoints)
    nparams = n + 13
    thetas = rng.uniform(0, 2 * np.pi, nparams)
    ops, pmap = _planner_circuit(n, thetas)
    obs = random_pauli_sum(n, 6, 3, np.random.default_rng(4))
    e, grads = _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints)
    assert abs(e - oracle_energy(n, ops, *obs)) < etol
    h = 1e-6
    for p in range(nparams):
        up, dn = thetas.copy(), thetas.copy()
        up[p] += h
        dn[p] -= h
        fd = (oracle_energy(n, _planner_circuit(n, up)[0], *obs) -
              oracle_energy(n, _planner_circuit(n, dn)[0], *obs)) / (2 * h)
        assert abs(grads[p] - fd) < gtol, (p, grads[p], fd)
