"""Adjoint layer passes (launch_adjoint_layer): layers of one-target gates on
distinct qubits applied 12 or up to 8 bits a pass in both sweeps, checked
against the CPU oracle and against the same gradient with the layer passes
switched off (register groups and single steps instead)."""
import numpy as np
import pytest

from helpers import device_state
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _energy_oracle(n, ops, paulis, qubits, coeffs):
    st = O.basis_state(n)
    for op in ops:
        O.apply(st, np.array(op.matrix()), op.targets, n)
    return sum(c * O.expect_pauli(st, p, qs, n).real for p, qs, c in zip(paulis, qubits, coeffs))


def _layer_circuit(n, thetas):
    """Layers of every kind: RotY on all qubits (own parameters), H on every
    other qubit (no derivative), RotX on a scattered subset with one shared
    parameter, a U layer (three parameters: not a layer gate) and CNOT
    ladders between them."""
    from snuqs_amd import _C as C

    g = C.operation.gate_operations
    ops, pmap = [], []

    def add(op, params=()):
        ops.append(op)
        pmap.append(list(params))

    k = 0
    for rep in range(2):
        for q in range(n):
            add(g.RotY([q], thetas[k]), [(0, k)])
            k += 1
        for q in range(0, n, 2):
            add(g.Hadamard([q]))
        for q in range(n - 1, -1, -3):
            add(g.RotX([q], thetas[k]), [(0, k)])
        k += 1
        for q in range(n - 1):
            add(g.CX([q, q + 1]))
        for q in range(1, n, 4):
            add(g.U([q], thetas[k], thetas[k + 1], thetas[k + 2]),
                [(0, k), (1, k + 1), (2, k + 2)])
        k += 3
        for q in range(n):
            add(g.RotZ([q], thetas[k]), [(0, k)])
            k += 1
    return ops, pmap, k


def _nparams(n):
    return _layer_circuit(n, np.zeros(10 * n))[2]


def _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, layers):
    saved = C.core.tune_get("adjoint_layers")
    C.core.tune_set("adjoint_layers", layers)
    try:
        prog = C.functionals.Program(n, ops, dtype=dtype)
        sv, w0, w1 = (device_state(C, n, dtype) for _ in range(3))
        cps = [device_state(C, n, dtype) for _ in range(checkpoints)]
        return prog.adjoint(sv, w0, w1, cps, pmap, nparams, *obs)
    finally:
        C.core.tune_set("adjoint_layers", saved)


@pytest.mark.parametrize("dtype,etol,gtol", [("complex128", 1e-10, 1e-7),
                                             ("complex64", 1e-5, 1e-4)])
@pytest.mark.parametrize("n,checkpoints", [(12, 0), (13, 3), (14, 2)])
def test_layer_passes_vs_oracle_finite_differences(C, n, checkpoints, dtype, etol, gtol):
    """12-14 qubits: the low pass (bits 0-11) and a 1-2 bit mid window, with
    checkpoint bounds moved to layer starts; energy and gradients against
    central differences of the FP64 oracle."""
    from snuqs_amd.circuits import random_pauli_sum

    nparams = _nparams(n)
    thetas = np.random.default_rng(n).uniform(0, 2 * np.pi, nparams)
    ops, pmap, _ = _layer_circuit(n, thetas)
    obs = random_pauli_sum(n, 6, 3, np.random.default_rng(5))
    e, grads = _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, 1)
    assert abs(e - _energy_oracle(n, ops, *obs)) < etol
    h = 1e-6
    for p in range(nparams):
        up, dn = thetas.copy(), thetas.copy()
        up[p] += h
        dn[p] -= h
        fd = (_energy_oracle(n, _layer_circuit(n, up)[0], *obs) -
              _energy_oracle(n, _layer_circuit(n, dn)[0], *obs)) / (2 * h)
        assert abs(grads[p] - fd) < gtol, (p, grads[p], fd)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-10), ("complex64", 5e-5)])
@pytest.mark.parametrize("n,checkpoints", [(16, 0), (20, 2), (22, 3), (23, 1)])
def test_layer_passes_match_groups(C, n, checkpoints, dtype, tol):
    """16-23 qubits (mid windows of 4 and 8 bits; at 22 and 23 qubits the two
    high windows overlap and each gate goes to one of them): layer passes on
    and off give the same energy and gradients."""
    from snuqs_amd.circuits import random_pauli_sum

    nparams = _nparams(n)
    thetas = np.random.default_rng(2 * n).uniform(0, 2 * np.pi, nparams)
    ops, pmap, _ = _layer_circuit(n, thetas)
    obs = random_pauli_sum(n, 8, 4, np.random.default_rng(n))
    e1, g1 = _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, 1)
    e0, g0 = _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, 0)
    print(f"{n}q {dtype}: energy diff {abs(e1 - e0):.2e}, max grad diff "
          f"{np.abs(g1 - g0).max():.2e}")
    assert abs(e1 - e0) < tol
    assert np.abs(g1 - g0).max() < tol


@pytest.mark.parametrize("n,dtype,tol", [(28, "complex128", 1e-11), (30, "complex64", 2e-5)])
def test_layer_passes_64bit_offsets(C, n, dtype, tol):
    """28 qubits complex128 (windows of 8 bits beside 4 low bits) and 30
    qubits complex64 (windows of 6 bits beside 6 low bits), 4 GiB and 8 GiB a
    state: the top window addresses past 2^31 bytes from a chunk base and
    takes the 64-bit path."""
    from snuqs_amd import _C

    g = _C.operation.gate_operations
    rng = np.random.default_rng(n)
    th = rng.uniform(0, 2 * np.pi, n)
    ops = [g.RotY([q], th[q]) for q in range(n)] + [g.CX([q, q + 1]) for q in range(0, n - 1, 2)]
    pmap = [[(0, q)] for q in range(n)] + [[] for _ in range(0, n - 1, 2)]
    obs = (["ZZ", "XZ", "YY"], [[0, n - 1], [3, 20], [11, 25]], [0.7, -1.1, 0.4])
    e1, g1 = _adjoint(C, n, ops, pmap, n, obs, dtype, 0, 1)
    e0, g0 = _adjoint(C, n, ops, pmap, n, obs, dtype, 0, 0)
    assert abs(e1 - e0) < tol
    assert np.abs(g1 - g0).max() < tol
    assert np.abs(g1).max() > 1e-3  # a non-trivial gradient


@pytest.mark.parametrize("n", [16, 22, 24])
def test_wide_windows_match_narrow(C, n):
    """Windows of up to 6 bits beside 6 contiguous low bits (512 B runs)
    against windows of up to 8 bits beside 4 (adjoint_wide_windows off) and
    the register groups (adjoint_layers off), complex128."""
    from snuqs_amd.circuits import random_pauli_sum

    nparams = _nparams(n)
    thetas = np.random.default_rng(3 * n).uniform(0, 2 * np.pi, nparams)
    ops, pmap, _ = _layer_circuit(n, thetas)
    obs = random_pauli_sum(n, 8, 4, np.random.default_rng(n + 1))
    saved = C.core.tune_get("adjoint_wide_windows")
    try:
        res = {}
        for wide in (1, 0):
            C.core.tune_set("adjoint_wide_windows", wide)
            res[wide] = _adjoint(C, n, ops, pmap, nparams, obs, "complex128", 2, 1)
    finally:
        C.core.tune_set("adjoint_wide_windows", saved)
    e0, g0 = _adjoint(C, n, ops, pmap, nparams, obs, "complex128", 2, 0)
    for e, g in res.values():
        assert abs(e - e0) < 1e-10
        assert np.abs(g - g0).max() < 1e-10


def _fuse_adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, fuse):
    saved = C.core.tune_get("adjoint_fuse_runs")
    C.core.tune_set("adjoint_fuse_runs", fuse)
    try:
        return _adjoint(C, n, ops, pmap, nparams, obs, dtype, checkpoints, 1)
    finally:
        C.core.tune_set("adjoint_fuse_runs", saved)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-10), ("complex64", 5e-5)])
@pytest.mark.parametrize("n,checkpoints", [(12, 0), (16, 3), (22, 2)])
def test_fused_runs_match_unfused(C, n, checkpoints, dtype, tol):
    """The ansatz's RotZ layer and CZ ring (a diagonal run) folded into the
    low pass of the RotY layer below it, forward, reverse and both (at 16 and
    22 qubits two CZs cross bit 12), against the unfused sweep and the
    register groups."""
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    layers = 3
    thetas = np.random.default_rng(n).uniform(0, 2 * np.pi, 2 * n * layers)
    ops, pmap = hea_ansatz(n, layers, thetas)
    obs = random_pauli_sum(n, 8, 4, np.random.default_rng(n + 7))
    e0, g0 = _adjoint(C, n, ops, pmap, thetas.size, obs, dtype, checkpoints, 0)
    for fuse in (0, 1, 2, 3):
        e, g = _fuse_adjoint(C, n, ops, pmap, thetas.size, obs, dtype, checkpoints, fuse)
        assert abs(e - e0) < tol, fuse
        assert np.abs(g - g0).max() < tol, fuse


def test_fused_runs_vs_oracle_and_declined(C):
    """13 qubits against central differences of the FP64 oracle; a run with
    more than four gates across bit 12 (CZ(q, q + 6) on 14 qubits) is not
    folded and still matches."""
    from snuqs_amd import _C
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    n, layers = 13, 2
    thetas = np.random.default_rng(13).uniform(0, 2 * np.pi, 2 * n * layers)
    ops, pmap = hea_ansatz(n, layers, thetas)
    obs = random_pauli_sum(n, 6, 3, np.random.default_rng(3))
    e, grads = _fuse_adjoint(C, n, ops, pmap, thetas.size, obs, "complex128", 2, 3)
    assert abs(e - _energy_oracle(n, ops, *obs)) < 1e-10
    h = 1e-6
    for p in range(thetas.size):
        up, dn = thetas.copy(), thetas.copy()
        up[p] += h
        dn[p] -= h
        fd = (_energy_oracle(n, hea_ansatz(n, layers, up)[0], *obs) -
              _energy_oracle(n, hea_ansatz(n, layers, dn)[0], *obs)) / (2 * h)
        assert abs(grads[p] - fd) < 1e-7, (p, grads[p], fd)
    g = _C.operation.gate_operations
    n = 14
    th = np.random.default_rng(14).uniform(0, 2 * np.pi, 2 * n)
    ops = [g.RotY([q], th[q]) for q in range(n)]
    ops += [g.RotZ([q], th[n + q]) for q in range(n)]
    ops += [g.CZ([q, q + 6]) for q in range(n - 6)]
    ops += [g.RotY([q], 0.3 * q) for q in range(n)]
    pmap = [[(0, q)] for q in range(2 * n)] + [[] for _ in range(2 * n - 6)]
    obs = random_pauli_sum(n, 6, 3, np.random.default_rng(4))
    e0, g0 = _fuse_adjoint(C, n, ops, pmap, 2 * n, obs, "complex128", 0, 0)
    e1, g1 = _fuse_adjoint(C, n, ops, pmap, 2 * n, obs, "complex128", 0, 3)
    assert abs(e1 - e0) < 1e-12 and np.abs(g1 - g0).max() < 1e-12
