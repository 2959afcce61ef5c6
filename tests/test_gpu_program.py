"""Programs re-parametrised in place (Program.set_gates), adjoint buffer
checks, derivative matrices of 4-5 target gates, and the cached workspaces of
the torch wrapper."""
import numpy as np
import pytest

from helpers import device_state, run_oracle, to_np
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _param_circuit(C, n, th):
    """Rotations, controlled rotations with 3-4 targets and a fused block."""
    g = C.operation.gate_operations
    ops = [g.Hadamard([q]) for q in range(n)]
    ops += [g.RotY([q], th[q]) for q in range(n)]
    ops += [g.RotX([0, 1, 2, 3], th[n], ctrl_modifiers=[1, 1, 1]),
            g.RotY([4, 2, 5], th[n + 1], ctrl_modifiers=[0, 1]),
            g.XX([1, n - 1], th[n + 2]), g.ZZ([2, 3], th[n + 3])]
    return ops


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_set_gates_equals_a_new_program(C, dtype):
    from snuqs_amd.transpile import fuse_operations

    n = 11
    rng = np.random.default_rng(3)
    th1, th2 = rng.uniform(0, 2 * np.pi, (2, n + 4))
    for fuse in (False, True):
        ops1, ops2 = _param_circuit(C, n, th1), _param_circuit(C, n, th2)
        if fuse:  # 4-5 target blocks: their device slots are rewritten
            ops1, ops2 = fuse_operations(ops1, 5), fuse_operations(ops2, 5)
            assert [o.targets for o in ops1] == [o.targets for o in ops2]
            assert max(len(o.targets) for o in ops1) >= 4
        prog = C.functionals.Program(n, ops1, dtype=dtype)
        sv = device_state(C, n, dtype)
        prog.run(sv, use_graph=True)  # a graph captured with the first matrices
        prog.set_gates(ops2)
        for use_graph in (False, True):
            C.functionals.initialize_basis_z(sv)
            prog.run(sv, use_graph=use_graph)
            fresh = device_state(C, n, dtype)
            C.functionals.Program(n, ops2, dtype=dtype).run(fresh)
            assert np.array_equal(to_np(C, sv), to_np(C, fresh)), (fuse, use_graph)
        tol = 1e-12 if dtype == "complex128" else 1e-5
        assert np.abs(to_np(C, sv) - run_oracle(n, ops2)).max() < tol


def test_set_gates_rejects_other_targets(C):
    g = C.operation.gate_operations
    prog = C.functionals.Program(3, [g.RotX([0], 0.1), g.CZ([0, 1])])
    with pytest.raises(ValueError):
        prog.set_gates([g.RotX([1], 0.1), g.CZ([0, 1])])
    with pytest.raises(ValueError):
        prog.set_gates([g.RotX([0], 0.1)])


def test_adjoint_rejects_mismatched_buffers(C):
    """Every buffer is checked against the program before a kernel runs."""
    g = C.operation.gate_operations
    n = 4
    ops = [g.RotY([q], 0.3) for q in range(n)]
    pmap = [[(0, q)] for q in range(n)]
    obs = (["Z"], [[0]], [1.0])
    prog = C.functionals.Program(n, ops, dtype="complex128")
    good = lambda: device_state(C, n, "complex128")  # noqa: E731
    sv, w0, w1 = good(), good(), good()
    bad = {
        "dtype": device_state(C, n, "complex64"),
        "size": device_state(C, n - 1, "complex128"),
        "host": C.result_types.StateVector(n),
    }
    for name, b in bad.items():
        with pytest.raises(ValueError):
            prog.adjoint(b, w0, w1, [], pmap, n, *obs)
        with pytest.raises(ValueError):
            prog.adjoint(sv, b, w1, [], pmap, n, *obs)
        with pytest.raises(ValueError):
            prog.adjoint(sv, w0, b, [], pmap, n, *obs)
        with pytest.raises(ValueError):
            prog.adjoint(sv, w0, w1, [good(), b], pmap, n, *obs)
    e, gr = prog.adjoint(sv, w0, w1, [good()], pmap, n, *obs)
    assert np.isfinite(e) and np.all(np.isfinite(gr))


def _oracle_energy(n, ops, paulis, qubits, coeffs):
    st = run_oracle(n, ops)
    return sum(c * O.expect_pauli(st, p, q, n).real for p, q, c in zip(paulis, qubits, coeffs))


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-8), ("complex64", 1e-4)])
def test_adjoint_derivatives_of_controlled_rotations(C, dtype, tol):
    """3- and 4-target controlled rotations: their derivative matrices are
    dense blocks (the uncontrolled part is zero, not the identity), the
    4-target ones staged in the library's workspace."""
    n = 8
    rng = np.random.default_rng(8)
    th = rng.uniform(0, 2 * np.pi, n + 4)
    pmap = [[] for _ in range(n)] + [[(0, q)] for q in range(n)] + \
        [[(0, n)], [(0, n + 1)], [(0, n + 2)], [(0, n + 3)]]
    obs = (["XZ", "YX", "Z"], [[0, 3], [1, 7], [5]], [0.7, -0.4, 1.1])
    ops = _param_circuit(C, n, th)
    prog = C.functionals.Program(n, ops, dtype=dtype)
    st = [device_state(C, n, dtype) for _ in range(3)]
    e, grads = prog.adjoint(*st, [], pmap, th.size, *obs)
    assert abs(e - _oracle_energy(n, ops, *obs)) < (1e-12 if dtype == "complex128" else 1e-5)
    h = 1e-6
    for p in range(th.size):
        up, dn = th.copy(), th.copy()
        up[p] += h
        dn[p] -= h
        fd = (_oracle_energy(n, _param_circuit(C, n, up), *obs) -
              _oracle_energy(n, _param_circuit(C, n, dn), *obs)) / (2 * h)
        assert abs(grads[p] - fd) < tol, (p, grads[p], fd)
    # the same program re-parametrised gives the new gradient
    th2 = th[::-1].copy()
    prog.set_gates(_param_circuit(C, n, th2))
    e2, g2 = prog.adjoint(*st, [], pmap, th.size, *obs)
    fresh = C.functionals.Program(n, _param_circuit(C, n, th2), dtype=dtype)
    e3, g3 = fresh.adjoint(*st, [], pmap, th.size, *obs)
    assert e2 == e3 and np.array_equal(g2, g3)


def test_circuit_energy_reuses_its_workspace(C):
    import torch

    from snuqs_amd import autograd
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    autograd.clear_cache()
    n, layers = 7, 2
    rng = np.random.default_rng(1)
    obs = random_pauli_sum(n, 6, 2, rng)
    builder = lambda th: hea_ansatz(n, layers, th)  # noqa: E731
    for step in range(3):
        th0 = rng.uniform(0, 2 * np.pi, 2 * n * layers)
        theta = torch.tensor(th0, requires_grad=True)
        e = autograd.circuit_energy(theta, builder, obs)
        e.backward()
        assert len(autograd._CACHE) == 1
        ws = next(iter(autograd._CACHE.values()))
        if step == 0:
            first = ws
        assert ws is first  # same program and states every step
        ops, pmap = builder(th0)
        prog = C.functionals.Program(n, ops)
        st = [device_state(C, n) for _ in range(3)]
        e_ref, g_ref = prog.adjoint(*st, [], pmap, th0.size, *obs)
        assert e.item() == e_ref
        assert np.array_equal(theta.grad.numpy(), g_ref)
    autograd.clear_cache()
