"""OpenQASM 3 frontend (snuqs_amd/openqasm.py) on the CPU.

Each program is run through the CPU oracle from a random entangled state and
compared with a second program that states the same operator through
closed-form identities (ctrl @ x = cnot, pow(1/2) @ z = s, inv @ t = ti, a
user-defined gate = the built-in it spells out, ...), so a dropped control, a
swapped control order, a wrong power branch or a reversed inverse shows up as
a state difference.  The reference's own OpenQASM interpreter cannot run here
(its ANTLR runtime and the Braket schemas are absent), so these identities,
not its outputs, pin the frontend.
"""
import math

import numpy as np
import pytest

from helpers import run_oracle
from snuqs_amd.openqasm import QasmError, basis_rotations, parse, unitary_power


def _random_state(n, seed=7):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


def _run(body, n, inputs=None, init=None):
    c = parse(f"OPENQASM 3.0;\nqubit[{n}] q;\n{body}", inputs)
    return run_oracle(n, c.instructions, init=_random_state(n) if init is None else init)


def _same(a, b, n=3, inputs=None, up_to_phase=False):
    x, y = _run(a, n, inputs), _run(b, n, inputs)
    if up_to_phase:
        ov = np.vdot(y, x)
        y = y * ov / abs(ov)
    assert np.abs(x - y).max() < 1e-12, (a, b)


IDENTITIES = [
    ("ctrl @ x q[0], q[1];", "cnot q[0], q[1];"),
    ("ctrl @ x q[2], q[0];", "cnot q[2], q[0];"),
    ("negctrl @ x q[0], q[1];", "x q[0]; cnot q[0], q[1]; x q[0];"),
    ("ctrl @ negctrl @ x q[0], q[1], q[2];", "x q[1]; ccnot q[0], q[1], q[2]; x q[1];"),
    ("negctrl @ ctrl @ x q[0], q[1], q[2];", "x q[0]; ccnot q[0], q[1], q[2]; x q[0];"),
    ("ctrl(2) @ x q[0], q[1], q[2];", "ccnot q[0], q[1], q[2];"),
    ("ctrl @ ctrl @ x q[2], q[1], q[0];", "ccnot q[2], q[1], q[0];"),
    ("ctrl @ cnot q[0], q[1], q[2];", "ccnot q[0], q[1], q[2];"),
    ("ctrl @ swap q[1], q[0], q[2];", "cswap q[1], q[0], q[2];"),
    ("ctrl @ z q[0], q[2];", "cz q[0], q[2];"),
    ("ctrl @ phaseshift(0.3) q[1], q[2];", "cphaseshift(0.3) q[1], q[2];"),
    ("negctrl @ phaseshift(0.3) q[1], q[2];", "cphaseshift01(0.3) q[1], q[2];"),
    ("inv @ s q[0];", "si q[0];"),
    ("inv @ t q[1];", "ti q[1];"),
    ("inv @ v q[2];", "vi q[2];"),
    ("inv @ inv @ t q[1];", "t q[1];"),
    ("inv @ rx(0.3) q[0];", "rx(-0.3) q[0];"),
    ("inv @ U(0.1, 0.2, 0.3) q[0];", "U(-0.1, -0.3, -0.2) q[0];"),
    ("inv @ xy(0.4) q[0], q[2];", "xy(-0.4) q[0], q[2];"),
    ("pow(2) @ s q[0];", "z q[0];"),
    ("pow(3) @ t q[0];", "s q[0]; t q[0];"),
    ("pow(-2) @ t q[0];", "si q[0];"),
    ("pow(1/2) @ z q[1];", "s q[1];"),
    ("pow(0.5) @ s q[1];", "t q[1];"),
    ("pow(1/2) @ x q[2];", "v q[2];"),
    ("pow(-1/2) @ x q[2];", "vi q[2];"),
    ("pow(1/3) @ rz(0.9) q[0];", "rz(0.3) q[0];"),
    ("pow(0.25) @ phaseshift(1.2) q[0];", "phaseshift(0.3) q[0];"),
    ("pow(1/2) @ zz(0.8) q[0], q[1];", "zz(0.4) q[0], q[1];"),
    ("ctrl @ pow(1/2) @ x q[0], q[1];", "ctrl @ v q[0], q[1];"),
    ("pow(1/2) @ ctrl @ x q[0], q[1];", "ctrl @ v q[0], q[1];"),
    ("inv @ ctrl @ s q[1], q[0];", "ctrl @ si q[1], q[0];"),
    ("ctrl @ gphase(0.4) q[1];", "phaseshift(0.4) q[1];"),
    ("negctrl @ gphase(0.4) q[1];", "x q[1]; phaseshift(0.4) q[1]; x q[1];"),
    ("ctrl(2) @ gphase(0.4) q[0], q[2];", "cphaseshift(0.4) q[0], q[2];"),
    ("inv @ ctrl @ gphase(0.4) q[0];", "phaseshift(-0.4) q[0];"),
    ("h q;", "h q[0]; h q[1]; h q[2];"),
    ("rx(0.2) q[0:1];", "rx(0.2) q[0]; rx(0.2) q[1];"),
    ("cnot q[0], q[1:2];", "cnot q[0], q[1]; cnot q[0], q[2];"),
    ("x q[-1];", "x q[2];"),
    ("#pragma braket unitary([[0, 1], [1, 0]]) q[1]", "x q[1];"),
    ("#pragma braket unitary([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]) "
     "q[0], q[2]", "cnot q[0], q[2];"),
    ("#pragma braket unitary([[0.7071067811865476, 0.7071067811865476], "
     "[0.7071067811865476im, -0.7071067811865476im]]) q[2]", "h q[2]; s q[2];"),
    ("rx(2*pi/3) q[0];", "rx(2.0943951023931953) q[0];"),
    ("rz(-τ/4) q[0];", "rz(-1.5707963267948966) q[0];"),
    ("ry(sin(0.3)**2 + cos(0.3)**2) q[1];", "ry(1.0) q[1];"),
    ("phaseshift(2**-1) q[1];", "phaseshift(0.5) q[1];"),
    ("rx(-(1 - 3) * π / 8) q[1];", "rx(0.7853981633974483) q[1];"),
    ("gpi(0) q[0];", "x q[0];"),
    ("gpi2(0) q[0];", "rx(pi/2) q[0];"),
    ("prx(0.7, 0) q[0];", "rx(0.7) q[0];"),
    ("prx(0.7, pi/2) q[0];", "ry(0.7) q[0];"),
    ("ms(0, 0, 0.9) q[0], q[1];", "xx(0.9) q[0], q[1];"),
    ("cv q[0], q[2];", "ctrl @ v q[0], q[2];"),
    ("barrier q; box { h q[0]; }", "h q[0];"),
    ("// h q[0];\n/* x q[1];\n y q[2]; */ z q[0];", "z q[0];"),
]


@pytest.mark.parametrize("a,b", IDENTITIES)
def test_modifier_and_expression_identities(a, b):
    _same(a, b)


USER_GATES = """
gate myzz(t) a, b { cnot a, b; rz(t) b; cnot a, b; }
gate myswap a, b { cnot a, b; cnot b, a; cnot a, b; }
gate g2 a, b { myswap a, b; h a; }
gate hs a { h a; s a; }
"""


@pytest.mark.parametrize("a,b", [
    ("myzz(0.7) q[0], q[2];", "zz(0.7) q[0], q[2];"),
    ("myswap q[1], q[2];", "swap q[1], q[2];"),
    ("ctrl @ myswap q[2], q[0], q[1];", "cswap q[2], q[0], q[1];"),
    ("negctrl @ myswap q[2], q[0], q[1];", "x q[2]; cswap q[2], q[0], q[1]; x q[2];"),
    ("inv @ g2 q[0], q[1];", "h q[0]; swap q[0], q[1];"),
    ("pow(2) @ g2 q[0], q[1];", "swap q[0], q[1]; h q[0]; swap q[0], q[1]; h q[0];"),
    ("pow(-1) @ myzz(0.7) q[0], q[1];", "zz(-0.7) q[0], q[1];"),
    ("pow(1/2) @ myzz(0.8) q[0], q[1];", "zz(0.4) q[0], q[1];"),
    ("ctrl @ myzz(0.5) q[2], q[0], q[1];",
     "ctrl @ cnot q[2], q[0], q[1]; ctrl @ rz(0.5) q[2], q[1]; ctrl @ cnot q[2], q[0], q[1];"),
    ("pow(2) @ hs q[1];", "h q[1]; s q[1]; h q[1]; s q[1];"),
    ("myzz(0.7) q;", None),
])
def test_user_defined_gates(a, b):
    if b is None:  # two-qubit gate broadcast over a 3-qubit register is an error
        with pytest.raises((ValueError, QasmError)):
            _run(USER_GATES + a, 3)
        return
    _same(USER_GATES + a, b)


def test_pow_of_a_gate_body_is_not_the_body_of_powers():
    # pow(2) @ (h; s) = h s h s; raising each statement instead gives h^2 s^2 = z,
    # which the reference's modify_body computes (_helpers/quantum.py:93-109)
    x = _run(USER_GATES + "pow(2) @ hs q[0];", 1)
    y = _run("z q[0];", 1)
    assert np.abs(x - y).max() > 0.1


def test_ecr_matches_its_echoed_cross_resonance_decomposition():
    # ECR = RZX(pi/4) (X on the control) RZX(-pi/4), RZX built from cnot-rz-cnot
    # conjugated by h on the target; equal up to a global phase
    _same("ecr q[0], q[1];",
          "h q[1]; cnot q[0], q[1]; rz(π/4) q[1]; cnot q[0], q[1]; h q[1]; x q[0]; "
          "h q[1]; cnot q[0], q[1]; rz(-π/4) q[1]; cnot q[0], q[1]; h q[1];",
          up_to_phase=True)


def test_global_phase():
    psi = _random_state(2)
    x = _run("gphase(0.7); h q[0];", 2, init=psi.copy())
    y = _run("h q[0];", 2, init=psi.copy())
    assert np.abs(x - np.exp(0.7j) * y).max() < 1e-12
    x = _run("inv @ gphase(0.7) q[0];", 2, init=psi.copy())
    assert np.abs(x - np.exp(-0.7j) * psi).max() < 1e-12


def test_inputs_consts_and_physical_qubits():
    _same("input float theta; const float h2 = theta / 2; rx(h2) q[0];", "rx(0.35) q[0];",
          inputs={"theta": 0.7})
    c = parse("OPENQASM 3.0;\nh $0;\ncnot $0, $3;\n")
    assert sorted(c.qubit_set) == [0, 3]
    with pytest.raises(QasmError):
        parse("OPENQASM 3.0;\ninput float theta;\nqubit q;\nrx(theta) q;\n")
    with pytest.raises(QasmError):
        parse("OPENQASM 3.0;\nqubit[2] q;\nh $0;\n")


def test_registers_are_allocated_in_declaration_order():
    c = parse("OPENQASM 3.0;\nqubit[2] a;\nqreg b[3];\nqubit c;\nx b[1];\nx c;\nx a[1];\n")
    assert [list(op.targets) for op in c.instructions] == [[3], [5], [1]]
    _same("qubit[1] r; cnot q[0:1], q[1:2];", "cnot q[0], q[1]; cnot q[1], q[2];")


def test_results_and_measurements():
    c = parse("""OPENQASM 3.0;
bit[3] b;
qubit[3] q;
h q[0];
b[0] = measure q[0];
b[1:2] = measure q[1:2];
#pragma braket result expectation z(q[0]) @ x(q[1])
#pragma braket result variance h(q[2])
#pragma braket result expectation y all
#pragma braket result probability q[0], q[2]
#pragma braket result probability
#pragma braket result amplitude "000", "101"
#pragma braket result state_vector
""")
    assert c.measured_qubits == [0, 1, 2]
    assert c.results == [
        {"type": "expectation", "observable": ["z", "x"], "targets": [0, 1]},
        {"type": "variance", "observable": ["h"], "targets": [2]},
        {"type": "expectation", "observable": ["y"]},
        {"type": "probability", "targets": [0, 2]},
        {"type": "probability"},
        {"type": "amplitude", "states": ["000", "101"]},
        {"type": "statevector"},
    ]
    c = parse("OPENQASM 3.0;\nqubit[2] q;\nbit[2] c;\nmeasure q[1];\nc = measure q[0];\n")
    assert c.measured_qubits == [1, 0]
    with pytest.raises(ValueError):
        parse("OPENQASM 3.0;\nqubit[2] q;\nmeasure q[1];\nmeasure q;\n")


def test_basis_rotations_map_eigenstates_to_zero():
    s2 = math.sqrt(0.5)
    cases = {"x": [s2, s2], "y": [s2, 1j * s2],
             "h": [math.cos(math.pi / 8), math.sin(math.pi / 8)]}
    for letter, vec in cases.items():
        rot = basis_rotations([{"type": "expectation", "observable": [letter],
                                "targets": [0]}], [0])
        out = run_oracle(1, rot, init=np.array(vec, dtype=np.complex128))
        assert abs(abs(out[0]) - 1) < 1e-12, letter
    assert basis_rotations([{"type": "expectation", "observable": ["z"], "targets": [0]}],
                           [0]) == []
    with pytest.raises(ValueError):
        basis_rotations([{"type": "expectation", "observable": ["x"], "targets": [0]},
                         {"type": "variance", "observable": ["y"], "targets": [0]}], [0])


def test_unitary_power_closed_forms():
    from scipy.stats import unitary_group

    u = unitary_group.rvs(8, random_state=3)
    r = unitary_power(u, 0.5)
    assert np.abs(r @ r - u).max() < 1e-12
    assert np.abs(unitary_power(u, -1) - u.conj().T).max() < 1e-12
    assert np.abs(unitary_power(u, 1) - u).max() < 1e-12
    # eigenvalue -1 takes the +pi branch: sqrt(Z) = S, not S^dagger
    assert np.allclose(unitary_power(np.diag([1, -1]), 0.5), np.diag([1, 1j]))


@pytest.mark.parametrize("src,exc", [
    ("foo q[0];", QasmError),
    ("cnot q[0];", ValueError),
    ("h q[3];", QasmError),
    ("cnot q[0], q[0];", ValueError),
    ("rx q[0];", ValueError),
    ("reset q[0];", NotImplementedError),
    ("if (b) x q[0];", NotImplementedError),
    ("#pragma braket noise bit_flip(0.1) q[0]", NotImplementedError),
    ("#pragma braket unitary([[1, 1], [0, 1]]) q[0]", ValueError),
    ("#pragma braket unitary([[0, 1], [1, 0]]) q[0], q[1]", ValueError),
    ("#pragma braket result density_matrix", NotImplementedError),
    ("#pragma braket result expectation hermitian([[1, 1], [0, -1]]) q[0]", ValueError),
    ("#pragma braket result expectation hermitian([[1, 0], [0, -1]]) q[0], q[1]", ValueError),
    ("#pragma braket result expectation foo(q[0])", NotImplementedError),
    ("h q[0]", QasmError),
    ("ctrl(3) @ x q[0], q[1];", ValueError),
])
def test_errors(src, exc):
    with pytest.raises(exc):
        parse(f"OPENQASM 3.0;\nqubit[3] q;\n{src}\n")


def test_dual_numbers_differentiate_angle_expressions():
    """d angle / d input recorded per gate matches central differences of the
    same expression (forward-mode rules for + - * / ** and the functions)."""
    exprs = ["2*theta", "sin(phi)", "theta/phi", "theta**2", "exp(-theta)*cos(phi)",
             "sqrt(phi) - arctan(theta)", "2**theta", "-(theta - 3*phi)/4", "log(phi)*π",
             "arcsin(theta) + arccos(theta/2) + tan(phi)", "abs(theta - phi) + floor(phi)",
             "pow(phi, 3) + mod(theta, 0.2) + ceiling(theta)",
             # the divisor's term of a % b: -floor(a/b) db
             "mod(2.0, theta)", "2.0 % theta", "phi % theta", "mod(theta + 1, phi)"]
    x0 = {"theta": 0.3, "phi": 0.7}
    h = 1e-6

    def angle(expr, vals):
        c = parse(f"OPENQASM 3.0;\ninput float theta;\ninput float phi;\nqubit q;\n"
                  f"rx({expr}) q;\n", vals)
        return c.instructions[0].angles[0], dict(c.input_deps[0][0][1]) if c.input_deps[0] else {}

    for expr in exprs:
        _, d = angle(expr, x0)
        for name in ("theta", "phi"):
            up, _ = angle(expr, {**x0, name: x0[name] + h})
            dn, _ = angle(expr, {**x0, name: x0[name] - h})
            fd = (up - dn) / (2 * h)
            assert abs(d.get(name, 0.0) - fd) < 1e-7, (expr, name, d, fd)


def test_dual_power_at_base_zero():
    """x**e at x = 0: slope 1 for e = 1, 0 for e > 1 or e = 0, unbounded below 1."""
    from snuqs_amd.openqasm import Dual

    x = Dual(0.0, {"t": 1.0})
    assert (x ** 1).d["t"] == 1.0
    assert (x ** 2).d["t"] == 0.0
    assert (x ** 0).d["t"] == 0.0
    assert (x ** 0.5).d["t"] == float("inf")
    assert (x ** 1).v == 0.0 and (x ** 0).v == 1.0
    # the exponent's derivative at base 0 (e > 0) is 0
    assert (0.0 ** Dual(2.0, {"e": 1.0})).d["e"] == 0.0


def test_input_dependencies_through_modifiers_and_user_gates():
    c = parse("""OPENQASM 3.0;
input float theta;
input float phi;
qubit[3] q;
gate layer(a) x, y { rx(2*a) x; ctrl @ rz(a) x, y; }
inv @ rz(theta + phi) q[2];
pow(2) @ xx(theta) q[1], q[2];
ctrl @ gphase(phi) q[2];
inv @ layer(phi) q[0], q[1];
pow(1/2) @ rx(phi) q[1];
h q[0];
#pragma braket result adjoint_gradient expectation(-6 * y(q[0]) @ i(q[1]) + 0.75 * y(q[2]) @ z(q[1]) - h(q[0])) theta, phi
""", {"theta": 0.3, "phi": 0.7})
    deps = c.input_deps
    assert deps[0] == [(0, {"theta": -1.0, "phi": -1.0})]      # inv @ rz = rz(-t)
    assert deps[1] == [(0, {"theta": 2.0})]                     # pow(2) @ xx = xx(2 t)
    assert deps[2] == [(0, {"phi": 1.0})]                       # controlled phase
    # inv @ layer: reversed body, each statement inverted
    assert [op.name for op in c.instructions[3:5]] == ["RotZ", "RotX"]
    assert deps[3] == [(0, {"phi": -1.0})] and deps[4] == [(0, {"phi": -2.0})]
    assert deps[5] == {"phi"}                                   # fractional power: not differentiable
    assert deps[6] == []
    assert c.results[0] == {"type": "adjoint_gradient", "parameters": ["theta", "phi"],
                            "observable": [(-6.0, ["y", "i"], [0, 1]),
                                           (0.75, ["y", "z"], [2, 1]), (-1.0, ["h"], [0])]}
    with pytest.raises(QasmError):
        parse("OPENQASM 3.0;\ninput float t;\nqubit q;\nrx(t) q;\n"
              "#pragma braket result adjoint_gradient expectation(z(q)) s\n", {"t": 0.1})


@pytest.mark.parametrize("a,b", [
    ("pow(3) @ rx(0.4) q[0];", "rx(1.2000000000000002) q[0];"),
    ("pow(-2) @ cphaseshift(0.3) q[0], q[1];", "cphaseshift(-0.6) q[0], q[1];"),
    ("pow(2) @ ctrl @ yy(0.25) q[2], q[0], q[1];", "ctrl @ yy(0.5) q[2], q[0], q[1];"),
    ("pow(2) @ pswap(0.3) q[0], q[1];", "pswap(0.3) q[0], q[1]; pswap(0.3) q[0], q[1];"),
])
def test_integer_powers_of_rotations_scale_the_angle(a, b):
    _same(a, b)


def test_hermitian_observables_by_diagonalization():
    """<psi|O|psi> of a tensor product with Hermitian factors equals the
    eigenvalue-weighted distribution after the factors' diagonalizing gates
    (what the GPU path computes), for random Hermitians on 1 and 2 qubits
    next to Pauli and H factors; and the OpenQASM form parses to the same."""
    from snuqs_amd.openqasm import observable_factor

    rng = np.random.default_rng(11)
    n = 4
    psi = _random_state(n, seed=3)

    def rand_herm(d):
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        return (a + a.conj().T) / 2

    def full_op(factors, targets):
        # Kronecker product in qubit order 0..n-1 (qubit 0 = MSB)
        mats = {"x": np.array([[0, 1], [1, 0]]), "y": np.array([[0, -1j], [1j, 0]]),
                "z": np.diag([1, -1]), "i": np.eye(2),
                "h": np.array([[1, 1], [1, -1]]) / np.sqrt(2)}
        ops, k = [], 0
        for f in factors:
            m = mats[f] if isinstance(f, str) else f
            nq = m.shape[0].bit_length() - 1
            ops.append((m, targets[k:k + nq]))
            k += nq
        big = np.eye(1 << n, dtype=complex)
        for m, qs in ops:
            big = _embed(m, qs, n) @ big
        return big

    for factors, targets in [([rand_herm(2)], [2]), ([rand_herm(4)], [3, 1]),
                             ([rand_herm(2), "x", rand_herm(4)], [1, 0, 3, 2]),
                             (["h", rand_herm(2), "y"], [3, 0, 2])]:
        want = np.vdot(psi, full_op(factors, targets) @ psi).real
        rot, lam, k = [], np.ones(1), 0
        for f in factors:
            nq, make, eig = observable_factor(f)
            rot += make(list(targets[k:k + nq]))
            k += nq
            lam = np.kron(lam, eig)
        out = run_oracle(n, rot, init=psi.copy())
        p = (np.abs(out) ** 2).reshape([2] * n).transpose(
            targets + [q for q in range(n) if q not in targets]).reshape(1 << len(targets), -1)
        assert abs(p.sum(axis=1) @ lam - want) < 1e-12
    c = parse("OPENQASM 3.0;\nqubit[3] q;\n#pragma braket result variance "
              "hermitian([[2, 1im], [-1im, 0]]) q[1] @ x(q[2])\n")
    r = c.results[0]
    assert r["targets"] == [1, 2] and r["observable"][1] == "x"
    assert np.allclose(r["observable"][0], [[2, 1j], [-1j, 0]])
    rot = basis_rotations(c.results, [0, 1, 2])
    assert [op.name for op in rot] == ["Unitary", "Hadamard"]


def _embed(m, qs, n):
    """m on qubits qs (qs[0] the most significant) as a 2^n matrix."""
    k = len(qs)
    others = [q for q in range(n) if q not in qs]
    full = np.kron(m, np.eye(1 << (n - k)))
    order = list(qs) + others  # axis j of `full` is qubit order[j]
    perm = [order.index(q) for q in range(n)]
    t = full.reshape([2] * (2 * n)).transpose(perm + [n + p for p in perm])
    return t.reshape(1 << n, 1 << n)
