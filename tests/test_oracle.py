"""Pins the CPU oracle (runs without a GPU).

The oracle restates cpu::applyGate (gate_operations_impl_cpu.cc:10-148); it is
checked against (1) the reference's own cpu::applyGate compiled from its
sources (oracle/_ref), bitwise; (2) golden vectors that build wrote; (3) a
Kronecker brute force and the tensordot formulation of linalg_utils.py:47-74;
(4) closed forms; (5) numpy's sampler through black-box fixtures.  A planted
mistake (transposed matrix, swapped target order) must fail the checks.
"""
import math
import os

import numpy as np
import pytest

from oracle import gates as G
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rand_state(n, rng):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def all_gates(rng):
    for name, (f, k, na) in G.GATES.items():
        yield name, f(*rng.uniform(0, 2 * np.pi, na)), k


def test_oracle_bitwise_equals_reference_cpu_path():
    if O.ref_lib() is None:
        pytest.skip("reference CPU build absent (oracle/_ref)")
    rng = np.random.default_rng(0)
    n = 9
    for name, m, k in all_gates(rng):
        for _ in range(4):
            t = [int(x) for x in rng.permutation(n)[:k]]
            sv = rand_state(n, rng)
            a, b = sv.copy(), sv.copy()
            O.apply(a, m, t, n)
            O.ref_apply(b, m, t, n)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (name, t)


def test_oracle_reproduces_reference_golden_vectors():
    d = np.load(os.path.join(GOLDEN, "ref_circuits.npz"))
    for n in (8, 11):
        ks, tg, mats = d[f"n{n}__k"], d[f"n{n}__targets"], d[f"n{n}__matrices"]
        st = O.basis_state(n)
        ti = mi = 0
        for k in ks:
            dim = 1 << int(k)
            m = mats[mi:mi + dim * dim].reshape(dim, dim)
            O.apply(st, m, list(tg[ti:ti + k]), n)
            ti += k
            mi += dim * dim
        assert np.array_equal(st, d[f"n{n}__state"])


@pytest.mark.parametrize("n", [5, 6])
def test_oracle_vs_kronecker_and_tensordot(n):
    rng = np.random.default_rng(n)
    for name, m, k in all_gates(rng):
        if k == 0:
            continue
        for _ in range(3):
            t = [int(x) for x in rng.permutation(n)[:k]]
            sv = rand_state(n, rng)
            a = sv.copy()
            O.apply(a, m, t, n)
            assert np.abs(a - O.kron_apply(sv, m, t, n)).max() < 1e-14, (name, t)
            assert np.abs(a - O.tensordot_apply(sv, m, t, n)).max() < 1e-14, (name, t)
    # 4- and 5-target dense blocks
    for k in (4, 5):
        z = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        q, _ = np.linalg.qr(z)
        t = [int(x) for x in rng.permutation(n)[:k]]
        sv = rand_state(n, rng)
        a = sv.copy()
        O.apply(a, q, t, n)
        assert np.abs(a - O.tensordot_apply(sv, q, t, n)).max() < 1e-14


def test_planted_mistakes_are_caught():
    """A transposed matrix or reversed target order must fail the brute force."""
    rng = np.random.default_rng(7)
    n = 5
    sv = rand_state(n, rng)
    m = G.u(0.3, 1.1, -0.4)
    bad = sv.copy()
    O.apply(bad, m.T, [2], n)
    assert np.abs(bad - O.kron_apply(sv, m, [2], n)).max() > 1e-3
    m2 = G.cy()
    bad = sv.copy()
    O.apply(bad, m2, [3, 1], n)
    assert np.abs(bad - O.kron_apply(sv, m2, [1, 3], n)).max() > 1e-3
    c64 = sv.astype(np.complex64)
    O.apply(c64, G.hadamard(), [0], n)
    ref = sv.copy()
    O.apply(ref, G.hadamard(), [0], n)
    assert np.abs(c64 - ref).max() < 1e-6


def test_gate_closed_forms_and_identities():
    I2 = np.eye(2)
    for name, (f, k, na) in G.GATES.items():
        m = f(*[0.37 * (j + 1) for j in range(na)])
        assert np.allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-15), name
    assert np.allclose(G.pauli_x() @ G.pauli_x(), I2)
    assert np.allclose(G.s() @ G.s(), G.pauli_z())
    assert np.allclose(G.t() @ G.t(), G.s())
    assert np.allclose(G.v() @ G.v(), G.pauli_x())
    assert np.allclose(G.vi() @ G.v(), I2)
    a = 0.81
    assert np.allclose(G.rot_z(a), np.exp(-0.5j * a) * G.phase_shift(a))
    assert np.allclose(G.u(math.pi, 0, math.pi), G.pauli_x())
    assert np.allclose(G.u(math.pi / 2, 0, math.pi), G.hadamard())
    assert np.allclose(G.u(a, 0, 0), G.rot_y(a))
    assert np.allclose(G.u(a, -math.pi / 2, math.pi / 2), G.rot_x(a))
    assert np.allclose(G.u(math.pi, math.pi / 2, math.pi / 2), G.pauli_y())
    assert G.hadamard()[0, 0].real == math.sqrt(0.5)


def _seq(*mats):
    """Product of 2-qubit operators applied left to right."""
    out = np.eye(4, dtype=complex)
    for m in mats:
        out = m @ out
    return out


def _on_a(m):
    return np.kron(m, np.eye(2))


def _on_b(m):
    return np.kron(np.eye(2), m)


def _equal_up_to_phase(a, b):
    i = np.unravel_index(np.argmax(np.abs(b)), b.shape)
    ph = a[i] / b[i]
    return abs(abs(ph) - 1) < 1e-12 and np.allclose(a, ph * b, atol=1e-12)


def test_braket_gate_decompositions():
    """braket_gates.inc:5-123 (up to global phase where the file adds gphase)."""
    a = 0.73
    cnot = G.cx()
    cnot_ba = G.swap() @ cnot @ G.swap()
    assert np.allclose(_seq(cnot, cnot_ba, cnot), G.swap())  # swap
    assert _equal_up_to_phase(
        _seq(cnot, _on_b(G.rot_z(a)), cnot), G.zz(a))  # zz
    h2 = np.kron(G.hadamard(), G.hadamard())
    assert _equal_up_to_phase(_seq(h2, cnot, _on_b(G.rot_z(a)), cnot, h2), G.xx(a))  # xx
    rx2 = np.kron(G.rot_x(math.pi / 2), G.rot_x(math.pi / 2))
    rxm2 = np.kron(G.rot_x(-math.pi / 2), G.rot_x(-math.pi / 2))
    assert _equal_up_to_phase(_seq(rx2, cnot, _on_b(G.rot_z(a)), cnot, rxm2), G.yy(a))  # yy
    x_a, x_b = _on_a(G.pauli_x()), _on_b(G.pauli_x())
    assert np.allclose(_seq(x_a, x_b, G.cphase_shift(a), x_a, x_b), G.cphase_shift00(a))
    assert np.allclose(_seq(x_b, G.cphase_shift(a), x_b), G.cphase_shift10(a))
    assert np.allclose(_seq(G.cphase_shift01(a), G.cphase_shift10(a), G.swap()), G.pswap(a))
    s2 = np.kron(G.s(), G.s())
    assert np.allclose(_seq(s2, _on_a(G.hadamard()), cnot, cnot_ba, _on_b(G.hadamard())),
                       G.iswap())  # iswap
    assert _equal_up_to_phase(
        _seq(_on_a(G.hadamard()), G.cy(), _on_a(G.rot_y(a / 2)), _on_b(G.rot_x(-a / 2)),
             G.cy(), _on_a(G.hadamard())), G.xy(a))  # xy


def test_numpy_choice_fixtures_pin_the_sampler():
    d = np.load(os.path.join(GOLDEN, "choice_fixtures.npz"))
    names = sorted({k.split("__")[0] for k in d.files})
    assert len(names) >= 5
    for name in names:
        p, u, idx = d[f"{name}__p"], d[f"{name}__u"], d[f"{name}__idx"]
        seed = int(d[f"{name}__seed"])
        # the fixture's uniforms are RandomState(seed).random_sample(k)
        assert np.array_equal(np.random.RandomState(seed).random_sample(u.size), u)
        assert np.array_equal(O.choice(p, u), idx), name
        # the restated cumsum equals numpy's cumsum bit for bit
        assert np.array_equal(O.cumsum(p), np.cumsum(p))


def test_expectation_formula_vs_gate_application():
    rng = np.random.default_rng(3)
    n = 6
    sv = rand_state(n, rng)
    for pauli, qs in [("XYZ", [0, 3, 5]), ("YY", [1, 2]), ("Z", [4]), ("XZYX", [5, 0, 2, 1])]:
        st = sv.copy()
        for ch, q in zip(pauli, qs):
            O.apply(st, {"X": G.pauli_x(), "Y": G.pauli_y(), "Z": G.pauli_z()}[ch], [q], n)
        want = np.vdot(sv, st)
        got = O.expect_pauli(sv, pauli, qs, n)
        assert abs(got - want) < 1e-14
    assert abs(O.inner(sv, sv) - 1) < 1e-14


def test_probabilities_match_numpy_form():
    """np.abs(sv)**2 (simulation.py:214): numpy's complex abs is its own
    routine (it differs from np.hypot in the last bit), so the restatement
    (hypot, then square) agrees to a few ulp, not bitwise."""
    rng = np.random.default_rng(4)
    sv = rand_state(10, rng)
    p, want = O.probabilities(sv), np.abs(sv) ** 2
    assert np.max(np.abs(p - want) / want) < 1e-15


def test_inner_product_conjugates_its_first_argument():
    """sqo_inner is <a|b> = sum conj(a_i) b_i, the convention of np.vdot.
    Checked on two different states with a large imaginary part, so that a
    conjugate on the wrong side (the planted <b|a>) or no conjugate at all
    fails the same comparison."""
    rng = np.random.default_rng(11)
    n = 9
    a = rand_state(n, rng)
    b = a * np.exp(0.7j) + 0.5 * rand_state(n, rng)  # <a|b> ~ exp(0.7i)
    b /= np.linalg.norm(b)
    want = np.vdot(a, b)
    assert abs(want.imag) > 0.3
    assert abs(O.inner(a, b) - want) < 1e-14
    # planted mistakes must fail the same bar
    assert abs(O.inner(b, a) - want) > 0.5           # conjugate-swap
    assert abs(complex(np.sum(a * b)) - want) > 0.5  # no conjugate
    # linearity in the second argument, antilinearity in the first
    c = 0.3 - 1.7j
    assert abs(O.inner(a, c * b) - c * want) < 1e-13
    assert abs(O.inner(c * a, b) - np.conj(c) * want) < 1e-13


class _Op:
    """A gate stand-in carrying only what the reorder reads."""

    def __init__(self, name, targets):
        self.name, self.targets = name, list(targets)

    def sliceable(self):
        return False


def _order(ops):
    return [op.name for op in O.pseudo_sort_operations_descending(ops)]


# Expected orders worked by hand from transpile.py:33-54 (pseudo_gt) and
# :74-89 (the insertion sort).  The sort passes (op, i) and (op, j) with
# j < i, so "pos1 < pos2" is always False there: gates sharing a qubit, or
# two zero-target gates, stop the insertion.
_PSEUDO_SORT_CASES = [
    # disjoint one-target gates: descending by target
    ("disjoint", [("a", [0]), ("b", [2]), ("c", [1])], ["b", "c", "a"]),
    # disjoint, min differs from max: B (min 1) passes A (min 0); by max
    # (2 vs 3) it would not
    ("min-not-max", [("A", [0, 3]), ("B", [1, 2])], ["B", "A"]),
    # insertion stops at the first gate that is not smaller
    ("stops", [("A", [3]), ("B", [1]), ("C", [2])], ["A", "C", "B"]),
    # zero targets on one side: the global phase moves to the front, and a
    # later gate does not pass it
    ("gphase", [("X0", [0]), ("G", []), ("H1", [1])], ["G", "H1", "X0"]),
    # two zero-target gates keep their order
    ("two-gphase", [("G1", []), ("X0", [0]), ("G2", [])], ["G1", "G2", "X0"]),
    # a shared qubit blocks the move: C stays after B
    ("shared", [("A", [0]), ("B", [2]), ("C", [2])], ["B", "C", "A"]),
    # shared qubit between two-target gates, with a free gate passing both
    ("shared2", [("X0", [0]), ("CX01", [0, 1]), ("H2", [2])], ["H2", "X0", "CX01"]),
]


@pytest.mark.parametrize("name,ops,want", _PSEUDO_SORT_CASES, ids=[c[0] for c in _PSEUDO_SORT_CASES])
def test_pseudo_sort_hand_worked_orders(name, ops, want):
    assert _order([_Op(nm, t) for nm, t in ops]) == want


def _planted_gt(kind):
    def gt(op1, pos1, op2, pos2, sliceable=lambda op: False):
        t1, t2 = op1.targets, op2.targets
        if len(t1) == 0 and len(t2) == 0:
            return pos1 < pos2
        if len(t1) == 0:
            return kind != "zero-flip"
        if len(t2) == 0:
            return kind == "zero-flip"
        if set(t1) & set(t2):
            return True if kind == "shared-moves" else pos1 < pos2
        if kind == "max":
            return max(t1) > max(t2)
        if kind == "lt":
            return min(t1) < min(t2)
        return min(t1) > min(t2)
    return gt


@pytest.mark.parametrize("kind", ["max", "lt", "zero-flip", "shared-moves"])
def test_pseudo_sort_planted_mistakes_fail(kind, monkeypatch):
    """Each planted slip in pseudo_gt changes at least one hand-worked order."""
    monkeypatch.setattr(O, "_pseudo_gt", _planted_gt(kind))
    failed = [name for name, ops, want in _PSEUDO_SORT_CASES
              if _order([_Op(nm, t) for nm, t in ops]) != want]
    assert failed, kind
