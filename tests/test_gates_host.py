"""Host gate objects of the product vs the oracle's restatement of
gate_operations.cc (no GPU needed: the module loads without touching it)."""
import numpy as np
import pytest

from oracle import gates as G


@pytest.fixture(scope="module")
def gm():
    from snuqs_amd import _C

    return _C.operation.gate_operations


def test_matrices_bitwise_equal_restatement(gm):
    rng = np.random.default_rng(0)
    for name, (f, k, na) in G.GATES.items():
        ang = list(rng.uniform(-7, 7, na))
        op = getattr(gm, name)(list(range(k)), *ang) if na else getattr(gm, name)(list(range(k)))
        assert np.array_equal(np.array(op, copy=False), f(*ang)), name
        assert op.name == name


def test_buffer_protocol_and_repr(gm):
    op = gm.RotX([3], 0.25)
    a = np.array(op, copy=False)
    assert a.shape == (2, 2) and a.dtype == np.complex128
    assert repr(op) == "RotX(0.25) {3}"
    assert op.targets == [3]
    op.targets = [1]
    assert op.targets == [1]
    assert gm.CX([0, 1]) == gm.CX([0, 1]) and not (gm.CX([0, 1]) == gm.CX([1, 0]))


def test_derivatives_vs_finite_differences(gm):
    rng = np.random.default_rng(1)
    h = 1e-6
    for name, (f, k, na) in G.GATES.items():
        for i in range(na):
            ang = rng.uniform(0, 6, na)
            d = getattr(gm, name)(list(range(k)), *ang).derivative(i)
            ap, am = ang.copy(), ang.copy()
            ap[i] += h
            am[i] -= h
            fd = (f(*ap) - f(*am)) / (2 * h)
            assert np.abs(d - fd).max() < 1e-8, (name, i)


def test_ctrl_modifiers_and_power(gm):
    x = G.pauli_x()
    cx = np.array(gm.PauliX([0, 1], ctrl_modifiers=[1]))
    assert np.array_equal(cx, G.cx())
    neg = np.array(gm.PauliX([0, 1], ctrl_modifiers=[0]))
    want = np.eye(4, dtype=complex)
    want[:2, :2] = x
    assert np.array_equal(neg, want)
    assert np.allclose(np.array(gm.T([0], power=2)), G.s())
    assert np.allclose(np.array(gm.S([0], power=-1)), G.si())
    cc = np.array(gm.PauliX([0, 1, 2], ctrl_modifiers=[1, 1]))
    assert np.array_equal(cc, G.ccnot())
    gp = np.array(gm.GPhase([0, 1], 0.3))
    assert np.allclose(gp, np.exp(0.3j) * np.eye(4))


def test_unitary_validation(gm):
    u = gm.Unitary([0, 2], np.eye(4))
    assert np.array_equal(np.array(u), np.eye(4))
    with pytest.raises(Exception):
        gm.Unitary([0], np.array([[1, 1], [0, 1]]))
    with pytest.raises(Exception):
        gm.Unitary([0, 1], np.eye(2))
    with pytest.raises(Exception):
        gm.CX([0])
