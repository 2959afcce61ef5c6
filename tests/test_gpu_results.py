"""Results taken from a device state vs the oracle: probabilities, norm,
amplitudes, inner products, Pauli-sum expectations and application, sampling."""
import numpy as np
import pytest

from helpers import device_state, random_state, run_engine, to_np
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _state(C, n, rng, dtype="complex128"):
    v = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    return v.astype(np.complex128), device_state(C, n, dtype, v)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_probabilities_norm_amplitudes(C, dtype):
    rng = np.random.default_rng(1)
    n = 14
    v, sv = _state(C, n, rng, dtype)
    p = C.functionals.probabilities(sv)
    # re^2 + im^2 in float64, no contraction: bitwise equal to the numpy form
    assert np.array_equal(p, v.real * v.real + v.imag * v.imag)
    assert abs(p.sum() - np.vdot(v, v).real) < 1e-12
    # the oracle's np.abs(sv)**2 form differs by at most an ulp or two
    assert np.abs(p - O.probabilities(v)).max() <= 4e-16 * p.max() + 1e-300
    assert abs(C.functionals.norm2(sv) - np.vdot(v, v).real) < 1e-13
    idx = [0, 5, (1 << n) - 1, 1234]
    amps = C.functionals.amplitudes(sv, idx)
    assert np.array_equal(amps, v[idx])


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_inner_and_scale(C, dtype):
    rng = np.random.default_rng(2)
    n = 13
    a, sa = _state(C, n, rng, dtype)
    b, sb = _state(C, n, rng, dtype)
    got = C.functionals.inner(sa, sb)
    assert abs(got - O.inner(a, b)) < 1e-13
    C.functionals.scale(sa, 0.5 - 0.25j)
    tol = 1e-15 if dtype == "complex128" else 1e-7
    assert np.abs(to_np(C, sa) - a * (0.5 - 0.25j)).max() < tol


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_renormalize(C, dtype):
    """renormalize scales by 1/sqrt(FP64 norm^2), the product formed in FP64 and
    rounded once: bitwise equal to numpy's complex128 product cast back."""
    rng = np.random.default_rng(3)
    n = 12
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    v = (random_state(n, rng) * 1.37).astype(npdt)
    sv = device_state(C, n, dtype, v)
    n2 = C.functionals.renormalize(sv)
    want_n2 = np.vdot(v.astype(np.complex128), v.astype(np.complex128)).real
    assert abs(n2 - want_n2) < 1e-13 * want_n2
    f = 1.0 / np.sqrt(n2)
    want = (v.astype(np.complex128) * f).astype(npdt)
    assert np.array_equal(np.asarray(sv.cpu()), want)
    p = C.functionals.probabilities(sv)
    assert abs(p.sum() - 1) < (1e-8 if dtype == "complex64" else 1e-14)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,targets", [(14, [3]), (14, [13, 0]), (14, [0, 1, 2, 3, 4, 5, 6, 7]),
                                       (14, [12, 13, 11, 10, 2, 9, 8, 0, 4]),
                                       (16, list(range(15, 3, -1))), (21, list(range(1, 21)))])
def test_marginal_probabilities(C, dtype, n, targets):
    """Marginal distribution over target qubits (targets[0] = most significant
    bit of the outcome) against numpy on the same amplitudes: <= 8 targets use
    the inserted-bit index map, more use the bit-by-bit deposit; 20 targets of
    21 qubits leave one block per outcome."""
    rng = np.random.default_rng(len(targets) + n)
    v, sv = _state(C, n, rng, dtype)
    got = C.functionals.marginal_probabilities(sv, targets)
    p = (v.real * v.real + v.imag * v.imag).reshape([2] * n)  # axis q = qubit q
    rest = tuple(q for q in range(n) if q not in targets)
    want = p.sum(axis=rest).transpose(np.argsort(np.argsort(targets))).reshape(-1)
    assert got.shape == (1 << len(targets),)
    assert np.abs(got - want).max() <= 1e-13 * max(1.0, want.max())
    assert abs(got.sum() - p.sum()) < 1e-12


def test_marginal_rejects_bad_targets(C):
    rng = np.random.default_rng(0)
    _, sv = _state(C, 6, rng)
    for bad in ([], [6], [1, 1], list(range(6)) * 4):
        with pytest.raises(ValueError):
            C.functionals.marginal_probabilities(sv, bad)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_probabilities_normalized(C, dtype):
    """normalize=True: |a|^2 / norm2 in FP64 with norm2 the device's FP64 norm,
    division correctly rounded (bitwise equal to numpy's), summing to 1."""
    rng = np.random.default_rng(6)
    n = 13
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    v = (random_state(n, rng) * 0.93).astype(npdt)
    sv = device_state(C, n, dtype, v)
    n2 = C.functionals.norm2(sv)
    p = C.functionals.probabilities(sv, normalize=True)
    w = v.astype(np.complex128)
    assert np.array_equal(p, (w.real * w.real + w.imag * w.imag) / n2)
    assert abs(p.sum() - 1) < 1e-14
    u = np.random.RandomState(8).random_sample(5000)
    assert np.array_equal(C.functionals.sample(sv, u, exact=True, normalize=True),
                          np.random.RandomState(8).choice(p.size, p=p, size=5000))


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_pauli_expectation(C, dtype):
    from snuqs_amd.circuits import random_pauli_sum

    rng = np.random.default_rng(3)
    n = 12
    v, sv = _state(C, n, rng, dtype)
    paulis, qubits, coeffs = random_pauli_sum(n, 40, 4, rng)
    paulis += ["ZZ", "X", "Y", "IZ", "XYZ"]
    qubits += [[0, n - 1], [3], [n - 1], [2, 5], [0, 1, 2]]
    coeffs += [0.3, -1.1, 0.7, 2.0, 0.5]
    total, terms = C.functionals.expectation(sv, paulis, qubits, coeffs)
    ref_terms = np.array([O.expect_pauli(v, p.replace("I", ""), [q for q, c in zip(qs, p) if c != "I"], n)
                          for p, qs in zip(paulis, qubits)])
    assert np.abs(ref_terms.imag).max() < 1e-12
    assert np.abs(terms - ref_terms.real).max() < 1e-12
    assert abs(total - float(np.dot(coeffs, ref_terms.real))) < 1e-11


def test_pauli_expectation_pinned_by_gate_application(C):
    """<P> from the closed-form kernel equals <psi| (X/Y/Z gates applied) psi>."""
    g = C.operation.gate_operations
    rng = np.random.default_rng(4)
    n = 9
    v, sv = _state(C, n, rng)
    for pauli, qs in [("XYZ", [0, 4, 8]), ("YY", [2, 3]), ("ZXZY", [8, 1, 5, 0])]:
        ops = [{"X": g.PauliX, "Y": g.PauliY, "Z": g.PauliZ}[c]([q]) for c, q in zip(pauli, qs)]
        pv = run_engine(C, n, ops, "complex128", v)
        want = np.vdot(v, pv)
        got, _ = C.functionals.expectation(sv, [pauli], [qs], [1.0])
        assert abs(got - want.real) < 1e-12 and abs(want.imag) < 1e-12


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_apply_pauli_sum(C, dtype):
    from snuqs_amd.circuits import random_pauli_sum

    rng = np.random.default_rng(5)
    n = 11
    v, sv = _state(C, n, rng, dtype)
    paulis, qubits, coeffs = random_pauli_sum(n, 16, 3, rng)
    out = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
    C.functionals.apply_pauli_sum(sv, out, paulis, qubits, coeffs)
    g = C.operation.gate_operations
    ref = np.zeros_like(v)
    for p, qs, c in zip(paulis, qubits, coeffs):
        ops = [{"X": g.PauliX, "Y": g.PauliY, "Z": g.PauliZ}[ch]([q]) for ch, q in zip(p, qs)]
        ref += c * run_engine(C, n, ops, "complex128", v)
    tol = 1e-12 if dtype == "complex128" else 1e-5
    assert np.abs(to_np(C, out) - ref).max() < tol


def _numpy_choice(p, seed, shots):
    return np.random.RandomState(seed).choice(len(p), p=p, size=shots)


@pytest.mark.parametrize("n", [4, 10, 16])
def test_sampling_exact_matches_numpy_choice(C, n):
    """Exact mode: identical indices to numpy RandomState.choice given the same
    probabilities and the same uniform draws (simulation.py:217-220)."""
    from snuqs_amd.circuits import random_gate_list

    rng = np.random.default_rng(n)
    ops = random_gate_list(n, 3 * n, rng)
    sv = device_state(C, n)
    for op in ops:
        C.functionals.apply(sv, op, n, op.targets)
    p = C.functionals.probabilities(sv)
    shots = 20000
    u = np.random.RandomState(77).random_sample(shots)
    want = _numpy_choice(p, 77, shots)
    got = C.functionals.sample(sv, u, exact=True)
    assert np.array_equal(got, want)
    # the oracle's restatement agrees too
    assert np.array_equal(O.choice(p, u), want)


def test_sampling_exact_adversarial_distribution(C):
    """Zeros, ties and tiny entries: uniform blocks, exact powers of two, a
    long run of zeros (GHZ-like) and subnormal-range values."""
    n = 12
    N = 1 << n
    rng = np.random.default_rng(9)
    p = np.zeros(N)
    p[5] = 0.25
    p[100:164] = 1.0 / 256
    p[1000:1512] = rng.random(512) * 1e-3
    p[3000] = 1e-300
    p[4000:4064] = 2.0 ** -20
    p /= p.sum()
    re = np.sqrt(p)
    sv = device_state(C, n, "complex128", re.astype(np.complex128))
    pd = C.functionals.probabilities(sv)
    u = np.random.RandomState(3).random_sample(50000)
    want = _numpy_choice(pd, 3, 50000)
    got = C.functionals.sample(sv, u, exact=True)
    assert np.array_equal(got, want)


def test_sampling_ghz_and_fast_mode(C):
    from snuqs_amd.circuits import ghz

    n = 20
    sv = device_state(C, n)
    for op in ghz(n):
        C.functionals.apply(sv, op, n, op.targets)
    u = np.random.RandomState(1).random_sample(10000)
    got = C.functionals.sample(sv, u, exact=True)
    assert set(np.unique(got)) <= {0, (1 << n) - 1}
    fast = C.functionals.sample(sv, u, exact=False)
    assert np.array_equal(fast, got)  # exact values: no rounding differences


def test_sampling_fast_mode_mismatches_are_rare(C):
    from snuqs_amd.circuits import random_gate_list

    n = 18
    rng = np.random.default_rng(18)
    sv = device_state(C, n)
    for op in random_gate_list(n, 60, rng):
        C.functionals.apply(sv, op, n, op.targets)
    p = C.functionals.probabilities(sv)
    u = np.random.RandomState(5).random_sample(100000)
    want = _numpy_choice(p, 5, 100000)
    fast = C.functionals.sample(sv, u, exact=False)
    mism = int(np.count_nonzero(fast != want))
    print("fast-mode mismatches:", mism)
    assert mism <= 20
    assert np.array_equal(C.functionals.sample(sv, u, exact=True), want)


def test_edge_cases(C):
    """1-qubit states, an empty program, empty queries, zero shots and a
    marginal over every qubit."""
    g = C.operation.gate_operations
    for dtype in ("complex128", "complex64"):
        sv = device_state(C, 1, dtype)
        C.functionals.apply(sv, g.Hadamard([0]), 1, [0])
        C.functionals.apply(sv, g.RotZ([0], 0.4), 1, [0])
        ref = np.array([np.exp(-0.2j), np.exp(0.2j)]) / np.sqrt(2)
        assert np.abs(to_np(C, sv) - ref).max() < 1e-7
        prog = C.functionals.Program(1, [], dtype=dtype)
        assert len(prog) == 0
        prog.run(sv)
        prog.run(sv, use_graph=True)
        assert np.abs(to_np(C, sv) - ref).max() < 1e-7
        assert C.functionals.amplitudes(sv, []).size == 0
        assert C.functionals.sample(sv, np.zeros(0)).size == 0
        p = C.functionals.probabilities(sv)
        assert np.allclose(p, [0.5, 0.5])
    rng = np.random.default_rng(2)
    v, sv = _state(C, 5, rng)
    m = C.functionals.marginal_probabilities(sv, [4, 3, 2, 1, 0])  # all qubits, reversed
    p = v.real * v.real + v.imag * v.imag
    rev = [int(format(i, "05b")[::-1], 2) for i in range(32)]
    assert np.abs(m - p[rev]).max() < 1e-15
