"""The Braket entry point (StateVectorSimulator.run) and StateVectorSimulation
on JAQCD programs given as JSON dicts (the Braket SDK is not installed)."""
import numpy as np
import pytest

from helpers import run_oracle
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def ghz_program(n):
    ins = [{"type": "h", "target": 0}] + [{"type": "cnot", "control": 0, "target": i}
                                           for i in range(1, n)]
    return {"instructions": ins, "results": [{"type": "statevector"}]}


def test_run_statevector_ghz():
    from snuqs_amd.simulator import StateVectorSimulator

    res = StateVectorSimulator().run(ghz_program(6))
    sv = res.resultTypes[0]["value"]
    assert sv.shape == (64,)
    assert sv[0] == sv[-1] == np.sqrt(0.5)
    assert np.count_nonzero(sv[1:-1]) == 0


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_run_mixed_program_against_oracle(precision):
    """Every JAQCD instruction kind, sparse qubit labels (remapped contiguously
    as in simulator.py:278-307), result types computed on the GPU."""
    from snuqs_amd.circuits import haar_unitary
    from snuqs_amd.gate_operations import from_braket_instruction
    from snuqs_amd.simulator import StateVectorSimulator

    rng = np.random.default_rng(1)
    U = haar_unitary(4, rng)
    umat = [[[z.real, z.imag] for z in row] for row in U]
    ins = [
        {"type": "h", "target": 3}, {"type": "h", "target": 10}, {"type": "x", "target": 7},
        {"type": "rx", "target": 3, "angle": 0.3}, {"type": "cnot", "control": 3, "target": 7},
        {"type": "ccnot", "controls": [3, 7], "target": 10},
        {"type": "cswap", "control": 10, "targets": [3, 7]},
        {"type": "xx", "targets": [3, 10], "angle": 1.1},
        {"type": "cphaseshift01", "control": 7, "target": 10, "angle": 0.4},
        {"type": "unitary", "targets": [10, 3], "matrix": umat},
        {"type": "ry", "target": 7, "angle": -0.8}, {"type": "iswap", "targets": [7, 3]},
    ]
    results = [{"type": "statevector"}, {"type": "probability", "targets": [10]},
               {"type": "amplitude", "states": ["000", "101"]},
               {"type": "expectation", "observable": ["z", "x"], "targets": [3, 10]},
               {"type": "variance", "observable": ["y", "z"], "targets": [7, 3]},
               {"type": "probability", "targets": [10, 3]},
               {"type": "probability", "targets": [7, 10, 3]}]
    res = StateVectorSimulator().run({"instructions": ins, "results": results},
                                     precision=precision)
    qmap = {3: 0, 7: 1, 10: 2}
    ops = []
    for i in ins:
        op = from_braket_instruction(i)
        op.targets = [qmap[q] for q in op.targets]
        ops.append(op)
    ref = run_oracle(3, ops)
    tol = 1e-12 if precision == "complex128" else 1e-5
    vals = [r["value"] for r in res.resultTypes]
    assert np.abs(vals[0].astype(np.complex128) - ref).max() < tol
    p = np.abs(ref) ** 2
    assert np.abs(vals[1] - np.array([p[0::2].sum(), p[1::2].sum()])).max() < tol
    assert abs(vals[2]["101"] - ref[5]) < tol and abs(vals[2]["000"] - ref[0]) < tol
    want = O.expect_pauli(ref, "ZX", [0, 2], 3).real
    assert abs(vals[3] - want) < tol
    e_yz = O.expect_pauli(ref, "YZ", [1, 0], 3).real
    assert abs(vals[4] - (1.0 - e_yz ** 2)) < 4 * tol
    # marginal over (q2, q0): outcome bit 1 = qubit 2, bit 0 = qubit 0
    pm = p.reshape(2, 2, 2)  # axes: qubit 0, 1, 2 (qubit 0 = MSB of the index)
    assert np.abs(vals[5] - pm.sum(axis=1).T.reshape(-1)).max() < tol
    # a permutation of all qubits reorders the distribution
    assert np.abs(vals[6] - pm.transpose(1, 2, 0).reshape(-1)).max() < tol


def test_shots_match_numpy_choice_with_global_seed():
    """retrieve_samples draws uniforms from numpy's global RandomState exactly as
    np.random.choice(len(sv), p=probabilities, size=shots) (simulation.py:220)."""
    from snuqs_amd import _C
    from snuqs_amd.simulation import StateVectorSimulation
    from snuqs_amd.subcircuit import Subcircuit
    from snuqs_amd.circuits import random_gate_list

    n = 9
    ops = random_gate_list(n, 40, np.random.default_rng(3))
    sub = Subcircuit(qubit_count=n, operations=ops)
    sim = StateVectorSimulation(n, 5000, sub)
    sim.evolve(sub)
    p = sim.probabilities
    np.random.seed(123)
    got = sim.retrieve_samples()
    np.random.seed(123)
    want = np.random.choice(len(p), p=p, size=5000)
    assert np.array_equal(got, want)


def test_shots_formatting_and_errors():
    from snuqs_amd.simulator import StateVectorSimulator

    sim = StateVectorSimulator()
    res = sim.run(ghz_program(4), shots=100)
    assert len(res.measurements) == 100
    assert res.resultTypes == []  # JAQCD with shots: no result types (simulator.py:497-510)
    assert all("".join(m) in ("0000", "1111") for m in res.measurements)
    with pytest.raises(TypeError):
        sim.run(ghz_program(2), accelerator="cpu")
    with pytest.raises(TypeError):
        sim.run(ghz_program(2), offload="storage")


QASM = """OPENQASM 3.0;
input float theta;
bit[3] b;
qubit[4] q;
gate entangle(t) a, c { h a; cnot a, c; rz(t) c; }
entangle(theta) q[0], q[3];
ctrl @ ry(0.4) q[3], q[1];
negctrl @ pow(1/2) @ x q[1], q[0];
inv @ entangle(0.3) q[1], q[0];
#pragma braket unitary([[0.6, 0.8im], [0.8im, 0.6]]) q[3]
gphase(0.25);
b[0] = measure q[3];
b[1] = measure q[0];
b[2] = measure q[1];
"""


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_run_openqasm_against_oracle(precision):
    """OpenQASM 3 through StateVectorSimulator.run: the declared-but-unused
    qubit q[2] is dropped by the contiguous remap (simulator.py:278-307), the
    result pragmas are computed on the GPU and checked against the CPU oracle
    on the frontend's own gate list."""
    from snuqs_amd.openqasm import parse
    from snuqs_amd.simulator import StateVectorSimulator

    results = """#pragma braket result state_vector
#pragma braket result expectation z(q[0]) @ h(q[3])
#pragma braket result variance x(q[1])
#pragma braket result expectation z all
#pragma braket result probability q[3], q[0]
#pragma braket result amplitude "010", "111"
"""
    src = QASM.replace("b[0] = measure q[3];\nb[1] = measure q[0];\nb[2] = measure q[1];\n",
                       results)
    res = StateVectorSimulator().run({"source": src, "inputs": {"theta": 0.9}},
                                     precision=precision)
    c = parse(src, {"theta": 0.9})
    qmap = {0: 0, 1: 1, 3: 2}
    for op in c.instructions:
        op.targets = [qmap[q] for q in op.targets]
    ref = run_oracle(3, c.instructions)
    tol = 1e-12 if precision == "complex128" else 1e-5
    vals = [r["value"] for r in res.resultTypes]
    assert np.abs(vals[0].astype(np.complex128) - ref).max() < tol
    s2 = np.sqrt(0.5)
    want = s2 * (O.expect_pauli(ref, "ZX", [0, 2], 3) + O.expect_pauli(ref, "ZZ", [0, 2], 3))
    assert abs(vals[1] - want.real) < tol
    assert abs(vals[2] - (1 - O.expect_pauli(ref, "X", [1], 3).real ** 2)) < 4 * tol
    assert np.abs(np.array(vals[3]) -
                  [O.expect_pauli(ref, "Z", [q], 3).real for q in range(3)]).max() < tol
    pm = (np.abs(ref) ** 2).reshape(2, 2, 2)
    assert np.abs(vals[4] - pm.sum(axis=1).T.reshape(-1)).max() < tol
    assert abs(vals[5]["010"] - ref[2]) < tol and abs(vals[5]["111"] - ref[7]) < tol


def test_run_openqasm_shots_and_measured_qubits():
    """Measured qubits order the bit characters; a register-wide observable
    adds its basis rotation before sampling (circuit.py:95-143)."""
    from snuqs_amd.openqasm import parse
    from snuqs_amd.simulator import StateVectorSimulator

    sim = StateVectorSimulator()
    np.random.seed(5)
    res = sim.run({"source": QASM, "inputs": {"theta": 0.9}}, shots=2000)
    assert res.measuredQubits == [3, 0, 1]
    c = parse(QASM, {"theta": 0.9})
    qmap = {0: 0, 1: 1, 3: 2}
    for op in c.instructions:
        op.targets = [qmap[q] for q in op.targets]
    p = np.abs(run_oracle(3, c.instructions)) ** 2
    np.random.seed(5)
    idx = np.random.choice(8, p=p, size=2000)
    want = [[str((i >> (2 - q)) & 1) for q in (2, 0, 1)] for i in idx]
    assert res.measurements == want
    # GHZ measured in the X basis on every qubit: parity of the outcomes is even
    ghz = ("OPENQASM 3.0;\nqubit[3] q;\nh q[0];\ncnot q[0], q[1];\ncnot q[1], q[2];\n"
           "#pragma braket result expectation x all\n")
    res = sim.run(ghz, shots=500)
    assert all(sum(int(ch) for ch in m) % 2 == 0 for m in res.measurements)
    # Sample of X x X x Y: the rotation runs before sampling; with shots the
    # program's result list comes back unchanged (simulator.py:437-447), the
    # values being derived from the measurements by the client
    ghz_s = ghz.replace("expectation x all", "sample x(q[0]) @ x(q[1]) @ y(q[2])")
    res = sim.run(ghz_s, shots=400)
    assert res.resultTypes == parse(ghz_s).results and len(res.measurements) == 400
    with pytest.raises(ValueError):
        sim.run(ghz_s, shots=0)


ADJ = """OPENQASM 3.0;
input float theta;
input float phi;
qubit[4] q;
gate layer(a) x, y { rx(2*a) x; ctrl @ rz(a) x, y; cnot y, x; }
h q;
ry(theta) q[0];
inv @ layer(phi) q[1], q[2];
pow(2) @ xx(theta * phi) q[0], q[3];
ctrl @ gphase(sin(phi)) q[2];
negctrl @ ry(theta / 3) q[3], q[1];
U(theta, 0.2, phi) q[2];
layer(theta - phi) q[3], q[0];
#pragma braket result adjoint_gradient expectation(-0.6 * y(q[0]) @ i(q[1]) + 0.75 * x(q[2]) @ z(q[1]) - 1.5 * h(q[3]) @ z(q[0])) theta, phi
"""


def _oracle_energy(src, inputs):
    from snuqs_amd.openqasm import parse

    c = parse(src, inputs)
    psi = run_oracle(4, [op for op in c.instructions if list(op.targets)])
    e = 0.0
    for coef, letters, qs in c.results[0]["observable"]:
        terms = [("", [], coef)]
        for L, q in zip(letters, qs):
            if L == "i":
                continue
            opts, w = (("X", "Z"), np.sqrt(0.5)) if L == "h" else ((L.upper(),), 1.0)
            terms = [(s + o, t + [q], cc * w) for s, t, cc in terms for o in opts]
        for s, t, cc in terms:
            e += cc * O.expect_pauli(psi, s, t, 4).real
    return e


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_openqasm_adjoint_gradient_matches_finite_differences(precision):
    """Braket's AdjointGradient result through StateVectorSimulator.run:
    expectation and d/d(theta, phi) against the CPU oracle's energy and its
    central differences (the chain rule through rx(2a), xx(theta*phi)^2,
    inverted user gates, a controlled phase, a negated control and U)."""
    from snuqs_amd.simulator import StateVectorSimulator

    x0 = {"theta": 0.37, "phi": -1.1}
    res = StateVectorSimulator().run({"source": ADJ, "inputs": x0}, precision=precision)
    v = res.resultTypes[0]["value"]
    tol = 1e-9 if precision == "complex128" else 1e-5
    assert abs(v["expectation"] - _oracle_energy(ADJ, x0)) < tol
    h = 1e-5
    for name in ("theta", "phi"):
        up = _oracle_energy(ADJ, {**x0, name: x0[name] + h})
        dn = _oracle_energy(ADJ, {**x0, name: x0[name] - h})
        fd = (up - dn) / (2 * h)
        assert abs(v["gradient"][name] - fd) < max(tol, 1e-8), (name, v, fd)
    # 'all' names every input; a gradient with shots is refused
    src_all = ADJ.replace(") theta, phi", ") all")
    v2 = StateVectorSimulator().run({"source": src_all, "inputs": x0},
                                    precision=precision).resultTypes[0]["value"]
    assert sorted(v2["gradient"]) == ["phi", "theta"]
    assert abs(v2["gradient"]["phi"] - v["gradient"]["phi"]) < 1e-12
    with pytest.raises(ValueError):
        StateVectorSimulator().run({"source": ADJ, "inputs": x0}, shots=10)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_hermitian_observables(precision):
    """Expectation and Variance of Hermitian-matrix observables (JAQCD matrix
    form and OpenQASM hermitian(...)), alone, on two qubits and in a tensor
    product with Pauli factors, against <psi|O|psi> and <O^2> - <O>^2 from
    the oracle's state; with shots, the diagonalizing gate runs before
    sampling and the sample mean of the eigenvalues tracks <O>."""
    from snuqs_amd.simulator import StateVectorSimulator

    rng = np.random.default_rng(2)
    h2 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    h2 = (h2 + h2.conj().T) / 2
    h4 = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    h4 = (h4 + h4.conj().T) / 2
    jm = lambda m: [[[z.real, z.imag] for z in row] for row in m]  # noqa: E731
    ins = [{"type": "h", "target": 0}, {"type": "ry", "target": 1, "angle": 0.7},
           {"type": "cnot", "control": 0, "target": 2}, {"type": "rx", "target": 2, "angle": -0.4},
           {"type": "xy", "targets": [1, 2], "angle": 0.9}]
    results = [{"type": "expectation", "observable": [jm(h2)], "targets": [1]},
               {"type": "variance", "observable": [jm(h4)], "targets": [2, 0]},
               {"type": "expectation", "observable": ["x", jm(h2)], "targets": [0, 2]},
               {"type": "expectation", "observable": [jm(h2)]}]
    res = StateVectorSimulator().run({"instructions": ins, "results": results},
                                     precision=precision)
    from snuqs_amd.gate_operations import from_braket_instruction
    psi = run_oracle(3, [from_braket_instruction(i) for i in ins])
    X = np.array([[0, 1], [1, 0]])

    def emb(m, qs):
        from test_openqasm import _embed
        return _embed(m, qs, 3)

    tol = 1e-10 if precision == "complex128" else 2e-5
    vals = [r["value"] for r in res.resultTypes]
    ev = lambda op: np.vdot(psi, op @ psi).real  # noqa: E731
    assert abs(vals[0] - ev(emb(h2, [1]))) < tol
    o = emb(h4, [2, 0])
    assert abs(vals[1] - (ev(o @ o) - ev(o) ** 2)) < 4 * tol
    assert abs(vals[2] - ev(emb(X, [0]) @ emb(h2, [2]))) < tol
    assert np.abs(np.array(vals[3]) - [ev(emb(h2, [q])) for q in range(3)]).max() < tol
    # OpenQASM form of the tensor product
    qasm = ("OPENQASM 3.0;\nqubit[3] q;\nh q[0];\nry(0.7) q[1];\ncnot q[0], q[2];\n"
            "rx(-0.4) q[2];\nxy(0.9) q[1], q[2];\n"
            "#pragma braket result expectation x(q[0]) @ hermitian(" +
            str([[complex(z) for z in row] for row in h2]).replace("j", "im").replace("(", "")
            .replace(")", "") + ") q[2]\n")
    v = StateVectorSimulator().run(qasm, precision=precision).resultTypes[0]["value"]
    assert abs(v - vals[2]) < tol
    # shots (OpenQASM derives the rotations from the result pragmas): the
    # eigenvalue sample mean after the diagonalizing rotation tracks <O>
    lam = np.linalg.eigvalsh(h2)
    head, pragma = qasm.rstrip("\n").rsplit("\n", 1)
    qasm1 = head + "\n" + pragma.replace("x(q[0]) @ ", "").replace(") q[2]", ") q[1]") + "\n"
    shot_res = StateVectorSimulator().run(qasm1, shots=20000, precision=precision)
    bits = np.array([[int(ch) for ch in m] for m in shot_res.measurements])
    mean = lam[bits[:, 1]].mean()
    assert abs(mean - vals[0]) < 5 * np.ptp(lam) / np.sqrt(20000)
