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
    assert all("".join(m) in ("0000", "1111") for m in res.measurements)
    with pytest.raises(TypeError):
        sim.run(ghz_program(2), accelerator="cpu")
    with pytest.raises(TypeError):
        sim.run(ghz_program(2), offload="storage")
