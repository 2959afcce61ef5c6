"""The host-memory tier on the GPU (sq_host_tier_pass): a state in pinned host
memory streamed through device slices, against the CPU oracle and against
the same circuit run on a device-resident state."""
import numpy as np
import pytest

from helpers import device_state, to_np
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
@pytest.mark.parametrize("n,m,k,buffers", [(14, 10, 6, 3), (13, 9, 1, 2), (12, 12, 12, 1)])
def test_host_tier_vs_oracle(dtype, tol, n, m, k, buffers):
    from snuqs_amd.circuits import random_gate_list
    from snuqs_amd.host_tier import HostStateSimulation

    rng = np.random.default_rng(n + m + k)
    ops = random_gate_list(n, 150, rng, with_unitary=True)
    sim = HostStateSimulation(n, m, dtype=dtype, contiguous_bits=k, buffers=buffers)
    sim.evolve(ops)
    ref = O.run_circuit(n, [(op.matrix(), op.targets) for op in ops])
    got = sim.state_vector.astype(np.complex128)
    assert np.abs(got - ref).max() <= tol
    if m < n:
        assert sim.passes_run > 1


def test_host_tier_buffer_rings_agree_and_evolve_twice():
    """1, 2 and 3 device buffers give the same bits; a second evolve continues
    from the state the first left (no zero-slice shortcut)."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.host_tier import HostStateSimulation

    n, m = 16, 12
    ops = random_layered(n, 6, seed=3)
    out = []
    for nb in (1, 2, 3):
        sim = HostStateSimulation(n, m, dtype="complex128", contiguous_bits=7, buffers=nb)
        sim.evolve(ops[:40])
        sim.evolve(ops[40:])
        out.append(sim.state_vector.copy())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    ref = O.run_circuit(n, [(op.matrix(), op.targets) for op in ops])
    assert np.abs(out[0] - ref).max() < 1e-12


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_host_tier_vs_device_state_26q(C, dtype, tol):
    """26 qubits through 20-qubit slices (64 slices, 64 chunks of 2^14
    amplitudes each way per slice) against the device-resident engine."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.host_tier import HostStateSimulation
    from snuqs_amd.transpile import fuse_operations

    n, m = 26, 20
    ops = random_layered(n, 8, seed=26)
    sim = HostStateSimulation(n, m, dtype=dtype, contiguous_bits=14, buffers=3)
    sim.evolve(ops)
    sv = device_state(C, n, dtype)
    C.functionals.Program(n, fuse_operations(ops, 5), dtype=dtype).run(sv)
    ref = to_np(C, sv)
    got = sim.state_vector.astype(np.complex128)
    print(f"host tier {n}q/{m}q {dtype}: {sim.passes_run} passes, max-abs vs device "
          f"{np.abs(got - ref).max():.2e}")
    assert np.abs(got - ref).max() <= tol
    assert sim.passes_run > 1


def test_host_tier_rejects_bad_arguments(C):
    from snuqs_amd._C.operation import gate_operations as g

    n, m = 10, 6
    host = C.result_types.StateVector(n, True, dtype="complex128")
    C.functionals.host_initialize_basis_z(host)
    prog = C.functionals.Program(m, [g.Hadamard([0])], dtype="complex128")
    buf = C.result_types.StateVector(C.DeviceType.CUDA, m, dtype="complex128")
    with pytest.raises(Exception):  # must start with host bit 0
        prog.run_host_pass(host, [1, 2, 3, 4, 5, 6], [buf])
    with pytest.raises(Exception):  # repeated bit
        prog.run_host_pass(host, [0, 1, 2, 3, 4, 4], [buf])
    with pytest.raises(ValueError):  # wrong number of local bits
        prog.run_host_pass(host, [0, 1, 2, 3, 4], [buf])
    with pytest.raises(ValueError):  # buffer of another size
        prog.run_host_pass(host, [0, 1, 2, 3, 4, 5],
                           [C.result_types.StateVector(C.DeviceType.CUDA, m + 1)])
    with pytest.raises(ValueError):  # device state instead of a host state
        prog.run_host_pass(buf, [0, 1, 2, 3, 4, 5], [buf])
    prog.run_host_pass(host, [0, 1, 2, 3, 4, 5], [buf])  # valid: H on device qubit 0 = bit 5
    a = np.asarray(host)
    assert abs(a[0] - 2 ** -0.5) < 1e-15 and abs(a[1 << 5] - 2 ** -0.5) < 1e-15


@pytest.mark.parametrize("mode", [{"offload": "cpu"}, {"accelerator": "hybrid"}])
def test_simulator_host_tier_results(mode):
    """StateVectorSimulator.run with the state in host memory (the
    reference's hybrid / cpu-offload options, qubit_count_cuda = slice size)
    gives the device path's state vector, probabilities and shots."""
    from snuqs_amd.simulator import StateVectorSimulator

    n = 12
    rng = np.random.default_rng(7)
    ins = []
    for _ in range(60):
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ins.append({"type": "h", "target": a})
        ins.append({"type": "cnot", "control": a, "target": b})
        ins.append({"type": "rz", "target": b, "angle": float(rng.uniform(0, 6.28))})
    prog = {"instructions": ins, "results": [{"type": "statevector"},
                                             {"type": "probability", "targets": [3, 0, 7]}]}
    sim = StateVectorSimulator()
    ref = sim.run(prog)
    got = sim.run(prog, qubit_count_cuda=8, qubit_count_slice=5, **mode)
    assert np.abs(got.resultTypes[0]["value"] - ref.resultTypes[0]["value"]).max() < 1e-12
    assert np.abs(got.resultTypes[1]["value"] - ref.resultTypes[1]["value"]).max() < 1e-12
    shots_prog = {"instructions": ins}
    np.random.seed(11)
    a = sim.run(shots_prog, shots=300, qubit_count_cuda=8, qubit_count_slice=5, **mode)
    np.random.seed(11)
    b = sim.run(shots_prog, shots=300)
    assert a.measurements == b.measurements
