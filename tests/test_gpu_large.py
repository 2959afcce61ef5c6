"""Full-size configurations, checked through size-independent properties and,
where the CPU oracle finishes in reasonable time, directly.

* config 5: 34-qubit complex64 and 33-qubit complex128 states (128 GiB each)
  of 20 layers of Haar 1-qubit U on every qubit + CX on random disjoint pairs,
  fused into 5-qubit blocks: the mirror circuit (C then C^dagger) returns
  |0...0>, the norm stays 1, and the queried bitstring probabilities are
  finite and sum below 1.  The same generator at 20 qubits is compared with
  the oracle amplitude by amplitude.
* config 2 at its full 30 qubits, complex128: the first layer against the CPU
  restatement over all 2^30 amplitudes.
"""
import gc

import numpy as np
import pytest

from helpers import device_state, run_oracle, to_np

pytestmark = pytest.mark.gpu


def _inverse(ops):
    from snuqs_amd._C.operation import gate_operations as g

    return [g.Unitary(list(op.targets), np.array(op.matrix()).conj().T) if op.targets
            else g.GPhase([], -float(np.angle(np.array(op.matrix())[0, 0])))
            for op in reversed(ops)]


def _free_gb(C):
    return C.core.cuda.mem_info()[0] / 2**30


@pytest.mark.parametrize("n,dtype,tol", [(34, "complex64", 2e-4), (33, "complex128", 1e-10)])
def test_config5_mirror_full_size(C, n, dtype, tol):
    from snuqs_amd.circuits import random_u_cx
    from snuqs_amd.transpile import fuse_operations

    need = (1 << n) * (8 if dtype == "complex64" else 16) / 2**30
    if _free_gb(C) < need + 4:
        pytest.skip(f"needs {need:.0f} GiB of free device memory")
    ops = random_u_cx(n, 20, seed=n)
    fused = fuse_operations(ops, 5)
    prog = C.functionals.Program(n, fused, dtype=dtype)
    inv = C.functionals.Program(n, _inverse(fused), dtype=dtype)
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
    C.functionals.initialize_basis_z(sv)
    prog.run(sv)
    q = np.random.default_rng(n).integers(0, 1 << n, 10**4, dtype=np.uint64)
    amps = C.functionals.amplitudes(sv, [int(x) for x in q])
    probs = np.abs(amps) ** 2
    assert np.isfinite(probs).all() and probs.sum() < 1.0
    assert abs(C.functionals.norm2(sv) - 1.0) < tol
    inv.run(sv)
    a0 = C.functionals.amplitudes(sv, [0])[0]
    assert abs(a0 - 1.0) < tol, a0
    del sv, prog, inv
    gc.collect()


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_config5_generator_vs_oracle_20q(C, dtype, tol):
    from snuqs_amd.circuits import random_u_cx
    from snuqs_amd.transpile import fuse_operations

    n = 20
    ops = random_u_cx(n, 20, seed=34)
    fused = fuse_operations(ops, 5)
    prog = C.functionals.Program(n, fused, dtype=dtype)
    sv = device_state(C, n, dtype)
    prog.run(sv)
    ref = run_oracle(n, ops)
    q = [int(x) for x in np.random.default_rng(1).integers(0, 1 << n, 10**4)]
    got = C.functionals.amplitudes(sv, q)
    assert np.abs(got - ref[q]).max() < tol
    assert np.abs(to_np(C, sv) - ref).max() < tol


def test_config2_full_30q_first_layer_vs_cpu_restatement(C):
    from snuqs_amd.circuits import random_layered

    n = 30
    if _free_gb(C) < 20:
        pytest.skip("needs 20 GiB of free device memory")
    ops = random_layered(n, 1, seed=1234)  # 15 Unitary + 30 RotZ
    sv = device_state(C, n, "complex128")
    C.functionals.Program(n, ops, dtype="complex128").run(sv)
    got = to_np(C, sv)
    del sv
    ref = run_oracle(n, ops)
    err = float(np.abs(got - ref).max())
    del got, ref
    gc.collect()
    assert err <= 1e-12, err


def test_config2_30q_c64_million_samples_exact(C):
    """Config 2 at 30 qubits complex64 (fused): 1e6 samples from the exact GPU
    sampler are identical to numpy's RandomState.choice on the same
    probabilities and seed; fast mode's mismatch count is reported."""
    import time

    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    n = 30
    if _free_gb(C) < 40:
        pytest.skip("needs 40 GiB of free device memory")
    sv = device_state(C, n, "complex64")
    C.functionals.Program(n, fuse_operations(random_layered(n, 30, seed=1234), 5),
                          dtype="complex64").run(sv)
    assert abs(C.functionals.norm2(sv) - 1) < 1e-5  # FP32 drift of the norm
    p = C.functionals.probabilities(sv, normalize=True)
    u = np.random.RandomState(1234).random_sample(10**6)
    t0 = time.perf_counter()
    got = C.functionals.sample(sv, u, exact=True, normalize=True)
    t_exact = time.perf_counter() - t0
    t0 = time.perf_counter()
    fast = C.functionals.sample(sv, u, exact=False, normalize=True)
    t_fast = time.perf_counter() - t0
    # numpy's choice rejects probabilities whose sum is off by more than
    # sqrt(eps); the device normalises by cumsum(p)[-1] exactly as choice does
    assert abs(p.sum() - 1) < 1e-12, p.sum()
    want = np.random.RandomState(1234).choice(p.size, p=p, size=10**6)
    print(f"exact {t_exact:.2f} s, fast {t_fast:.2f} s, fast-mode mismatches "
          f"{int(np.count_nonzero(fast != want))} of 1e6")
    assert np.array_equal(got, want)
    del sv
    gc.collect()
