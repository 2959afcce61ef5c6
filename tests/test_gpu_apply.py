"""Gate application on the GPU vs the CPU oracle (cpu::applyGate restatement).

Bars: complex128 max-abs <= 1e-12; complex64 max-abs <= 1e-5 against the
complex128 oracle run on the same (complex64-rounded) input.
"""
import struct

import numpy as np
import pytest

from helpers import device_state, random_state, run_engine, run_oracle, to_np
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {"complex128": 1e-12, "complex64": 1e-5}


def _kinds(C):
    from snuqs_amd.circuits import GATE_KINDS

    return GATE_KINDS


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n", [5, 11, 16])
def test_every_gate_kind_every_position(C, dtype, n):
    """Each of the 33 reference gate kinds on several target sets (lowest,
    highest and random bit positions, both target orders).  Targets or
    controls on the lowest bits take the LDS-tiled kernels at n = 11 and 16;
    permutations stay exact moves there (checked bitwise below)."""
    from snuqs_amd.circuits import make_gate

    rng = np.random.default_rng(100 + n)
    init = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    g = C.operation.gate_operations
    perms = (g.CX, g.Swap, g.CCNot, g.CSwap, g.PauliX)
    worst = 0.0
    for kind in _kinds(C):
        k = kind[1]
        choices = [list(range(k)), list(range(n - k, n))[::-1]] + [
            list(rng.permutation(n)[:k]) for _ in range(3)]
        for t in choices:
            op = make_gate(kind, t, rng.random(3) * 2 * np.pi)
            got = run_engine(C, n, [op], dtype, init)
            ref = run_oracle(n, [op], init.astype(np.complex128))
            err = np.abs(got - ref).max()
            worst = max(worst, err)
            assert err <= TOL[dtype], (op, t, err)
            if kind[0] in perms:
                # pure permutations: every amplitude moves unchanged
                assert np.array_equal(got, ref.astype(got.dtype)), (op, t)
    print(f"n={n} {dtype}: worst {worst:.2e}")


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("n", [7, 12])
def test_unitary_k_targets(C, dtype, k, n):
    """Dense Unitary on k targets: k<=3 VALU kernel, k=4,5 MFMA (n=12) or the
    LDS kernel (small states)."""
    from snuqs_amd.circuits import haar_unitary

    g = C.operation.gate_operations
    rng = np.random.default_rng(7 * k + n)
    init = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    for trial in range(4):
        if trial == 0:
            t = list(range(k))  # highest strides
        elif trial == 1:
            t = list(range(n - k, n))  # lowest strides, including bit 0
        else:
            t = [int(x) for x in rng.permutation(n)[:k]]
        op = g.Unitary(t, haar_unitary(1 << k, rng))
        got = run_engine(C, n, [op], dtype, init)
        ref = run_oracle(n, [op], init.astype(np.complex128))
        err = np.abs(got - ref).max()
        assert err <= TOL[dtype], (t, err)


# (grid cap, prefetch, tile order 0 cyclic / 1 blocked / 2 scattered, wide)
_NARROW = [(256, 1, 0, 0), (1, 1, 0, 0), (1, 0, 0, 0), (1, 1, 1, 0), (3, 0, 1, 0), (1024, 1, 1, 0),
           (256, 1, 2, 0), (1, 1, 2, 0), (3, 0, 2, 0)]
_WIDE = [(256, 1, 0, 1), (1, 1, 0, 1), (1, 0, 1, 1), (1024, 0, 1, 1), (1, 1, 2, 1),
         (256, 0, 2, 1)]  # complex64 only


@pytest.mark.parametrize("dtype,grid,pipe,order,wide",
                         [(d,) + c for d in ("complex128", "complex64") for c in _NARROW]
                         + [("complex64",) + c for c in _WIDE])
def test_unitary_5_target_tiles(C, dtype, grid, pipe, order, wide):
    """5-target MFMA kernels (LDS-staged tiles) on target sets that put the
    tile's bits everywhere: the lowest five address bits, the highest, spread
    and given out of order, plus 4-target and controlled 4-target blocks; a
    1-block-per-CU grid cap makes every wave loop over several tiles with and
    without the prefetch.  wide = 64-group complex64 tiles (22 qubits, so the
    1-block cap still loops)."""
    from snuqs_amd.circuits import haar_unitary

    g = C.operation.gate_operations
    n = 22 if wide else 21
    rng = np.random.default_rng(55)
    init = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    ref0 = init.astype(np.complex128)
    suffix = "c64" if dtype == "complex64" else "c128"
    keys = (f"mfma_blocks_per_cu_{suffix}", f"mfma5_pipeline_{suffix}",
            f"mfma5_order_{suffix}", "mfma5_wide_c64")
    saved = [C.core.tune_get(k) for k in keys]
    try:
        C.core.tune_set(keys[0], grid)
        C.core.tune_set(keys[1], pipe)
        C.core.tune_set(keys[2], order)
        C.core.tune_set(keys[3], wide)
        top = n - 1
        ops = [g.Unitary(t, haar_unitary(32, rng))
               for t in ([16, 17, 18, 19, top], [0, 1, 2, 3, 4], [top, 0, 13, 6, 17],
                         [18, 11, 5, 19, 2], [top, 18, 16, 14, 12])]
        # 4-target blocks (widened to 5 targets with an identity factor on the
        # lowest free bit) and controlled 4-target blocks: control on the
        # lowest bit (not widened) and on a middle bit (widened)
        ops += [g.Unitary(t, haar_unitary(16, rng))
                for t in ([17, 18, 19, top], [0, 7, 19, 3], [1, 2, 3, top])]
        for t in ([top, 0, 5, 11, 16], [9, top, 19, 2, 14]):
            m = np.eye(32, dtype=complex)
            m[16:, 16:] = haar_unitary(16, rng)  # targets[0] is the control
            ops.append(g.Unitary(t, m))
        for op in ops:
            t = list(op.targets)
            got = run_engine(C, n, [op], dtype, init)
            ref = run_oracle(n, [op], ref0)
            err = np.abs(got - ref).max()
            assert err <= TOL[dtype], (t, err)
    finally:
        for k, v in zip(keys, saved):
            C.core.tune_set(k, v)


def test_wide_tiles_are_bitwise_narrow_tiles(C):
    """The 64-group complex64 tile multiplies the same operands in the same
    MFMA order as the 32-group tile: outputs are bitwise equal."""
    from snuqs_amd.circuits import haar_unitary

    g = C.operation.gate_operations
    n = 22
    rng = np.random.default_rng(8)
    init = random_state(n, rng, np.complex64)
    saved = C.core.tune_get("mfma5_wide_c64")
    try:
        for t in ([21, 3, 9, 15, 0], [5, 6, 7, 8, 9], [20, 17, 12, 1, 10]):
            op = g.Unitary(t, haar_unitary(32, rng))
            out = []
            for wide in (0, 1):
                C.core.tune_set("mfma5_wide_c64", wide)
                out.append(run_engine(C, n, [op], "complex64", init))
            assert np.array_equal(out[0], out[1]), t
    finally:
        C.core.tune_set("mfma5_wide_c64", saved)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_tile_orders_are_bitwise_equal(C, dtype):
    """Cyclic, blocked and scattered (bit-reversed) tile orders only change
    which wave takes which tile: outputs are bitwise equal."""
    from snuqs_amd.circuits import haar_unitary

    g = C.operation.gate_operations
    n = 22
    rng = np.random.default_rng(81)
    init = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    key = "mfma5_order_" + ("c64" if dtype == "complex64" else "c128")
    saved = C.core.tune_get(key)
    try:
        for t in ([21, 3, 9, 15, 0], [5, 6, 7, 8, 9], [20, 17, 12, 1, 10]):
            op = g.Unitary(t, haar_unitary(32, rng))
            out = []
            for order in (0, 1, 2):
                C.core.tune_set(key, order)
                out.append(run_engine(C, n, [op], dtype, init))
            assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2]), t
        with pytest.raises(ValueError):
            C.core.tune_set(key, 3)
    finally:
        C.core.tune_set(key, saved)


def test_ghz20_closed_form_bitwise(C):
    """GHZ-20 complex128: amplitudes 0 and 2^20-1 are bitwise M_SQRT1_2
    (0x3FE6A09E667F3BCD), every other amplitude is exactly zero."""
    from snuqs_amd.circuits import ghz

    n = 20
    got = run_engine(C, n, ghz(n), "complex128")
    bits = struct.unpack("<Q", struct.pack("<d", got[0].real))[0]
    assert bits == 0x3FE6A09E667F3BCD
    assert got[-1] == got[0] and got[0].imag == 0.0
    rest = got[1:-1]
    assert np.count_nonzero(rest) == 0
    p = C.functionals.probabilities(device_state(C, 1))  # smoke of the API on a tiny state
    assert p.shape == (2,)


def test_ghz20_matches_reference_cpu_path(C):
    """GHZ-20 against the reference's own cpu::applyGate (oracle/_ref), bitwise."""
    from snuqs_amd.circuits import ghz

    if O.ref_lib() is None:
        pytest.skip("reference CPU build absent")
    n = 20
    ops = ghz(n)
    st = O.basis_state(n)
    for op in ops:
        O.ref_apply(st, np.array(op.matrix()), op.targets, n)
    got = run_engine(C, n, ops, "complex128")
    assert np.array_equal(got, st)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_random_circuit(C, dtype):
    from snuqs_amd.circuits import random_gate_list

    n = 12
    rng = np.random.default_rng(2024)
    ops = random_gate_list(n, 400, rng, with_unitary=True)
    got = run_engine(C, n, ops, dtype)
    ref = run_oracle(n, ops)
    assert np.abs(got - ref).max() <= TOL[dtype]


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_ctrl_and_power_modifiers(C, dtype):
    g = C.operation.gate_operations
    n = 6
    rng = np.random.default_rng(3)
    init = random_state(n, rng)
    ops = [
        g.PauliX([0, 3], ctrl_modifiers=[1]),
        g.PauliX([1, 2, 4], ctrl_modifiers=[0, 1]),
        g.RotY([5, 0], 0.7, ctrl_modifiers=[0]),
        g.Hadamard([2], power=3),
        g.T([4], power=-1),
        g.U([1, 3], 0.3, 0.2, 0.1, ctrl_modifiers=[1]),
        g.GPhase([0, 1], 0.4),
        g.GPhase([], 0.9),
    ]
    init_d = init.astype(np.complex64 if dtype == "complex64" else np.complex128)
    got = run_engine(C, n, ops, dtype, init_d)
    ref = run_oracle(n, ops, init_d.astype(np.complex128))
    assert np.abs(got - ref).max() <= TOL[dtype]


def test_slice_with_larger_num_qubits(C):
    """As in the reference, num_qubits may exceed log2(num_elems): index math
    uses num_qubits, the sweep covers the slice (functionals.cc:15-28)."""
    g = C.operation.gate_operations
    n, m = 12, 9  # 2^9-amplitude slice of a 12-qubit index space
    rng = np.random.default_rng(11)
    full = C.result_types.StateVector(C.DeviceType.CUDA, m, dtype="complex128")
    init = random_state(m, rng)
    host = C.result_types.StateVector(m)
    np.array(host, copy=False)[:] = init
    full.copy(host)
    ops = [g.Hadamard([5]), g.CX([3, 11]), g.RotZ([8], 0.3), g.CCNot([4, 6, 10])]
    ref = init.copy()
    for op in ops:
        C.functionals.apply(full, op, n, op.targets)
        O.apply(ref, np.array(op.matrix()), op.targets, n)
    assert np.abs(to_np(C, full) - ref).max() <= 1e-12
    with pytest.raises(ValueError):
        C.functionals.apply(full, g.Hadamard([2]), n, [2])  # stride beyond the slice


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_large_slices_take_the_tiled_kernels(C, dtype):
    """2^16-amplitude slices of a 20-qubit state, gates given with
    num_qubits = 20 (targets t >= 4 fall inside the slice): the 5-target MFMA
    kernel, the LDS-tiled low-bit kernels and the permutation path run on the
    slice's own amplitudes and leave the other slices untouched."""
    from snuqs_amd.circuits import haar_unitary

    g = C.operation.gate_operations
    n, m = 20, 16
    rng = np.random.default_rng(12)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    init = random_state(n, rng, npdt)
    full = device_state(C, n, dtype, init)
    ops = [g.Unitary([19, 5, 12, 8, 17], haar_unitary(32, rng)),
           g.Unitary([18, 19], haar_unitary(4, rng)), g.Hadamard([19]),
           g.CX([19, 7]), g.CCNot([18, 19, 6]), g.Unitary([16, 17, 18], haar_unitary(8, rng))]
    for idx in (1, 10):
        sl = full.slice(m, idx)
        for op in ops:
            C.functionals.apply(sl, op, n, op.targets)
    ref = init.astype(np.complex128).reshape(1 << (n - m), 1 << m).copy()
    for idx in (1, 10):
        part = ref[idx].copy()
        for op in ops:
            O.apply(part, np.array(op.matrix()), [t - (n - m) for t in op.targets], m)
        ref[idx] = part
    got = to_np(C, full).reshape(1 << (n - m), 1 << m)
    assert np.abs(got - ref).max() <= TOL[dtype]


def test_invalid_targets_raise(C):
    g = C.operation.gate_operations
    sv = device_state(C, 4)
    with pytest.raises(ValueError):
        C.functionals.apply(sv, g.CX([0, 0]), 4, [0, 0])
    with pytest.raises(ValueError):
        C.functionals.apply(sv, g.Hadamard([4]), 4, [4])
    with pytest.raises(ValueError):
        C.functionals.apply(sv, g.CX([0, 1]), 4, [0])
    host = C.result_types.StateVector(4)
    with pytest.raises(Exception):
        C.functionals.apply(host, g.Hadamard([0]), 4, [0])  # no CPU path


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_program_matches_per_gate_apply(C, dtype):
    from snuqs_amd.circuits import random_gate_list

    n = 13
    rng = np.random.default_rng(8)
    ops = random_gate_list(n, 300, rng, with_unitary=True)
    prog = C.functionals.Program(n, ops, dtype=dtype)
    sv = device_state(C, n, dtype)
    prog.run(sv)
    got = to_np(C, sv)
    ref = run_oracle(n, ops)
    assert np.abs(got - ref).max() <= TOL[dtype]
    st = prog.stats()
    assert st["launches"] <= len(ops)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_fused_layered_circuit(C, dtype):
    """The benchmark circuit family at 16 qubits: fused 5-qubit blocks on the
    matrix cores vs the unfused oracle run."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    n = 16
    ops = random_layered(n, 8, seed=1234)
    fused = fuse_operations(ops, 5)
    prog = C.functionals.Program(n, fused, dtype=dtype)
    sv = device_state(C, n, dtype)
    prog.run(sv)
    got = to_np(C, sv)
    ref = run_oracle(n, ops)
    assert np.abs(got - ref).max() <= TOL[dtype]
    classes = {prog.gate_info(i)[0] for i in range(len(prog))}
    assert 5 in classes  # MFMA kernel used
    ms = prog.run_timed(sv)
    assert ms.shape == (len(fused),) and (ms >= 0).all()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_program_tune_keeps_results_bitwise(C, dtype):
    """Program.tune picks a launch shape per 5-target gate; the arithmetic is
    the same in every shape, so the tuned program's output is bitwise equal to
    the untuned one's."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    n = 21
    fused = fuse_operations(random_layered(n, 6, seed=99), 5)
    prog = C.functionals.Program(n, fused, dtype=dtype)
    sv = device_state(C, n, dtype)
    prog.run(sv)
    before = np.asarray(sv.cpu()).copy()
    scratch = device_state(C, n, dtype)
    tuned = prog.tune(scratch, reps=1)
    assert tuned == sum(1 for i in range(len(prog)) if prog.gate_info(i)[:2] == (5, 5))
    assert tuned > 0
    assert prog.tune(scratch, reps=1) == tuned  # second call: cached shapes
    sv2 = device_state(C, n, dtype)
    prog.run(sv2)
    assert np.array_equal(np.asarray(sv2.cpu()), before)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_program_graph_replays_match_eager(C, dtype):
    """Program.run(use_graph=True) captures the launch sequence once per state
    and replays it: every replay (with eager work and syncs between) is
    bitwise equal to the eager run; another state or a re-tune recaptures."""
    from snuqs_amd.circuits import random_gate_list, random_layered
    from snuqs_amd.transpile import fuse_operations

    n = 18
    rng = np.random.default_rng(77)
    ops = fuse_operations(random_layered(n, 3, seed=5), 5) + random_gate_list(n, 12, rng)
    prog = C.functionals.Program(n, ops, dtype=dtype)
    eager = device_state(C, n, dtype)
    prog.run(eager)
    want = np.asarray(eager.cpu()).copy()
    sv = device_state(C, n, dtype)
    other = device_state(C, n, dtype)
    for rep in range(4):
        C.functionals.initialize_basis_z(sv)  # eager kernel between replays
        prog.run(sv, use_graph=True)
        C.core.cuda.device_synchronize()
        assert np.array_equal(np.asarray(sv.cpu()), want), rep
        prog.run(other)  # eager launches of the same program in between
        assert C.functionals.norm2(other) > 0
    C.functionals.initialize_basis_z(other)
    prog.run(other, use_graph=True)  # another state: recaptured
    assert np.array_equal(np.asarray(other.cpu()), want)
    prog.tune(other, reps=1)  # drops the graph
    C.functionals.initialize_basis_z(sv)
    prog.run(sv, use_graph=True)
    assert np.array_equal(np.asarray(sv.cpu()), want)


def test_copy_kernel(C):
    rng = np.random.default_rng(0)
    n = 15
    v = random_state(n, rng)
    a = device_state(C, n, "complex128", v)
    b = device_state(C, n, "complex128")
    C.functionals.copy_kernel(b, a)
    assert np.array_equal(to_np(C, b), v)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_fused_permutation_blocks(C, dtype):
    """Blocks that are pure permutations on 4-5 qubits (fused CX / Swap / CCNot)."""
    g = C.operation.gate_operations
    from snuqs_amd.transpile import fuse_operations

    n = 12
    rng = np.random.default_rng(4)
    init = random_state(n, rng, np.complex64 if dtype == "complex64" else np.complex128)
    ops = [g.CX([0, 1]), g.CX([2, 3]), g.Swap([1, 2]), g.CCNot([4, 5, 6]), g.CX([6, 7]),
           g.CSwap([8, 9, 10]), g.CX([10, 11])]
    fused = fuse_operations(ops, 5)
    assert max(len(op.targets) for op in fused) >= 4
    got = run_engine(C, n, fused, dtype, init)
    ref = run_oracle(n, ops, init.astype(np.complex128))
    assert np.abs(got - ref).max() <= TOL[dtype]


@pytest.mark.parametrize("dtype,n", [("complex64", 26), ("complex128", 24)])
def test_program_tiles_per_wave_bitwise(C, dtype, n):
    """Untuned programs large enough for the 4-5 target kernels to walk 4
    tiles per wave (mfma_tiles_per_wave 4) give the same bits as 1 tile per
    wave (narrow and, for complex64, wide tiles)."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    ops = fuse_operations(random_layered(n, 4, seed=n), 5)
    assert any(len(op.targets) == 5 for op in ops)
    keys = ("mfma_tiles_per_wave", "mfma5_wide_c64")
    saved = [C.core.tune_get(k) for k in keys]
    out = {}
    try:
        for wide in ((0, 1) if dtype == "complex64" else (0,)):
            C.core.tune_set(keys[1], wide)
            for tiles in (1, 4):
                C.core.tune_set(keys[0], tiles)
                sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
                C.functionals.initialize_basis_z(sv)
                C.functionals.Program(n, ops, dtype=dtype).run(sv)
                out[(wide, tiles)] = np.array(sv.cpu(), copy=True)
                del sv
    finally:
        for k, v in zip(keys, saved):
            C.core.tune_set(k, v)
    for wide in ((0, 1) if dtype == "complex64" else (0,)):
        assert np.array_equal(out[(wide, 1)], out[(wide, 4)]), wide
