"""The drop-in boundary on the GPU: the C ABI called through ctypes (the
binding a maintainer would add to the reference), error reporting through
sq_last_error, stream semantics of the pybind layer (the reference's apply,
initialize_* and copies take `stream=`, pymodule.cc:143-151,167-171), and the
StateVector constructors and transfers of pymodule.cc:154-177."""
import ctypes
import os

import numpy as np
import pytest

from helpers import device_state, run_oracle, to_np
from oracle import gates as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SQ_OK, SQ_ERR_INVALID = 0, 1
SQ_COMPLEX64, SQ_COMPLEX128 = 0, 1


@pytest.fixture(scope="module")
def lib(C):
    L = ctypes.CDLL(os.path.join(ROOT, "snuqs_amd", "libsnuqs_hip.so"))
    P, U32, U64, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    L.sq_malloc.argtypes = [ctypes.POINTER(P), SZ]
    L.sq_free.argtypes = [P]
    L.sq_initialize_basis_z.argtypes = [P, ctypes.c_int, U64, P]
    L.sq_initialize_zero.argtypes = [P, ctypes.c_int, U64, P]
    L.sq_apply.argtypes = [P, ctypes.c_int, U64, U32, P, U32, P, P]
    L.sq_memcpy_d2h.argtypes = [P, P, SZ, P]
    L.sq_stream_synchronize.argtypes = [P]
    L.sq_program_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, U32, U32, P, P, P]
    L.sq_program_run.argtypes = [P, P, U64, P]
    L.sq_program_destroy.argtypes = [P]
    L.sq_norm2.argtypes = [P, ctypes.c_int, U64, ctypes.POINTER(ctypes.c_double), P]
    L.sq_last_error.restype = ctypes.c_char_p
    return L


def _readback(lib, sv, n, dtype=np.complex128):
    out = np.empty(1 << n, dtype=dtype)
    assert lib.sq_memcpy_d2h(out.ctypes.data, sv, out.nbytes, None) == SQ_OK
    return out


def test_ctypes_apply_and_program(lib):
    """INTEGRATION.md's ctypes binding: sq_apply gate by gate and the same
    gates as one sq_program, both against the oracle."""
    n = 12
    rng = np.random.default_rng(2)
    gates = [(G.hadamard(), [0]), (G.cx(), [0, 5]), (G.rot_y(0.7), [11]),
             (G.ccnot(), [3, 11, 1]), (G.xx(1.3), [2, 9]), (G.u(0.1, 0.2, 0.3), [7])]
    z = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    gates.append((np.linalg.qr(z)[0], [4, 6, 8, 10]))  # 4 targets: device-staged matrix
    ref = O.run_circuit(n, [(m, t) for m, t in gates])
    sv = ctypes.c_void_p()
    assert lib.sq_malloc(ctypes.byref(sv), 16 << n) == SQ_OK
    try:
        assert lib.sq_initialize_basis_z(sv, SQ_COMPLEX128, 1 << n, None) == SQ_OK
        for m, t in gates:
            mm = np.ascontiguousarray(m, dtype=np.complex128)
            tt = np.array(t, dtype=np.uint32)
            rc = lib.sq_apply(sv, SQ_COMPLEX128, 1 << n, n, tt.ctypes.data, len(t),
                              mm.ctypes.data, None)
            assert rc == SQ_OK, lib.sq_last_error()
        assert lib.sq_stream_synchronize(None) == SQ_OK
        assert np.abs(_readback(lib, sv, n) - ref).max() < 1e-12

        nt = np.array([len(t) for _, t in gates], dtype=np.uint32)
        tg = np.array([q for _, t in gates for q in t], dtype=np.uint32)
        mats = np.concatenate([np.ascontiguousarray(m, dtype=np.complex128).ravel()
                               for m, _ in gates])
        prog = ctypes.c_void_p()
        rc = lib.sq_program_create(ctypes.byref(prog), SQ_COMPLEX128, n, len(gates),
                                   nt.ctypes.data, tg.ctypes.data, mats.ctypes.data)
        assert rc == SQ_OK, lib.sq_last_error()
        assert lib.sq_initialize_basis_z(sv, SQ_COMPLEX128, 1 << n, None) == SQ_OK
        assert lib.sq_program_run(prog, sv, 1 << n, None) == SQ_OK
        assert lib.sq_stream_synchronize(None) == SQ_OK
        assert np.abs(_readback(lib, sv, n) - ref).max() < 1e-12
        # a program for another state size is refused
        assert lib.sq_program_run(prog, sv, 1 << (n - 1), None) == SQ_ERR_INVALID
        assert b"qubits" in lib.sq_last_error()
        assert lib.sq_program_destroy(prog) == SQ_OK
    finally:
        assert lib.sq_free(sv) == SQ_OK


@pytest.mark.parametrize("targets,what", [([3, 3], b"duplicate"), ([0, 12], b"range")])
def test_ctypes_bad_targets_report_an_error(lib, targets, what):
    """Bad targets return SQ_ERR_INVALID with a message; the process and the
    state are untouched (the reference exits, utils_cuda.h:8-17)."""
    n = 12
    sv = ctypes.c_void_p()
    assert lib.sq_malloc(ctypes.byref(sv), 16 << n) == SQ_OK
    try:
        assert lib.sq_initialize_basis_z(sv, SQ_COMPLEX128, 1 << n, None) == SQ_OK
        m = np.ascontiguousarray(G.cx(), dtype=np.complex128)
        tt = np.array(targets, dtype=np.uint32)
        rc = lib.sq_apply(sv, SQ_COMPLEX128, 1 << n, n, tt.ctypes.data, 2, m.ctypes.data, None)
        assert rc == SQ_ERR_INVALID
        assert what in lib.sq_last_error().lower(), lib.sq_last_error()
        # a 3-target matrix with 2 targets' worth of entries is not readable
        # through this ABI; a bad dtype and a non-power-of-two size are
        rc = lib.sq_apply(sv, 7, 1 << n, n, tt.ctypes.data, 1, m.ctypes.data, None)
        assert rc == SQ_ERR_INVALID
        rc = lib.sq_apply(sv, SQ_COMPLEX128, (1 << n) - 1, n, tt.ctypes.data, 1,
                          m.ctypes.data, None)
        assert rc == SQ_ERR_INVALID
        assert lib.sq_stream_synchronize(None) == SQ_OK
        st = _readback(lib, sv, n)
        assert st[0] == 1 and np.count_nonzero(st) == 1
    finally:
        assert lib.sq_free(sv) == SQ_OK


def test_ctypes_initialize_zero_and_norm(lib):
    n = 16
    sv = ctypes.c_void_p()
    assert lib.sq_malloc(ctypes.byref(sv), 8 << n) == SQ_OK
    try:
        nrm = ctypes.c_double()
        assert lib.sq_initialize_basis_z(sv, SQ_COMPLEX64, 1 << n, None) == SQ_OK
        assert lib.sq_norm2(sv, SQ_COMPLEX64, 1 << n, ctypes.byref(nrm), None) == SQ_OK
        assert nrm.value == 1.0
        assert lib.sq_initialize_zero(sv, SQ_COMPLEX64, 1 << n, None) == SQ_OK
        assert lib.sq_norm2(sv, SQ_COMPLEX64, 1 << n, ctypes.byref(nrm), None) == SQ_OK
        assert nrm.value == 0.0
        assert not _readback(lib, sv, n, np.complex64).any()
    finally:
        assert lib.sq_free(sv) == SQ_OK


@pytest.mark.parametrize("nonblocking", [False, True])
def test_streams_ordered_by_events(C, nonblocking):
    """Gates queued on two streams, ordered by an event: stream 1 prepares
    the state, stream 2 waits on the event and continues; the result is the
    whole sequence in order.  Every call honours its stream (the reference
    ignores it, functionals.cc:15-28)."""
    from snuqs_amd.circuits import random_gate_list

    make = C.stream.Stream.create_nonblocking if nonblocking else C.stream.Stream.create
    s1, s2 = make(), make()
    n = 20
    ops = random_gate_list(n, 120, np.random.default_rng(7))
    half = len(ops) // 2
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n)
    C.functionals.initialize_zero(sv, stream=s1)
    C.functionals.initialize_basis_z(sv, stream=s1)
    for op in ops[:half]:
        C.functionals.apply(sv, op, n, op.targets, stream=s1)
    ev = C.event.Event.create()
    ev.record(s1)
    s2.wait_event(ev)
    prog = C.functionals.Program(n, ops[half:])
    prog.run(sv, stream=s2)
    for op in ops[:4]:  # and back on the per-gate path, still on stream 2
        C.functionals.apply(sv, op, n, op.targets, stream=s2)
    host = sv.cpu(stream=s2)  # the copy waits for stream 2's work
    ref = run_oracle(n, ops + ops[:4])
    assert np.abs(np.asarray(host) - ref).max() < 1e-12
    # timing events on a stream
    e0, e1 = C.event.Event.create(), C.event.Event.create()
    e0.record(s1)
    prog.run(sv, stream=s1)
    e1.record(s1)
    e1.synchronize()
    assert e0.elapsed_ms(e1) >= 0.0
    s1.synchronize()
    s2.synchronize()


def test_pinned_host_state_round_trip(C):
    """StateVector(n, pinned) (pymodule.cc:154-177): the buffer protocol
    writes a pinned host state, .cuda() / .hip() copy it to the device, the
    gates run there, .cpu() copies back; .sync() and copy() between tiers."""
    from snuqs_amd.circuits import random_gate_list

    n = 14
    rng = np.random.default_rng(9)
    init = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    init /= np.linalg.norm(init)
    host = C.result_types.StateVector(n, True)
    assert "CPU" in repr(host) and host.num_elems() == 1 << n
    np.asarray(host)[:] = init
    dev = host.cuda()
    dev2 = host.hip()
    assert dev.device == C.DeviceType.CUDA and dev2.device == C.DeviceType.CUDA
    assert dev.cuda() is dev  # already on the device
    ops = random_gate_list(n, 60, rng)
    for op in ops:
        C.functionals.apply(dev, op, n, op.targets)
    dev.sync()
    got = np.asarray(dev.cpu())
    assert np.abs(got - run_oracle(n, ops, init)).max() < 1e-12
    assert np.array_equal(np.asarray(dev2.cpu()), init)  # an independent copy
    back = C.result_types.StateVector(n, True)
    back.copy(dev)                 # device -> pinned host
    assert np.array_equal(np.asarray(back), got)
    dev2.copy(back)                # pinned host -> device
    assert np.array_equal(to_np(C, dev2), got)
    with pytest.raises(ValueError):
        C.functionals.apply(host, ops[0], n, ops[0].targets)  # no CPU path
    with pytest.raises(ValueError):
        dev.copy(C.result_types.StateVector(n - 1, True))


def test_slices_share_the_parent_buffer(C):
    """slice(n_sliced, idx) views part of the parent (pbuffer.cc:78-81):
    gates on a slice change only that part of the parent state."""
    n, k = 12, 8
    sv = device_state(C, n)
    g = C.operation.gate_operations
    for q in range(n):
        C.functionals.apply(sv, g.Hadamard([q]), n, [q])
    sl = sv.slice(k, 3)  # amplitudes 3*256 .. 4*256-1
    C.functionals.apply(sl, g.PauliZ([k - 1]), k, [k - 1])
    got = to_np(C, sv)
    want = np.full(1 << n, 2 ** (-n / 2))
    want[3 * 256 + 1:4 * 256:2] *= -1
    assert np.abs(got - want).max() < 1e-15
