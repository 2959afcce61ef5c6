"""sq_expect_pauli_sum_noalloc (the expectation entry point that allocates
nothing: its scratch is a fixed per-device array) against sq_expect_pauli_sum,
both called through the C ABI, on two calls in a row, the second on a smaller
state with fewer terms, in both precisions; and against the oracle."""
import ctypes
import os

import numpy as np
import pytest

from helpers import device_state, random_state
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SQ_OK, SQ_ERR_INVALID = 0, 1
DT = {"complex64": 0, "complex128": 1}


@pytest.fixture(scope="module")
def lib(C):
    L = ctypes.CDLL(os.path.join(ROOT, "snuqs_amd", "libsnuqs_hip.so"))
    P, U32, U64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    D = ctypes.POINTER(ctypes.c_double)
    sig = [P, ctypes.c_int, U64, U32, P, P, P, P, D, P, P]
    L.sq_expect_pauli_sum.argtypes = sig
    L.sq_expect_pauli_sum_noalloc.argtypes = sig
    L.sq_last_error.restype = ctypes.c_char_p
    return L


def _terms(n, count, rng):
    """Random Pauli strings of weight 1-4: (x, z, ny, coeffs, labels)."""
    x = np.zeros(count, np.uint64)
    z = np.zeros(count, np.uint64)
    ny = np.zeros(count, np.uint32)
    labels = []
    for t in range(count):
        qs = rng.choice(n, int(rng.integers(1, 5)), replace=False)
        ps = rng.choice(list("XYZ"), qs.size)
        labels.append(("".join(ps), [int(q) for q in qs]))
        for p, q in zip(ps, qs):
            bit = np.uint64(1 << (n - 1 - int(q)))
            if p in "XY":
                x[t] |= bit
            if p in "YZ":
                z[t] |= bit
            ny[t] += p == "Y"
    return x, z, ny, rng.standard_normal(count), labels


def _call(lib, fn, sv, dt, n, x, z, ny, c):
    out = ctypes.c_double()
    vals = np.empty(max(1, x.size))
    rc = getattr(lib, fn)(sv.data_ptr, DT[dt], 1 << n, x.size, x.ctypes.data, z.ctypes.data,
                          ny.ctypes.data, c.ctypes.data, ctypes.byref(out), vals.ctypes.data,
                          None)
    assert rc == SQ_OK, lib.sq_last_error()
    return out.value, vals[:x.size].copy()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_noalloc_expectation_matches_the_workspace_entry_point(C, lib, dtype):
    rng = np.random.default_rng(11)
    # first call: 18 qubits and 200 terms (four launches of up to 64 terms);
    # second call: a smaller state and fewer terms on the same array
    for n, count in ((18, 200), (9, 7)):
        psi = random_state(n, rng)
        sv = device_state(C, n, dtype, init=psi.astype(dtype))
        x, z, ny, c, labels = _terms(n, count, rng)
        e_ws, v_ws = _call(lib, "sq_expect_pauli_sum", sv, dtype, n, x, z, ny, c)
        e_na, v_na = _call(lib, "sq_expect_pauli_sum_noalloc", sv, dtype, n, x, z, ny, c)
        # the same kernels, launch shapes and summation order: equal bits
        assert e_na == e_ws and np.array_equal(v_na, v_ws), (n, count)
        ref_in = psi.astype(dtype).astype(np.complex128)
        ref = np.array([O.expect_pauli(ref_in, p, q, n).real for p, q in labels])
        tol = 1e-12 if dtype == "complex128" else 1e-5
        assert np.abs(v_na - ref).max() <= tol, (n, count)
        assert abs(e_na - float(np.dot(c, ref))) <= tol * np.abs(c).sum()


def test_noalloc_expectation_limits(C, lib):
    n = 6
    sv = device_state(C, n)
    x, z, ny, c, _ = _terms(n, 257, np.random.default_rng(3))
    out = ctypes.c_double(7.0)
    rc = lib.sq_expect_pauli_sum_noalloc(sv.data_ptr, 1, 1 << n, 257, x.ctypes.data,
                                         z.ctypes.data, ny.ctypes.data, c.ctypes.data,
                                         ctypes.byref(out), None, None)
    assert rc == SQ_ERR_INVALID and b"256" in lib.sq_last_error()
    # no terms: zero, no launch
    rc = lib.sq_expect_pauli_sum_noalloc(sv.data_ptr, 1, 1 << n, 0, None, None, None, None,
                                         ctypes.byref(out), None, None)
    assert rc == SQ_OK and out.value == 0.0
    # the largest allowed call still works after the refusal (|0..0>: Z terms give +1)
    zz = np.full(256, 1 << (n - 1), np.uint64)
    zero = np.zeros(256, np.uint64)
    ny0 = np.zeros(256, np.uint32)
    ones = np.ones(256)
    rc = lib.sq_expect_pauli_sum_noalloc(sv.data_ptr, 1, 1 << n, 256, zero.ctypes.data,
                                         zz.ctypes.data, ny0.ctypes.data, ones.ctypes.data,
                                         ctypes.byref(out), None, None)
    assert rc == SQ_OK, lib.sq_last_error()
    assert out.value == 256.0
