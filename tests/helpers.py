"""Shared helpers for the tests (run the engine, run the oracle)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def to_np(C, sv):
    """Device state -> complex128 numpy array (host copy)."""
    a = np.asarray(sv.cpu())
    return a if a.dtype == np.complex128 else a.astype(np.complex128)


def device_state(C, n, dtype="complex128", init=None):
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
    if init is None:
        C.functionals.initialize_basis_z(sv)
    else:
        host = C.result_types.StateVector(n, dtype=dtype)
        np.array(host, copy=False)[:] = init
        sv.copy(host)
    return sv


def random_state(n, rng, dtype=np.complex128):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def run_engine(C, n, ops, dtype="complex128", init=None):
    sv = device_state(C, n, dtype, init)
    for op in ops:
        C.functionals.apply(sv, op, n, op.targets)
    return to_np(C, sv)


def run_oracle(n, ops, init=None):
    st = O.basis_state(n) if init is None else np.array(init, dtype=np.complex128)
    for op in ops:
        O.apply(st, np.array(op.matrix()), op.targets, n)
    return st
