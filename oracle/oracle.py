"""CPU oracle for the SnuQS state-vector path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product (snuqs_amd) never does.

Layers (each cites the reference behaviour it restates):

* ``apply`` / ``apply_c64`` -- ctypes calls into oracle/libsv_oracle.so, the C
  restatement of cpu::applyGate
  (src/braket/snuqs/csrc/operation/gate_operations_impl_cpu.cc:10-148).
* ``ref_apply`` -- the reference's own cpu::applyGate compiled from its sources
  into oracle/_ref/libref_cpu.so (complex128, k <= 3); ``None`` when absent.
* ``tensordot_apply`` -- an independent numpy formulation restating
  src/braket/snuqs/linalg_utils.py:47-74 (_multiply_matrix: reshape the state to
  [2]*n, tensordot the gate on the target axes, transpose back).
* ``kron_apply`` -- brute force with the full 2^n x 2^n Kronecker matrix.
* ``pseudo_sort_operations_descending`` -- src/braket/snuqs/transpile.py:33-89.
* ``probabilities`` / ``choice`` -- src/braket/snuqs/simulation.py:208-220.
* ``expect_pauli`` / ``inner`` -- textbook formulas (no reference counterpart).
* ``run_circuit`` -- init |0..0> (initialization_impl_cpu.cc:15-22) then the
  per-gate loop of simulation.py:274-278.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = False  # False = not yet tried; None = unavailable


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsv_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        P, S = ctypes.c_void_p, ctypes.c_size_t
        for name in ("sqo_apply_c128", "sqo_apply_c64"):
            getattr(lib, name).argtypes = [P, P, P, S, S, S]
            getattr(lib, name).restype = None
        lib.sqo_probabilities_c128.argtypes = [P, P, S]
        lib.sqo_cumsum.argtypes = [P, P, S]
        lib.sqo_choice.argtypes = [P, S, P, S, P, P]
        lib.sqo_expect_pauli_c128.argtypes = [P, S, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]
        lib.sqo_inner_c128.argtypes = [P, P, S, P]
        _LIB = lib
    return _LIB


def ref_lib():
    """The reference's own CPU gate path (oracle/_ref/libref_cpu.so) or None."""
    global _REF
    if _REF is False:
        path = os.path.join(_HERE, "_ref", "libref_cpu.so")
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            P, S = ctypes.c_void_p, ctypes.c_size_t
            lib.ref_apply_gate.argtypes = [P, P, P, S, S, S]
            lib.ref_initialize_zero.argtypes = [P, S]
            lib.ref_initialize_basis_z.argtypes = [P, S]
            _REF = lib
        else:
            _REF = None
    return _REF


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def apply(sv: np.ndarray, gate: np.ndarray, targets, nqubits: int) -> None:
    """In-place cpu::applyGate restatement; dtype of ``sv`` picks c128 or c64."""
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.uintp))
    k = len(targets)
    assert sv.flags.c_contiguous and sv.ndim == 1
    assert gate.shape == (1 << k, 1 << k)
    if sv.dtype == np.complex128:
        g = np.ascontiguousarray(gate, dtype=np.complex128)
        _lib().sqo_apply_c128(_ptr(sv), _ptr(g), _ptr(t), k, nqubits, sv.size)
    elif sv.dtype == np.complex64:
        g = np.ascontiguousarray(gate, dtype=np.complex64)
        _lib().sqo_apply_c64(_ptr(sv), _ptr(g), _ptr(t), k, nqubits, sv.size)
    else:
        raise TypeError(sv.dtype)


def ref_apply(sv: np.ndarray, gate: np.ndarray, targets, nqubits: int) -> None:
    lib = ref_lib()
    if lib is None:
        raise RuntimeError("reference CPU build absent")
    assert sv.dtype == np.complex128 and len(targets) <= 3
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.uintp))
    g = np.ascontiguousarray(gate, dtype=np.complex128)
    lib.ref_apply_gate(_ptr(sv), _ptr(g), _ptr(t), len(targets), nqubits, sv.size)


def basis_state(n: int, dtype=np.complex128) -> np.ndarray:
    sv = np.zeros(1 << n, dtype=dtype)
    sv[0] = 1
    return sv


def run_circuit(n: int, gates, dtype=np.complex128, state=None) -> np.ndarray:
    """gates: iterable of (matrix, targets)."""
    sv = basis_state(n, dtype) if state is None else state
    for m, t in gates:
        apply(sv, m, t, n)
    return sv


def tensordot_apply(sv: np.ndarray, gate: np.ndarray, targets, nqubits: int) -> np.ndarray:
    """linalg_utils._multiply_matrix formulation (returns a new array)."""
    k = len(targets)
    if k == 0:
        return sv * gate[0, 0]
    state = sv.reshape([2] * nqubits)
    g = gate.reshape([2] * (2 * k))
    prod = np.tensordot(g, state, axes=(np.arange(k, 2 * k), list(targets)))
    unused = [i for i in range(nqubits) if i not in targets]
    inv = np.argsort(list(targets) + unused)
    return np.transpose(prod, inv).reshape(-1)


def kron_apply(sv: np.ndarray, gate: np.ndarray, targets, nqubits: int) -> np.ndarray:
    """Brute force: build the full operator column by column (n <= 10)."""
    N = 1 << nqubits
    k = len(targets)
    full = np.zeros((N, N), dtype=np.complex128)
    for col in range(N):
        bits = [(col >> (nqubits - 1 - q)) & 1 for q in range(nqubits)]
        c_loc = 0
        for j, q in enumerate(targets):
            c_loc |= bits[q] << (k - 1 - j)
        for r_loc in range(1 << k):
            row_bits = list(bits)
            for j, q in enumerate(targets):
                row_bits[q] = (r_loc >> (k - 1 - j)) & 1
            row = 0
            for q in range(nqubits):
                row |= row_bits[q] << (nqubits - 1 - q)
            full[row, col] += gate[r_loc, c_loc]
    return full @ sv


def probabilities(sv: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(sv, dtype=np.complex128)
    p = np.empty(s.size, dtype=np.float64)
    _lib().sqo_probabilities_c128(_ptr(s), _ptr(p), s.size)
    return p


def cumsum(p: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    c = np.empty_like(p)
    _lib().sqo_cumsum(_ptr(p), _ptr(c), p.size)
    return c


def choice(p: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    scratch = np.empty_like(p)
    out = np.empty(u.size, dtype=np.int64)
    _lib().sqo_choice(_ptr(p), p.size, _ptr(u), u.size, _ptr(scratch), _ptr(out))
    return out


def pauli_masks(pauli: str, qubits, nqubits: int):
    """(xmask, zmask, ny) in amplitude-index bit space for a Pauli string."""
    x = z = ny = 0
    for ch, q in zip(pauli, qubits):
        b = 1 << (nqubits - 1 - q)
        if ch in "XY":
            x |= b
        if ch in "ZY":
            z |= b
        if ch == "Y":
            ny += 1
    return x, z, ny


def expect_pauli(sv: np.ndarray, pauli: str, qubits, nqubits: int) -> complex:
    s = np.ascontiguousarray(sv, dtype=np.complex128)
    x, z, ny = pauli_masks(pauli, qubits, nqubits)
    out = np.zeros(2)
    _lib().sqo_expect_pauli_c128(_ptr(s), s.size, x, z, ny, _ptr(out))
    return complex(out[0], out[1])


def inner(a: np.ndarray, b: np.ndarray) -> complex:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    out = np.zeros(2)
    _lib().sqo_inner_c128(_ptr(a), _ptr(b), a.size, _ptr(out))
    return complex(out[0], out[1])


# --- transpile.py:33-89 -----------------------------------------------------

def _pseudo_gt(op1, pos1, op2, pos2, sliceable=lambda op: False):
    t1, t2 = op1.targets, op2.targets
    if len(t1) == 0 and len(t2) == 0:
        return pos1 < pos2
    if len(t1) == 0:
        return True
    if len(t2) == 0:
        return False
    if set(t1) & set(t2):
        return pos1 < pos2
    if sliceable(op1) and sliceable(op2):
        return pos1 < pos2
    if sliceable(op1):
        return True
    if sliceable(op2):
        return False
    return min(t1) > min(t2)


def pseudo_sort_operations_descending(operations):
    """Insertion sort of transpile.py:74-89; pseudo_gt sees (op, i) and (op, j), i and j being indices into the list being sorted."""
    ops = list(operations)
    for i in range(1, len(operations)):
        op1 = ops[i]
        j = i - 1
        while j >= 0:
            op2 = ops[j]
            if _pseudo_gt(op1, i, op2, j):
                ops[j + 1] = ops[j]
            else:
                break
            j -= 1
        if i != j + 1:
            ops[j + 1] = op1
    return ops
