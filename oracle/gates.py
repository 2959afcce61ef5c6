"""Gate matrices of the SnuQS reference, restated in numpy.

TEST INFRASTRUCTURE ONLY (the oracle): the product builds its own matrices in
snuqs_amd/csrc/host/gates.cc; tests compare the two and pin both against
closed-form identities and the Braket decompositions in
src/braket/snuqs/openqasm/braket_gates.inc:5-123.

Each entry restates the constructor in
src/braket/snuqs/csrc/operation/gate_operations.cc (line ranges below): a
row-major complex128 2^k x 2^k matrix whose local index has targets[0] as the
MSB.
"""
import cmath
import math

import numpy as np

S2 = math.sqrt(0.5)  # M_SQRT1_2 (gate_operations.cc:26-29)


def _m(rows):
    return np.array(rows, dtype=np.complex128)


def identity():  # gate_operations.cc:9-19
    return _m([[1, 0], [0, 1]])


def hadamard():  # :21-31
    return _m([[S2, S2], [S2, -S2]])


def pauli_x():  # :33-43
    return _m([[0, 1], [1, 0]])


def pauli_y():  # :45-55
    return _m([[0, -1j], [1j, 0]])


def pauli_z():  # :57-67
    return _m([[1, 0], [0, -1]])


def cx():  # :69-91
    return _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])


def cy():  # :93-115
    return _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])


def cz():  # :117-139
    return _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]])


def s():  # :141-149
    return _m([[1, 0], [0, 1j]])


def si():  # :151-159
    return _m([[1, 0], [0, -1j]])


def t():  # :160-169
    return _m([[1, 0], [0, cmath.exp(1j * math.pi / 4.0)]])


def ti():  # :171-180
    return _m([[1, 0], [0, cmath.exp(-1j * math.pi / 4.0)]])


def v():  # :182-191
    return _m([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])


def vi():  # :193-202
    return _m([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])


def phase_shift(a):  # :204-215
    return _m([[1, 0], [0, cmath.exp(1j * a)]])


def _diag4(d):
    return np.diag(np.array(d, dtype=np.complex128))


def cphase_shift(a):  # :217-241
    return _diag4([1, 1, 1, cmath.exp(1j * a)])


def cphase_shift00(a):  # :243-267
    return _diag4([cmath.exp(1j * a), 1, 1, 1])


def cphase_shift01(a):  # :269-293
    return _diag4([1, cmath.exp(1j * a), 1, 1])


def cphase_shift10(a):  # :295-315
    return _diag4([1, 1, cmath.exp(1j * a), 1])


def rot_x(a):  # :317-330
    c, s_ = math.cos(a / 2), 1j * math.sin(a / 2)
    return _m([[c, -s_], [-s_, c]])


def rot_y(a):  # :332-345
    c, s_ = math.cos(a / 2), math.sin(a / 2)
    return _m([[c, -s_], [s_, c]])


def rot_z(a):  # :347-360
    return _m([[cmath.exp(-1j * a / 2.0), 0], [0, cmath.exp(1j * a / 2.0)]])


def swap():  # :362-384
    return _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


def iswap():  # :386-408
    return _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])


def pswap(a):  # :410-433
    e = cmath.exp(1j * a)
    return _m([[1, 0, 0, 0], [0, 0, e, 0], [0, e, 0, 0], [0, 0, 0, 1]])


def xy(a):  # :435-460
    c, s_ = math.cos(a / 2), 1j * math.sin(a / 2)
    return _m([[1, 0, 0, 0], [0, c, s_, 0], [0, s_, c, 0], [0, 0, 0, 1]])


def xx(a):  # :462-487
    c, s_ = math.cos(a / 2), 1j * math.sin(a / 2)
    return _m([[c, 0, 0, -s_], [0, c, -s_, 0], [0, -s_, c, 0], [-s_, 0, 0, c]])


def yy(a):  # :489-514
    c, s_ = math.cos(a / 2), 1j * math.sin(a / 2)
    return _m([[c, 0, 0, s_], [0, c, -s_, 0], [0, -s_, c, 0], [s_, 0, 0, c]])


def zz(a):  # :516-541
    p, m = cmath.exp(1j * a / 2.0), cmath.exp(-1j * a / 2.0)
    return _diag4([m, p, p, m])


def ccnot():  # :543-614
    g = np.eye(8, dtype=np.complex128)
    g[[6, 7]] = g[[7, 6]]
    return g


def cswap():  # :616-685
    g = np.eye(8, dtype=np.complex128)
    g[[5, 6]] = g[[6, 5]]
    return g


def u(theta, phi, lam):  # :687-700
    c, s_ = math.cos(theta / 2), math.sin(theta / 2)
    return _m([[c, -cmath.exp(1j * lam) * s_],
               [cmath.exp(1j * phi) * s_, cmath.exp(1j * (phi + lam)) * c]])


def gphase(a):  # :702-710 (zero targets: a 1x1 matrix)
    return _m([[cmath.exp(1j * a)]])


# name -> (constructor, number of targets, number of angles)
GATES = {
    "Identity": (identity, 1, 0), "Hadamard": (hadamard, 1, 0),
    "PauliX": (pauli_x, 1, 0), "PauliY": (pauli_y, 1, 0), "PauliZ": (pauli_z, 1, 0),
    "CX": (cx, 2, 0), "CY": (cy, 2, 0), "CZ": (cz, 2, 0),
    "S": (s, 1, 0), "Si": (si, 1, 0), "T": (t, 1, 0), "Ti": (ti, 1, 0),
    "V": (v, 1, 0), "Vi": (vi, 1, 0),
    "PhaseShift": (phase_shift, 1, 1), "CPhaseShift": (cphase_shift, 2, 1),
    "CPhaseShift00": (cphase_shift00, 2, 1), "CPhaseShift01": (cphase_shift01, 2, 1),
    "CPhaseShift10": (cphase_shift10, 2, 1),
    "RotX": (rot_x, 1, 1), "RotY": (rot_y, 1, 1), "RotZ": (rot_z, 1, 1),
    "Swap": (swap, 2, 0), "ISwap": (iswap, 2, 0), "PSwap": (pswap, 2, 1),
    "XY": (xy, 2, 1), "XX": (xx, 2, 1), "YY": (yy, 2, 1), "ZZ": (zz, 2, 1),
    "CCNot": (ccnot, 3, 0), "CSwap": (cswap, 3, 0),
    "U": (u, 1, 3), "GPhase": (gphase, 0, 1),
}
