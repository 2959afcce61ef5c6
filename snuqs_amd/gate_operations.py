"""Instruction -> gate-object translation (src/braket/snuqs/gate_operations.py
and operation_helpers.py of the reference).

``BRAKET_GATES`` maps Braket gate names to the engine's gate classes
(gate_operations.py:202-239).  ``from_braket_instruction`` turns a JAQCD
instruction -- a ``braket.ir.jaqcd`` object when the Braket SDK is installed,
or the same instruction as a JSON dict -- into a gate object; controls come
first in the target list (CX([control, target]), CCNot([c0, c1, t]),
CSwap([c, t0, t1])).  Unlike the reference, ``unitary`` instructions work.
"""
from __future__ import annotations

import numpy as np

from ._C.operation import gate_operations as g

BRAKET_GATES = {
    "i": g.Identity, "h": g.Hadamard, "x": g.PauliX, "y": g.PauliY, "z": g.PauliZ,
    "cnot": g.CX, "cy": g.CY, "cz": g.CZ, "s": g.S, "si": g.Si, "t": g.T, "ti": g.Ti,
    "v": g.V, "vi": g.Vi, "phaseshift": g.PhaseShift, "cphaseshift": g.CPhaseShift,
    "cphaseshift00": g.CPhaseShift00, "cphaseshift01": g.CPhaseShift01,
    "cphaseshift10": g.CPhaseShift10, "rx": g.RotX, "ry": g.RotY, "rz": g.RotZ,
    "swap": g.Swap, "iswap": g.ISwap, "pswap": g.PSwap, "xy": g.XY, "xx": g.XX,
    "yy": g.YY, "zz": g.ZZ, "ccnot": g.CCNot, "cswap": g.CSwap, "U": g.U,
    "gphase": g.GPhase, "unitary": g.Unitary,
}


def _get(ins, name, default=None):
    if isinstance(ins, dict):
        return ins.get(name, default)
    return getattr(ins, name, default)


def _type(ins) -> str:
    t = _get(ins, "type")
    t = getattr(t, "value", t)
    if t is None:
        t = type(ins).__name__
    return str(t).lower()


def ir_matrix_to_ndarray(matrix) -> np.ndarray:
    """[[[re, im], ...], ...] -> complex ndarray (operation_helpers.py:94-104)."""
    return np.array([[complex(e[0], e[1]) for e in row] for row in matrix])


def from_braket_instruction(ins):
    """JAQCD instruction (object or dict) -> gate object."""
    t = _type(ins)
    ctor = BRAKET_GATES.get(t)
    if ctor is None:
        raise NotImplementedError(f"Instruction {ins} not recognized")
    control = _get(ins, "control")
    controls = _get(ins, "controls")
    target = _get(ins, "target")
    targets = _get(ins, "targets")
    qubits = []
    if controls:
        qubits += [int(q) for q in controls]
    if control is not None:
        qubits.append(int(control))
    if target is not None:
        qubits.append(int(target))
    if targets:
        qubits += [int(q) for q in targets]
    if t == "unitary":
        return g.Unitary(qubits, ir_matrix_to_ndarray(_get(ins, "matrix")))
    angle = _get(ins, "angle")
    if angle is not None:
        return ctor(qubits, float(angle))
    return ctor(qubits)
