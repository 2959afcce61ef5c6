"""Braket local-simulator entry point (src/braket/snuqs/simulator.py of the reference).

``StateVectorSimulator`` keeps the reference's device id ("snuqs"), its
``run(ir, *, shots, qubit_count_cpu, qubit_count_cuda, qubit_count_slice,
accelerator, prefetch, offload, path, block_count)`` signature
(simulator.py:152-155, 450-461) and its steps: contiguous-qubit remap
(:278-307), transpile, simulation.evolve, result assembly (:309-381).  The
Braket SDK and schemas are optional: without them ``run`` takes a JAQCD
program as a JSON dict (or string), or an OpenQASM 3 program (an object or dict
with ``source`` and ``inputs``, or the source text; read by
``snuqs_amd.openqasm``, run_openqasm at simulator.py:380-437), and returns a
``TaskResult`` with the same fields as GateModelTaskResult.  Extra keyword arguments: ``precision``
("complex64" / "complex128"), ``fuse`` (largest fused block, 0 = off) and
``exact_sampling``.

Differences from the reference, by design: the engine runs on the GPU
(accelerator None / "cuda" / "hip"); accelerator "hybrid" or offload "cpu"
keeps the state in pinned host memory and streams it through device slices
of ``qubit_count_cuda`` qubits (``HostStateVectorSimulation``, for states
larger than HBM; StateVector, Amplitude, Probability and shots only);
"cpu" and the storage tier raise.  Besides StateVector the result types Amplitude, Probability
(marginals over up to 20 targets reduced on the GPU), Expectation and Variance
(Pauli products with h and i factors, and Hermitian matrices through their
diagonalizing gates and eigenvalues; an observable without targets gives one
value per qubit) are computed on the GPU, and OpenQASM programs may ask
for Braket's AdjointGradient (expectation of a Pauli sum and its gradient
with respect to the program's inputs, from the engine's adjoint sweep).
"""
from __future__ import annotations

import json
import math
import uuid
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _C, openqasm
from .gate_operations import BRAKET_GATES, from_braket_instruction
from .simulation import HostStateVectorSimulation, StateVectorSimulation
from .subcircuit import Subcircuit
from .transpile import pseudo_sort_operations_descending
from .types import AcceleratorType, OffloadType, PrefetchType


@dataclass
class TaskResult:
    """GateModelTaskResult's fields, for use without the Braket schemas."""
    measurements: list
    measuredQubits: list
    resultTypes: list
    taskMetadata: dict = field(default_factory=dict)
    additionalMetadata: dict = field(default_factory=dict)


def _get(x, name, default=None):
    if isinstance(x, dict):
        return x.get(name, default)
    return getattr(x, name, default)


class BaseSimulator:
    DEVICE_ID = "snuqs"

    @staticmethod
    def _validate_simulation_type(accelerator: Optional[str]) -> AcceleratorType:
        supported = {None: AcceleratorType.CUDA, "cuda": AcceleratorType.CUDA,
                     "hip": AcceleratorType.HIP, "gpu": AcceleratorType.CUDA,
                     "hybrid": AcceleratorType.HYBRID}
        if accelerator == "cpu":
            raise TypeError(f"Accelerator {accelerator} is not supported: this engine runs "
                            "on the GPU (accelerator='cuda', 'hip' or 'hybrid')")
        if accelerator not in supported:
            raise TypeError(f"Accelerator {accelerator} is not supported")
        return supported[accelerator]

    @staticmethod
    def _validate_offload_type(offload: Optional[str]) -> OffloadType:
        if offload in (None, "none"):
            return OffloadType.NONE
        if offload == "cpu":
            return OffloadType.CPU
        if offload == "storage":
            raise TypeError("Offload storage is not supported: the state lives in one GPU's "
                            "HBM or, with offload='cpu', in pinned host memory")
        raise TypeError(f"Offload {offload} is not supported")

    @staticmethod
    def _validate_prefetch_type(prefetch: Optional[str]) -> PrefetchType:
        table = {None: PrefetchType.NONE, "none": PrefetchType.NONE,
                 "cuda": PrefetchType.CUDA, "storage": PrefetchType.STORAGE}
        if prefetch not in table:
            raise TypeError(f"Prefetch {prefetch} is not supported")
        return table[prefetch]

    @staticmethod
    def _validate_qubit_count(qubit_count: int, precision: str) -> None:
        free, _ = _C.core.cuda.mem_info()
        need = (1 << qubit_count) * (8 if precision == "complex64" else 16)
        if need > free:
            raise MemoryError(f"{qubit_count} qubits in {precision} need {need} bytes; "
                              f"{free} bytes of device memory are free")

    @staticmethod
    def _contiguous_qubit_mapping(qubits) -> dict:
        return {q: i for i, q in enumerate(sorted(qubits))}

    @staticmethod
    def _formatted_measurements(samples: np.ndarray, qubit_count: int,
                                measured_qubits=None) -> list:
        """Sample index -> list of bit characters, qubit q = character q
        (simulator.py:309-343)."""
        bits = ((samples[:, None] >> (qubit_count - 1 - np.arange(qubit_count))) & 1)
        if measured_qubits:
            mq = np.array(measured_qubits)
            inside = mq[mq < qubit_count]
            bits = bits[:, inside]
            pad = int(np.count_nonzero(mq >= qubit_count))
            if pad:
                bits = np.pad(bits, ((0, 0), (0, pad)))
        return [[str(b) for b in row] for row in bits.tolist()]

    @staticmethod
    def _is_openqasm(ir) -> bool:
        if isinstance(ir, bytes):
            ir = ir.decode()
        if isinstance(ir, str):
            return not ir.lstrip().startswith("{")
        if isinstance(ir, dict):
            return "source" in ir
        return _get(ir, "source") is not None

    @staticmethod
    def _parse_program(ir):
        if isinstance(ir, (str, bytes)):
            ir = json.loads(ir)
        if isinstance(ir, dict):
            return ir.get("instructions", []), ir.get("results") or [], \
                ir.get("basis_rotation_instructions") or []
        return (_get(ir, "instructions", []), _get(ir, "results") or [],
                _get(ir, "basis_rotation_instructions") or [])

    def run(self, ir, *args, **kwargs):
        """Dispatch on the program kind (simulator.py:152-155)."""
        if self._is_openqasm(ir):
            return self.run_openqasm(ir, *args, **kwargs)
        return self.run_jaqcd(ir, *args, **kwargs)

    def run_jaqcd(self, ir, *, shots: int = 0, **kwargs):
        instructions, results, basis = self._parse_program(ir)
        ops = [from_braket_instruction(i) for i in instructions]
        if shots > 0 and basis:
            ops += [from_braket_instruction(i) for i in basis]
        used = {q for op in ops for q in op.targets}
        for r in results:
            for q in (_get(r, "targets") or []):
                used.add(int(q))
        # with shots the reference returns no result types for JAQCD
        # (simulator.py:497-510)
        return self._simulate(ir, ops, results, used, shots, None, shots_results=[], **kwargs)

    def run_openqasm(self, ir, *, shots: int = 0, inputs: Optional[dict] = None, **kwargs):
        """OpenQASM 3 program (simulator.py:380-437): gates, result pragmas and
        measured qubits from ``snuqs_amd.openqasm``; with shots, the
        observables' basis rotations run before sampling."""
        if isinstance(ir, bytes):
            ir = ir.decode()
        source = ir if isinstance(ir, str) else _get(ir, "source")
        given = {} if isinstance(ir, str) else dict(_get(ir, "inputs") or {})
        given.update(inputs or {})
        circuit = openqasm.parse(source, given)
        ops = list(circuit.instructions)
        if shots > 0:
            ops += openqasm.basis_rotations(circuit.results, sorted(circuit.qubit_set))
        ctx = {"deps": list(circuit.input_deps), "inputs": list(circuit.inputs)}
        # with shots the reference returns the program's result list itself
        # (simulator.py:437-447)
        return self._simulate(ir, ops, circuit.results, set(circuit.qubit_set), shots,
                              circuit.measured_qubits or None, adjoint_ctx=ctx,
                              shots_results=list(circuit.results), **kwargs)

    def _simulate(self, ir, ops, results, used, shots, measured_qubits, *,
                  qubit_count_cpu: Optional[int] = None, qubit_count_cuda: Optional[int] = None,
                  qubit_count_slice: Optional[int] = None, accelerator: Optional[str] = None,
                  prefetch: Optional[str] = None, offload: Optional[str] = None,
                  path=None, block_count: Optional[int] = None, qubit_count: Optional[int] = None,
                  precision: str = "complex128", fuse: int = 5, exact_sampling: bool = True,
                  adjoint_ctx: Optional[dict] = None, shots_results: Optional[list] = None):
        qmap = self._contiguous_qubit_mapping(used)
        for op in ops:
            op.targets = [qmap[q] for q in op.targets]
        in_order = list(ops)
        qubit_count = max(len(qmap), 1)
        acc = self._validate_simulation_type(accelerator)
        self._validate_prefetch_type(prefetch)
        off = self._validate_offload_type(offload)
        if acc != AcceleratorType.HYBRID and off != OffloadType.CPU:
            self._validate_qubit_count(qubit_count, precision)
        ops = pseudo_sort_operations_descending(ops)
        sub = Subcircuit(qubit_count=qubit_count, accelerator=acc, offload=off, operations=ops,
                         qubit_count_cuda=qubit_count_cuda, qubit_count_slice=qubit_count_slice)
        if acc == AcceleratorType.HYBRID or off == OffloadType.CPU:
            # the state in pinned host memory, streamed through device slices
            sim = HostStateVectorSimulation(qubit_count, shots, sub, precision=precision,
                                            fuse=fuse, local_qubits=qubit_count_cuda,
                                            contiguous_bits=qubit_count_slice)
        else:
            sim = StateVectorSimulation(qubit_count, shots, sub, precision=precision, fuse=fuse,
                                        exact_sampling=exact_sampling)
        sim.evolve(sub)
        values = []
        kinds = [str(getattr(_get(r, "type"), "value", _get(r, "type"))).lower() for r in results]
        if shots and "adjoint_gradient" in kinds:
            raise ValueError("adjoint_gradient needs shots=0")
        if not shots and "sample" in kinds:
            raise ValueError("the Sample result type needs shots > 0")
        if not shots:
            for r in results:
                if _get(r, "type") == "adjoint_gradient":
                    v = self._adjoint_gradient(r, sim, in_order, adjoint_ctx, qmap, qubit_count,
                                               precision)
                else:
                    v = self._calculate(r, sim, qmap, qubit_count)
                values.append({"type": r, "value": v})
        measurements = []
        if shots:
            mapped = [qmap[q] for q in measured_qubits] if measured_qubits else None
            measurements = self._formatted_measurements(sim.retrieve_samples(), qubit_count,
                                                        mapped)
        measured = list(measured_qubits) if measured_qubits else list(range(qubit_count))
        meta = {"id": str(uuid.uuid4()), "shots": shots, "deviceId": self.DEVICE_ID}
        try:  # the Braket schemas, when installed
            from braket.task_result import (AdditionalMetadata, GateModelTaskResult,
                                            ResultTypeValue, TaskMetadata)

            return GateModelTaskResult.construct(
                taskMetadata=TaskMetadata(**meta),
                additionalMetadata=AdditionalMetadata(action=ir),
                resultTypes=(list(shots_results or []) if shots else
                             [ResultTypeValue.construct(type=v["type"], value=v["value"])
                              for v in values]),
                measurements=measurements, measuredQubits=measured)
        except ImportError:
            return TaskResult(measurements=measurements, measuredQubits=measured,
                              resultTypes=list(shots_results or []) if shots else values,
                              taskMetadata=meta)

    @staticmethod
    def _calculate(r, sim: StateVectorSimulation, qmap, n) -> Any:
        t = str(getattr(_get(r, "type"), "value", _get(r, "type"))).lower()
        targets = [qmap[int(q)] for q in (_get(r, "targets") or [])]
        if t == "statevector":
            return sim.state_vector
        if t == "amplitude":
            states = _get(r, "states")
            amps = sim.amplitudes([int(s, 2) for s in states])
            return {s: complex(a) for s, a in zip(states, amps)}
        if t == "probability":
            if not targets:
                return sim.probabilities
            if len(targets) > 20 and sorted(targets) != list(range(n)):
                raise ValueError("Probability over more than 20 of the qubits: request all "
                                 "qubits or at most 20")
            return sim.marginal_probabilities(targets)
        if t in ("expectation", "variance"):
            obs = _get(r, "observable")
            if any(not isinstance(o, str) for o in obs):
                return BaseSimulator._hermitian_value(t, sim, obs, targets, n)
            letters = [str(o).upper() for o in obs]
            if any(L not in ("X", "Y", "Z", "I", "H") for L in letters):
                raise NotImplementedError("only Pauli products with H and I factors and "
                                          "Hermitian matrices are supported")
            if len(letters) == 1 and not targets:
                # no targets: the observable on each qubit in turn
                return [BaseSimulator._observable_value(t, sim, letters, [q]) for q in range(n)]
            if len(letters) == 1 and len(targets) > 1:
                letters = letters * len(targets)
            return BaseSimulator._observable_value(t, sim, letters, targets)
        raise NotImplementedError(f"Result type {t} not supported")

    @staticmethod
    def _adjoint_gradient(r, sim: StateVectorSimulation, ops, ctx, qmap, n, precision) -> dict:
        """Braket's AdjointGradient result: the expectation of a sum of Pauli
        products and its gradient with respect to program inputs.  Every
        (gate, angle) that depends on a requested input is one parameter of
        the engine's adjoint sweep (Program.adjoint); the recorded
        d angle / d input combine those into d<H>/d input (chain rule)."""
        if ctx is None:
            raise NotImplementedError("adjoint_gradient needs an OpenQASM program with inputs")
        names = r["parameters"] if r["parameters"] is not None else list(ctx["inputs"])
        want = set(names)
        keep, pmap, slots = [], [], []
        for op, d in zip(ops, ctx["deps"]):
            if isinstance(d, set):
                if d & want:
                    raise NotImplementedError(
                        f"{op} depends on {sorted(d & want)} through a fractional power or a "
                        "matrix gate; its gradient is not supported")
                d = []
            if not list(op.targets):
                continue  # a global phase changes no expectation
            entry = []
            for a, coefs in d:
                if want.intersection(coefs):
                    entry.append((a, len(slots)))
                    slots.append(coefs)
            keep.append(op)
            pmap.append(entry)
        paulis, qubits, coeffs, const = [], [], [], 0.0
        for coef, letters, qs in r["observable"]:
            strings = [("", [], float(coef))]
            for L, q in zip(letters, qs):
                L = L.upper()
                if L == "I":
                    continue
                opts, w = (("X", "Z"), math.sqrt(0.5)) if L == "H" else ((L,), 1.0)
                strings = [(st + o, t + [qmap[q]], c * w) for st, t, c in strings for o in opts]
            for st, t, c in strings:
                if st:
                    paulis.append(st)
                    qubits.append(t)
                    coeffs.append(c)
                else:
                    const += c
        grads = [0.0] * len(slots)
        if not paulis:
            e = 0.0
        elif not slots:
            e = sim.expectation(paulis, qubits, coeffs)
        else:
            mk = lambda: _C.result_types.StateVector(_C.DeviceType.CUDA, n,  # noqa: E731
                                                     dtype=precision)
            prog = _C.functionals.Program(n, keep, dtype=precision)
            e, grads = prog.adjoint(mk(), mk(), mk(), [], pmap, len(slots), paulis, qubits,
                                    coeffs)
        gradient = {nm: 0.0 for nm in names}
        for k, coefs in enumerate(slots):
            for nm, c in coefs.items():
                if nm in gradient:
                    gradient[nm] += c * float(grads[k])
        return {"expectation": float(e) + const, "gradient": gradient}

    @staticmethod
    def _hermitian_value(t, sim: StateVectorSimulation, obs, targets, n):
        """Expectation or variance of a tensor product with Hermitian-matrix
        factors (JAQCD: a matrix as [[[re, im], ...], ...]; OpenQASM:
        hermitian(...)), each factor taking log2(dim) consecutive targets."""
        factors = [openqasm.observable_factor(o) for o in obs]
        if not targets:
            if len(factors) != 1 or factors[0][0] != 1:
                raise ValueError(f"a {t} without targets needs a one-qubit observable")
            return [BaseSimulator._hermitian_value(t, sim, obs, [q], n) for q in range(n)]
        if sum(f[0] for f in factors) != len(targets):
            raise ValueError("observable factors do not match the number of targets")
        if len(set(targets)) != len(targets) or len(targets) > 20:
            raise ValueError("observable targets must be distinct and at most 20")
        rotations, lam, k = [], np.ones(1), 0
        for nq, make_rotation, eig in factors:
            qs = list(targets[k:k + nq])
            k += nq
            rotations += make_rotation(qs)
            lam = np.kron(lam, eig)
        e, e2 = sim.diagonalized_moments(rotations, targets, lam)
        return e if t == "expectation" else e2 - e * e

    @staticmethod
    def _observable_value(t, sim: StateVectorSimulation, letters, targets) -> float:
        # H = (X + Z)/sqrt(2) per factor: a sum of Pauli strings
        strings, coeffs = [""], [1.0]
        for L in letters:
            opts = ("X", "Z") if L == "H" else (L,)
            w = math.sqrt(0.5) if L == "H" else 1.0
            strings = [s + o for s in strings for o in opts]
            coeffs = [c * w for c in coeffs for _ in opts]
        e = sim.expectation(strings, [targets] * len(strings), coeffs)
        # every factor squares to I: Var = <O^2> - <O>^2 = 1 - <O>^2
        return e if t == "expectation" else 1.0 - e * e


class StateVectorSimulator(BaseSimulator):
    DEVICE_ID = "snuqs"

    def __init__(self):
        self.max_qubits = 34

    @property
    def properties(self) -> dict:
        results = [{"name": n} for n in
                   ("StateVector", "Amplitude", "Probability", "Expectation", "Variance")]
        return {
            "service": {"shotsRange": [0, 1 << 20]},
            "action": {
                "braket.ir.jaqcd.program": {
                    "actionType": "braket.ir.jaqcd.program", "version": ["1"],
                    "supportedOperations": sorted(k for k in BRAKET_GATES
                                                  if k not in ("U", "gphase")),
                    "supportedResultTypes": results,
                },
                "braket.ir.openqasm.program": {
                    "actionType": "braket.ir.openqasm.program", "version": ["1"],
                    "supportedOperations": sorted(
                        [k for k in BRAKET_GATES if k != "unitary"] +
                        ["cv", "ecr", "gpi", "gpi2", "ms", "prx"]),
                    "supportedModifiers": [
                        {"name": "ctrl"}, {"name": "negctrl"},
                        {"name": "pow", "exponent_types": ["int", "float"]}, {"name": "inv"}],
                    "supportedPragmas": ["braket_unitary_matrix", "braket_verbatim",
                                         "braket_result_type_adjoint_gradient",
                                         "braket_result_type_state_vector",
                                         "braket_result_type_amplitude",
                                         "braket_result_type_probability",
                                         "braket_result_type_expectation",
                                         "braket_result_type_variance"],
                    "supportedResultTypes": results,
                },
            },
            "paradigm": {"qubitCount": self.max_qubits},
        }

    def initialize_simulation(self, *, qubit_count: int, shots: int, subcircuit: Subcircuit,
                              **kw) -> StateVectorSimulation:
        return StateVectorSimulation(qubit_count, shots, subcircuit, **kw)
