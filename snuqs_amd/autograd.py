"""torch.autograd integration of the adjoint gradients.

* ``CircuitEnergy`` -- E(theta) = <psi(theta)| H |psi(theta)> for a
  parametrised gate list and a Pauli-sum H.  The forward pass runs the
  engine's adjoint (``Program.adjoint``: forward sweep, lambda = H psi, reverse
  sweep with one FP64 inner product per parameter) and keeps the gradient for
  backward, so a training step costs one adjoint evaluation.  The program and
  its states are cached per circuit shape and re-parametrised in place.
* ``QaoaEnergies`` -- the energies of a ``QaoaBatch`` (many MaxCut instances of
  uneven size); backward uses the batch's adjoint gradients.

Parameters and results are CPU float64 tensors; the states live on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _C


class _Workspace:
    """A program and its device states for one circuit shape (qubits, dtype,
    checkpoints, every gate's targets), reused across evaluations: a new
    parameter set only replaces the gate matrices (Program.set_gates), so a
    training step allocates no device memory."""

    def __init__(self, n, ops, dtype, checkpoints):
        self.prog = _C.functionals.Program(n, ops, dtype=dtype)
        mk = lambda: _C.result_types.StateVector(_C.DeviceType.CUDA, n, dtype=dtype)  # noqa: E731
        self.sv, self.w0, self.w1 = mk(), mk(), mk()
        self.cps = [mk() for _ in range(checkpoints)]


_CACHE: "dict[tuple, _Workspace]" = {}
_CACHE_SIZE = 4  # circuit shapes kept (each holds 3 + checkpoints states)


def _workspace(n, ops, dtype, checkpoints) -> _Workspace:
    key = (n, dtype, checkpoints, tuple(tuple(op.targets) for op in ops))
    ws = _CACHE.pop(key, None)
    if ws is None:
        while len(_CACHE) >= _CACHE_SIZE:
            _CACHE.pop(next(iter(_CACHE)))
        ws = _Workspace(n, ops, dtype, checkpoints)
    else:
        ws.prog.set_gates(ops)
    _CACHE[key] = ws  # most recent last
    return ws


def clear_cache() -> None:
    """Release the cached programs and states."""
    _CACHE.clear()


class CircuitEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thetas, builder, observable, dtype="complex128", checkpoints=0):
        """builder(thetas: np.ndarray) -> (ops, param_map); observable =
        (paulis, qubits, coeffs)."""
        th = thetas.detach().cpu().double().numpy()
        ops, pmap = builder(th)
        n = 1 + max((q for op in ops for q in op.targets), default=0)
        n = max(n, 1 + max((q for qs in observable[1] for q in qs), default=0))
        ws = _workspace(n, ops, dtype, checkpoints)
        e, grads = ws.prog.adjoint(ws.sv, ws.w0, ws.w1, ws.cps, pmap, th.size, *observable)
        ctx.save_for_backward(torch.from_numpy(np.asarray(grads, dtype=np.float64)))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None, None


class QaoaEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gammas, betas, batch, use_graph=True):
        out = batch.run(gammas.detach().cpu().double().numpy(),
                        betas.detach().cpu().double().numpy(), use_graph=use_graph)
        p = gammas.shape[1]
        ctx.save_for_backward(torch.from_numpy(out[:, 1:1 + p].copy()),
                              torch.from_numpy(out[:, 1 + p:].copy()))
        return torch.from_numpy(out[:, 0].copy())

    @staticmethod
    def backward(ctx, grad_out):
        dg, db = ctx.saved_tensors
        return grad_out[:, None] * dg, grad_out[:, None] * db, None, None


def circuit_energy(thetas, builder, observable, dtype="complex128", checkpoints=0):
    return CircuitEnergy.apply(thetas, builder, observable, dtype, checkpoints)


def qaoa_energies(gammas, betas, batch, use_graph=True):
    return QaoaEnergies.apply(gammas, betas, batch, use_graph)
