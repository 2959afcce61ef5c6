"""Seeded circuit generators for the benchmark configurations and tests.

Gate lists are lists of ``_C.operation.GateOperation`` objects (targets in the
reference's order: controls first, qubit 0 = MSB of the amplitude index).

* ``random_gate_list`` -- H on every qubit, then random instructions drawn
  uniformly from the 33 gate kinds with random distinct targets in random
  order and angles U[0, 2pi) (test/braket/test_random_circuit.py:13-99,
  112-148 of the reference; seeded here).
* ``ghz`` -- H(0), CX(0, i) for i = 1..n-1.
* ``random_layered`` -- per layer a seeded qubit permutation, a Haar-random
  4x4 Unitary on each pair (benchmark/quantum_volume.py:30-41), then RotZ with
  U[0, 2pi) angles on every qubit.
* ``hea_ansatz`` -- layers of RotY and RotZ on every qubit and a CZ ring.
* ``random_pauli_sum`` -- weight-w Pauli strings with N(0, 1) coefficients.
* ``qaoa_maxcut`` -- QAOA on a random 3-regular graph.
* ``random_u_cx`` -- layers of Haar 1-qubit U on every qubit and CX on
  disjoint random pairs.
"""
from __future__ import annotations

import math

import numpy as np

from ._C.operation import gate_operations as g

# (constructor, number of targets, number of angles); order of the reference test
GATE_KINDS = [
    (g.Identity, 1, 0), (g.Hadamard, 1, 0), (g.PauliX, 1, 0), (g.PauliY, 1, 0),
    (g.PauliZ, 1, 0), (g.CX, 2, 0), (g.CY, 2, 0), (g.CZ, 2, 0), (g.S, 1, 0), (g.Si, 1, 0),
    (g.T, 1, 0), (g.Ti, 1, 0), (g.V, 1, 0), (g.Vi, 1, 0), (g.PhaseShift, 1, 1),
    (g.CPhaseShift, 2, 1), (g.CPhaseShift00, 2, 1), (g.CPhaseShift01, 2, 1),
    (g.CPhaseShift10, 2, 1), (g.RotX, 1, 1), (g.RotY, 1, 1), (g.RotZ, 1, 1),
    (g.Swap, 2, 0), (g.ISwap, 2, 0), (g.PSwap, 2, 1), (g.XY, 2, 1), (g.XX, 2, 1),
    (g.YY, 2, 1), (g.ZZ, 2, 1), (g.CCNot, 3, 0), (g.CSwap, 3, 0), (g.U, 1, 3),
    (g.GPhase, 0, 1),
]


def make_gate(kind, targets, angles):
    ctor, _, na = kind
    targets = [int(t) for t in targets]
    if na == 0:
        return ctor(targets)
    if na == 1:
        return ctor(targets, float(angles[0]))
    return ctor(targets, *[float(a) for a in angles[:3]])


def random_gate_list(n: int, ngates: int, rng: np.random.Generator, h_layer: bool = True,
                     with_unitary: bool = False):
    ops = [g.Hadamard([q]) for q in range(n)] if h_layer else []
    kinds = list(GATE_KINDS)
    for _ in range(ngates):
        idx = int(rng.integers(0, len(kinds) + (1 if with_unitary else 0)))
        angles = rng.random(3) * 2 * math.pi
        if idx == len(kinds):
            k = int(rng.integers(1, min(3, n) + 1))
            t = rng.permutation(n)[:k]
            ops.append(g.Unitary([int(x) for x in t], haar_unitary(1 << k, rng)))
            continue
        kind = kinds[idx]
        t = rng.permutation(n)[: kind[1]]
        ops.append(make_gate(kind, t, angles))
    return ops


def ghz(n: int):
    return [g.Hadamard([0])] + [g.CX([0, i]) for i in range(1, n)]


def haar_unitary(d: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random U(d): QR of a complex Ginibre matrix with the phase fix."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r) / np.abs(np.diagonal(r))
    return q * ph


def random_layered(n: int, depth: int, seed: int = 1234):
    """Config: each layer pairs qubits by a seeded permutation, applies a Haar
    4x4 Unitary to each pair, then RotZ(U[0, 2pi)) on every qubit."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(depth):
        perm = rng.permutation(n)
        for w in range(n // 2):
            ops.append(g.Unitary([int(perm[2 * w]), int(perm[2 * w + 1])], haar_unitary(4, rng)))
        for q, a in enumerate(rng.uniform(0, 2 * math.pi, n)):
            ops.append(g.RotZ([q], float(a)))
    return ops


def hea_ansatz(n: int, layers: int, thetas):
    """RotY(theta) and RotZ(theta) on every qubit, then CZ(i, (i+1) % n), per
    layer.  Returns (ops, param_map) with param_map[g] = [(angle, param)]."""
    thetas = np.asarray(thetas, dtype=float)
    assert thetas.size == 2 * n * layers
    ops, pmap = [], []
    p = 0
    for _ in range(layers):
        for ctor in (g.RotY, g.RotZ):
            for q in range(n):
                ops.append(ctor([q], float(thetas[p])))
                pmap.append([(0, p)])
                p += 1
        for q in range(n):
            ops.append(g.CZ([q, (q + 1) % n]))
            pmap.append([])
    return ops, pmap


def random_pauli_sum(n: int, nterms: int, weight: int, rng: np.random.Generator):
    paulis, qubits, coeffs = [], [], []
    for _ in range(nterms):
        qs = rng.choice(n, weight, replace=False)
        paulis.append("".join("XYZ"[int(i)] for i in rng.integers(0, 3, weight)))
        qubits.append([int(q) for q in qs])
        coeffs.append(float(rng.standard_normal()))
    return paulis, qubits, coeffs


def random_3_regular(n: int, rng: np.random.Generator):
    """Random 3-regular simple graph by the pairing model with rejection."""
    assert n % 2 == 0 and n >= 4
    while True:
        stubs = rng.permutation(np.repeat(np.arange(n), 3))
        edges = set()
        ok = True
        for a, b in zip(stubs[0::2], stubs[1::2]):
            a, b = int(min(a, b)), int(max(a, b))
            if a == b or (a, b) in edges:
                ok = False
                break
            edges.add((a, b))
        if ok:
            return sorted(edges)


def qaoa_maxcut(n: int, edges, gammas, betas):
    """H on every qubit, then p layers of ZZ(2 gamma) on every edge and
    RotX(2 beta) on every qubit.  Energy <C> = sum_edges (1 - <Z_i Z_j>) / 2."""
    ops, pmap = [g.Hadamard([q]) for q in range(n)], [[] for _ in range(n)]
    p = len(gammas)
    for layer in range(p):
        for a, b in edges:
            ops.append(g.ZZ([a, b], 2.0 * float(gammas[layer])))
            pmap.append([(0, layer)])
        for q in range(n):
            ops.append(g.RotX([q], 2.0 * float(betas[layer])))
            pmap.append([(0, p + layer)])
    return ops, pmap


def random_u_cx(n: int, layers: int, seed: int):
    """Layers of Haar-random 1-qubit U on every qubit, then CX on n/2 disjoint
    pairs from a seeded permutation."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        for q in range(n):
            ops.append(g.Unitary([q], haar_unitary(2, rng)))
        perm = rng.permutation(n)
        for w in range(n // 2):
            ops.append(g.CX([int(perm[2 * w]), int(perm[2 * w + 1])]))
    return ops
