"""Batched QAOA MaxCut workloads (benchmark configuration 4).

``make_batch(num_instances, sizes, p, seed)`` draws instances with qubit
counts uniform over ``sizes``, seeded random 3-regular graphs (pairing model
with rejection; 3-regular graphs need an even qubit count, so sizes are the
even values of 12..26) and parameters gamma ~ U[0, pi], beta ~ U[0, pi/2].
``QaoaBatch`` (the engine, ``_C.functionals.QaoaBatch``) returns per instance
E = sum_edges (1 - <Z_a Z_b>) / 2 and dE/dgamma, dE/dbeta by the adjoint
method.  Graphs must be simple (an edge given twice raises).
"""
from __future__ import annotations

import math

import numpy as np

from . import _C
from .circuits import random_3_regular

EVEN_SIZES = tuple(range(12, 27, 2))


def make_batch(num_instances=256, sizes=EVEN_SIZES, p=4, seed=7):
    rng = np.random.default_rng(seed)
    nq = [int(rng.choice(sizes)) for _ in range(num_instances)]
    graphs = [random_3_regular(n, rng) for n in nq]
    gammas = rng.uniform(0, math.pi, (num_instances, p))
    betas = rng.uniform(0, math.pi / 2, (num_instances, p))
    return nq, graphs, gammas, betas


def qaoa_batch(nq, graphs, p, dtype="complex64"):
    return _C.functionals.QaoaBatch([int(n) for n in nq],
                                    [[(int(a), int(b)) for a, b in gph] for gph in graphs],
                                    int(p), dtype=dtype)
