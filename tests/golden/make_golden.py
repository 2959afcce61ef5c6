"""Writes the golden fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py

* choice_fixtures.npz -- black-box vectors of numpy's legacy sampler, the
  algorithm behind np.random.choice in src/braket/snuqs/simulation.py:220:
  for each case a probability vector p, a seed s, and
  RandomState(s).choice(len(p), p=p, size=k) together with
  RandomState(s).random_sample(k).  Cases include zeros, exact ties, tiny
  entries and a random state's |psi|^2.
* ref_circuits.npz -- final states of seeded random circuits computed with the
  reference's own CPU gate path (cpu::applyGate compiled from
  src/braket/snuqs/csrc/operation/gate_operations_impl_cpu.cc into
  oracle/_ref/libref_cpu.so), with each gate's matrix and targets, so the
  vectors are checkable without the reference tree.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))


def choice_cases():
    rng = np.random.default_rng(2024)
    cases = []
    p = rng.random(1000)
    cases.append(("uniform_random_1000", p / p.sum(), 11, 5000))
    p = np.zeros(4096)
    p[[0, 4095]] = 0.5
    cases.append(("ghz_4096", p, 12, 3000))
    p = np.zeros(512)
    p[100:164] = 1 / 256
    p[300] = 0.75
    cases.append(("ties_and_zeros_512", p, 13, 4000))
    v = rng.standard_normal(1 << 14) + 1j * rng.standard_normal(1 << 14)
    v /= np.linalg.norm(v)
    cases.append(("haar_like_state_16q", np.abs(v) ** 2, 14, 20000))
    p = rng.random(2000) ** 8
    p[::7] = 0
    cases.append(("heavy_tailed_2000", p / p.sum(), 15, 5000))
    return cases


def main():
    out = {}
    for name, p, seed, k in choice_cases():
        out[f"{name}__p"] = p
        out[f"{name}__seed"] = np.array(seed)
        out[f"{name}__u"] = np.random.RandomState(seed).random_sample(k)
        out[f"{name}__idx"] = np.random.RandomState(seed).choice(len(p), p=p, size=k)
    np.savez_compressed(os.path.join(OUT, "choice_fixtures.npz"), **out)

    from oracle import oracle as O
    from snuqs_amd.circuits import random_gate_list

    if O.ref_lib() is None:
        print("reference CPU build absent: ref_circuits.npz not regenerated")
        return
    circ = {}
    for n, ngates, seed in ((8, 120, 1), (11, 200, 2)):
        ops = random_gate_list(n, ngates, np.random.default_rng(seed))
        st = O.basis_state(n)
        mats, tgts, k = [], [], []
        for op in ops:
            m = np.array(op.matrix())
            t = list(op.targets)
            O.ref_apply(st, m, t, n)  # k = 0 (GPhase) takes applyGlobalPhase
            mats.append(m.reshape(-1))
            tgts.append(t)
            k.append(len(t))
        circ[f"n{n}__state"] = st
        circ[f"n{n}__k"] = np.array(k)
        circ[f"n{n}__targets"] = np.array([x for t in tgts for x in t], dtype=np.int64)
        circ[f"n{n}__matrices"] = np.concatenate(mats)
    np.savez_compressed(os.path.join(OUT, "ref_circuits.npz"), **circ)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
