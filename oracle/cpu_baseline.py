"""CPU baseline for bench.py (TEST/MEASUREMENT INFRASTRUCTURE ONLY).

Times the reference's own CPU gate path -- cpu::applyGate from
src/braket/snuqs/csrc/operation/gate_operations_impl_cpu.cc:131-148, compiled
from the reference sources into oracle/_ref/libref_cpu.so by oracle/Makefile
-- on a bounded prefix of the benchmark circuit, complex128 (the reference's
only precision), with the reference's own OpenMP loop.  Falls back to the C
restatement (oracle/libsv_oracle.so, kind "port") when the reference build is
absent.  Prints one JSON object.

Run as a child process of bench.py so the OpenMP team and the 2^n-amplitude
host buffer live in their own process:
    python -m oracle.cpu_baseline --qubits 30 --depth 30 --seconds 15
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=30)
    ap.add_argument("--seconds", type=float, default=15.0)
    ap.add_argument("--max-gates", type=int, default=200)
    a = ap.parse_args()

    from oracle import oracle as O
    from snuqs_amd.circuits import random_layered

    n = a.qubits
    ops = random_layered(n, a.depth, seed=1234)
    lib = O.ref_lib()
    kind = "reference" if lib is not None else "port"
    sv = np.empty(1 << n, dtype=np.complex128)
    if lib is not None:
        lib.ref_initialize_basis_z(sv.ctypes.data, sv.size)
    else:
        sv[:] = 0
        sv[0] = 1
    apply = O.ref_apply if lib is not None else O.apply
    N = 1 << n
    done, nbytes = 0, 0.0
    per_class = {}  # gate class -> [gates, seconds]
    t0 = time.perf_counter()
    for op in ops[: a.max_gates]:
        m = np.array(op.matrix())
        tg = time.perf_counter()
        apply(sv, m, list(op.targets), n)
        cls = "diagonal_k1" if len(op.targets) == 1 else f"dense_k{len(op.targets)}"
        c = per_class.setdefault(cls, [0, 0.0])
        c[0] += 1
        c[1] += time.perf_counter() - tg
        done += 1
        nbytes += 2.0 * N * 16  # Unitary and RotZ both change every amplitude
        if time.perf_counter() - t0 > a.seconds:
            break
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    print(json.dumps({
        "value": nbytes / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": kind,
        "sample": f"first {done} gates of the {n}-qubit depth-{a.depth} random layered circuit, "
                  f"complex128, cpu::applyGate, {dt:.1f} s",
        "gates": done, "seconds": dt, "cpu_model": model,
        "per_class_GBps": {k: v[0] * 2.0 * N * 16 / v[1] / 1e9 for k, v in per_class.items()},
        "host_cpus": os.cpu_count(),
    }))


if __name__ == "__main__":
    main()
