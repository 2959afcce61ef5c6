"""Autotuner for the engine's process-wide launch-shape knobs (``sq_tune_set``).

For each dtype it builds a fused random layered circuit (the benchmark's
circuit family), times every combination of the knobs with per-kernel HIP
events and keeps the fastest; ``save`` writes the winners to a JSON file that
``load`` (called when the package is imported, for snuqs_amd/tuning.json)
applies.  Run on the target GPU:

    python -m snuqs_amd.autotune --qubits 28 --out snuqs_amd/tuning.json

These knobs are the defaults; ``Program.tune(sv)`` (``sq_program_tune``)
picks a shape per gate of one program on top of them.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os

from . import _C

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
SPACE = {
    "mfma_blocks_per_cu": [32, 64, 128, 256, 512, 1024],
    "mfma5_pipeline": [0, 1],
    "mfma5_order": [0, 1, 2],
}
SUFFIX = {"complex64": "c64", "complex128": "c128"}


def load(path: str = DEFAULT_PATH) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        table = json.load(f)
    for k, v in table.get("knobs", {}).items():
        _C.core.tune_set(k, int(v))
    return table


def _time_program(prog, sv, reps=2):
    prog.run(sv)
    _C.core.cuda.device_synchronize()
    return min(float(prog.run_timed(sv).sum()) for _ in range(reps))


def tune(qubits: int = 28, depth: int = 12, dtypes=("complex64", "complex128"), log=print):
    from .circuits import random_layered
    from .transpile import fuse_operations

    knobs, report = {}, {}
    for dtype in dtypes:
        n = qubits if dtype == "complex64" else qubits - 1
        fused = fuse_operations(random_layered(n, depth, seed=99), 5)
        prog = _C.functionals.Program(n, fused, dtype=dtype)
        sv = _C.result_types.StateVector(_C.DeviceType.CUDA, n, dtype=dtype)
        _C.functionals.initialize_basis_z(sv)
        sfx = SUFFIX[dtype]
        best = None
        for combo in itertools.product(*SPACE.values()):
            for key, val in zip(SPACE, combo):
                _C.core.tune_set(f"{key}_{sfx}", val)
            ms = _time_program(prog, sv)
            report[f"{dtype} {dict(zip(SPACE, combo))}"] = ms
            log(f"{dtype} {dict(zip(SPACE, combo))}: {ms:.2f} ms")
            if best is None or ms < best[0]:
                best = (ms, combo)
        for key, val in zip(SPACE, best[1]):
            knobs[f"{key}_{sfx}"] = val
            _C.core.tune_set(f"{key}_{sfx}", val)
        del sv, prog
    return {"knobs": knobs, "qubits": qubits, "depth": depth, "timings_ms": report}


def save(table: dict, path: str = DEFAULT_PATH) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    table = tune(a.qubits, a.depth)
    save(table, a.out)
    print(json.dumps(table["knobs"]))


if __name__ == "__main__":
    main()
