#!/usr/bin/env python3
"""Benchmark: HBM GB/s of gate sweeps on one MI355X, plus adjoint gradients/s
against parameter shift.

Workload (one "step"): the random layered circuit on 30 qubits, depth 30,
complex64 -- each layer pairs the qubits by a seeded permutation, applies a
Haar-random 4x4 Unitary to each pair and RotZ(U[0, 2pi)) to every qubit
(1350 gates) -- applied to a device-resident |0...0> state (8 GiB).  The
scheduler fuses the gates into <= 5-qubit dense blocks that run on the matrix
cores; the unfused per-gate sweep is measured too.

value = algorithmic HBM bytes of the launched kernels / wall time of the timed
steps (2 x touched amplitudes x 8 B per kernel).  Per-kernel durations come
from HIP events recorded on the launch stream around every kernel of every
timed step.

Also reported: the dominant kernel's roofline, a plain copy kernel's
bandwidth, the 24-qubit hardware-efficient ansatz gradient by adjoint vs
parameter shift on the same GPU, and the reference CPU path
(cpu::applyGate) timed on a bounded sample on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "HBM GB/s of gate sweeps at 30q c64; adjoint grads/s vs param-shift"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
CLASS_NAMES = {0: "none", 1: "phase", 2: "diagonal", 3: "permutation", 4: "dense_1_3",
               5: "dense_4_5_mfma", 6: "dense_4_5_valu"}
KERNEL_NAMES = {1: "k_phase", 2: "k_diag", 3: "k_perm", 4: "k_dense", 5: "k_mfma{k}_{d}",
                6: "k_dense_lds"}


def kernel_name(cls, kk, dtype):
    name = KERNEL_NAMES.get(cls, "?").format(k=kk, d="c64" if dtype == "complex64" else "c128")
    if name == "k_mfma5_c64":  # the tuner picks the 32- or 64-group tile per gate
        name += "/k_mfma5w_c64"
    return name


def pmc_traffic(name):
    """HBM bytes per launch of `name` from the committed PMC summary
    (profiles/pmc_traffic.json: FETCH_SIZE and WRITE_SIZE runs, scaled by the
    copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(name)
        return None if k is None else k["read_bytes"] + k["write_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def sweep_stats(C, prog, sv, steps, warmup):
    """Warm up, then time `steps` runs with per-kernel events.  Returns wall
    seconds of the timed steps, per-gate ms (summed over steps) and the
    per-gate (class, bytes)."""
    for _ in range(warmup):
        prog.run(sv)
    C.core.cuda.device_synchronize()
    ng = len(prog)
    acc = np.zeros(ng)
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        acc += prog.run_timed(sv)
        log(f"step {s + 1}/{steps} done")
    C.core.cuda.device_synchronize()
    wall = time.perf_counter() - t0
    info = [prog.gate_info(g) for g in range(ng)]
    return wall, acc, info


def class_table(acc_ms, info, steps):
    tab = {}
    for ms, (cls, kk, b) in zip(acc_ms, info):
        if cls == 0:
            continue
        key = (cls, kk)
        t = tab.setdefault(key, {"launches": 0, "ms": 0.0, "bytes": 0.0})
        t["launches"] += steps
        t["ms"] += float(ms)
        t["bytes"] += b * steps
    return tab


def run_adjoint_bench(C, n=24, layers=8, nterms=64, weight=4, shift_params=12):
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    rng = np.random.default_rng(2024)
    thetas = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    obs = random_pauli_sum(n, nterms, weight, rng)
    ops, pmap = hea_ansatz(n, layers, thetas)
    res = {}
    for dtype in ("complex128", "complex64"):
        prog = C.functionals.Program(n, ops, dtype=dtype)
        st = [C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype) for _ in range(3)]
        cps = [C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype) for _ in range(4)]
        prog.adjoint(st[0], st[1], st[2], cps, pmap, thetas.size, *obs)  # warm-up
        C.core.cuda.device_synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            e, gr = prog.adjoint(st[0], st[1], st[2], cps, pmap, thetas.size, *obs)
        dt = (time.perf_counter() - t0) / reps
        res[dtype] = {"energy": e, "grads": gr, "seconds": dt}
        log(f"adjoint {dtype}: {dt * 1e3:.1f} ms per {thetas.size}-component gradient")
    # parameter shift on the same GPU (complex64): each shifted circuit runs
    # from |0..0> through the engine's fastest forward path (the circuit fused
    # into <= 5-qubit blocks; the per-gate program is timed too) and its
    # energy is reduced on the device; programs are built before the clock.
    from snuqs_amd.transpile import fuse_operations

    sub = np.random.default_rng(1).choice(thetas.size, shift_params, replace=False)
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype="complex64")
    rates, es_by = {}, {}
    for mode in ("fused", "per_gate"):
        progs = []
        for j in sub:
            for sgn in (+1, -1):
                th = thetas.copy()
                th[j] += sgn * np.pi / 2
                ops_j = hea_ansatz(n, layers, th)[0]
                if mode == "fused":
                    ops_j = fuse_operations(ops_j, 5)
                progs.append(C.functionals.Program(n, ops_j, dtype="complex64"))
        C.functionals.initialize_basis_z(sv)
        progs[0].run(sv)
        C.functionals.expectation(sv, *obs)
        C.core.cuda.device_synchronize()
        t0 = time.perf_counter()
        es = []
        for p in progs:
            C.functionals.initialize_basis_z(sv)
            p.run(sv)
            es.append(C.functionals.expectation(sv, *obs)[0])
        rates[mode] = len(sub) / (time.perf_counter() - t0)
        es_by[mode] = es
    es = es_by["fused"]
    ps = 0.5 * (np.array(es[0::2]) - np.array(es[1::2]))
    g64, g128 = res["complex64"]["grads"], res["complex128"]["grads"]
    adj_rate = thetas.size / res["complex64"]["seconds"]
    ps_rate = rates["fused"]
    return {
        "config": f"{n}-qubit hardware-efficient ansatz, {layers} layers of RotY+RotZ on every "
                  f"qubit + CZ ring, {thetas.size} parameters; observable: {nterms} random "
                  f"weight-{weight} Pauli strings with N(0,1) coefficients",
        "adjoint_grads_per_s": adj_rate,
        "adjoint_ms_per_gradient_c64": res["complex64"]["seconds"] * 1e3,
        "adjoint_ms_per_gradient_c128": res["complex128"]["seconds"] * 1e3,
        "param_shift_grads_per_s": ps_rate,
        "param_shift_sample": f"{len(sub)} of {thetas.size} components ({2 * len(sub)} shifted "
                              f"circuits), complex64, circuits fused into <= 5-qubit blocks",
        "speedup": adj_rate / ps_rate,
        "param_shift_per_gate_grads_per_s": rates["per_gate"],
        "speedup_vs_per_gate_param_shift": adj_rate / rates["per_gate"],
        "max_abs_grad_err_c64_vs_c128": float(np.abs(g64 - g128).max()),
        "max_abs_param_shift_vs_adjoint_c128": float(np.abs(ps - g128[sub]).max()),
        "energy_c128": res["complex128"]["energy"],
    }


def state_error(a64, a128, chunk=1 << 26):
    """Error of a complex64 state against the complex128 state of the same
    circuit, in FP64: max-abs, max-abs relative to the largest amplitude,
    norm-wise relative ||a64 - a128|| / ||a128|| and 1 - fidelity
    |<a64|a128>|^2 / (<a64|a64><a128|a128>)."""
    mx = amax = d2 = n128 = n64 = 0.0
    ov = 0j
    for s in range(0, a128.size, chunk):
        x = a64[s:s + chunk].astype(np.complex128)
        y = a128[s:s + chunk]
        d = x - y
        mx = max(mx, float(np.abs(d).max()))
        amax = max(amax, float(np.abs(y).max()))
        d2 += float(np.vdot(d, d).real)
        n128 += float(np.vdot(y, y).real)
        n64 += float(np.vdot(x, x).real)
        ov += complex(np.vdot(x, y))
    return {"max_abs": mx, "max_abs_over_max_amplitude": mx / amax,
            "normwise_relative": (d2 / n128) ** 0.5,
            "one_minus_fidelity": 1.0 - abs(ov) ** 2 / (n64 * n128)}


def run_other_configs(C, depth):
    """The remaining configurations, one timed run each after a warm-up:
    config 2 in complex128 (30 qubits, 16 GiB), config 1 (GHZ-20 complex128,
    closed-form check) and config 5 (34-qubit complex64 and 33-qubit
    complex128 random U + CX circuits of 20 layers, fused, with the
    probabilities of 1e4 seeded bitstrings)."""
    import struct

    from snuqs_amd.circuits import ghz, random_layered, random_u_cx
    from snuqs_amd.transpile import fuse_operations

    out = {}

    def timed_program(n, ops, dtype, tune=True):
        prog = C.functionals.Program(n, ops, dtype=dtype)
        sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
        if tune:
            prog.tune(sv, reps=1)
        C.functionals.initialize_basis_z(sv)
        C.core.cuda.device_synchronize()
        t0 = time.perf_counter()
        prog.run(sv)
        C.core.cuda.device_synchronize()
        dt = time.perf_counter() - t0
        nbytes = prog.stats()["algorithmic_bytes"]
        return prog, sv, dt, nbytes

    # config 2, complex128
    ops = fuse_operations(random_layered(30, depth, seed=1234), 5)
    prog, sv, dt, nbytes = timed_program(30, ops, "complex128")
    out["config2_c128_30q"] = {"ms": dt * 1e3, "GBps": nbytes / dt / 1e9, "kernels": len(ops)}
    log(f"config 2 c128: {dt * 1e3:.0f} ms, {nbytes / dt / 1e9:.0f} GB/s")
    del prog
    # the complex64 state of the same circuit against this complex128 one
    sv64 = C.result_types.StateVector(C.DeviceType.CUDA, 30, dtype="complex64")
    C.functionals.initialize_basis_z(sv64)
    C.functionals.Program(30, ops, dtype="complex64").run(sv64)
    out["config2_c64_vs_c128_30q"] = state_error(np.asarray(sv64.cpu()), np.asarray(sv.cpu()))
    log("config 2 c64 vs c128:", out["config2_c64_vs_c128_30q"])
    del sv, sv64
    # config 1: GHZ-20 complex128
    prog, sv, dt, nbytes = timed_program(20, ghz(20), "complex128", tune=False)
    a = C.functionals.amplitudes(sv, [0, (1 << 20) - 1])
    bits = {struct.pack("<d", float(z.real)).hex() for z in a}
    out["config1_ghz20_c128"] = {"ms": dt * 1e3, "launches": 20,
                                 "amplitudes_bitwise_M_SQRT1_2": bits == {struct.pack("<d", 0.5 ** 0.5).hex()}}
    del prog, sv
    # config 5: largest single-GPU states
    for n, dtype in ((34, "complex64"), (33, "complex128")):
        free = C.core.cuda.mem_info()[0]
        need = (1 << n) * (8 if dtype == "complex64" else 16)
        if free < need + (4 << 30):
            out[f"config5_{dtype}_{n}q"] = "skipped: not enough device memory"
            continue
        ops = fuse_operations(random_u_cx(n, 20, seed=n), 5)
        # default launch shapes: tuning at 128 GiB would take ~50 s per state
        prog, sv, dt, nbytes = timed_program(n, ops, dtype, tune=False)
        q = np.random.default_rng(n).integers(0, 1 << n, 10**4, dtype=np.uint64)
        t0 = time.perf_counter()
        amps = C.functionals.amplitudes(sv, [int(x) for x in q])
        tq = time.perf_counter() - t0
        out[f"config5_{dtype}_{n}q"] = {
            "ms": dt * 1e3, "GBps": nbytes / dt / 1e9, "kernels": len(ops),
            "state_GiB": need / 2**30, "query_1e4_ms": tq * 1e3,
            "sum_queried_probabilities": float(np.sum(np.abs(amps) ** 2))}
        log(f"config 5 {dtype} {n} q: {dt * 1e3:.0f} ms, {nbytes / dt / 1e9:.0f} GB/s")
        del prog, sv
    return out


def run_batch_bench(C, n=20, layers=8, nsets=64, nterms=64, weight=4, reps=3):
    """Parameter batch: 64 parameter sets of a 20-qubit, 8-layer
    hardware-efficient ansatz (320 parameters), complex64, energies of a
    64-string weight-4 Pauli sum.  One batch = |0..0> in all 64 states, the
    fused circuit as one launch per block over all states (hipGraph) and one
    batched reduction; against the same 64 circuits run one after another
    (one fused program and one expectation call per set)."""
    from snuqs_amd.batch import ParameterBatch, fuse_operations_batch
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum

    rng = np.random.default_rng(2020)
    thetas = rng.uniform(0, 2 * np.pi, (nsets, 2 * n * layers))
    obs = random_pauli_sum(n, nterms, weight, rng)
    build = lambda t: hea_ansatz(n, layers, t)[0]  # noqa: E731
    pb = ParameterBatch(n, build, thetas, dtype="complex64")
    e_b = pb.energies(*obs, use_graph=True)  # warm-up and capture
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        e_b = pb.energies(*obs, use_graph=True)
    dt_b = (time.perf_counter() - t0) / reps
    # the same circuits one after another (programs built before the clock)
    fused = fuse_operations_batch([build(t) for t in thetas], 5)
    progs = [C.functionals.Program(n, f, dtype="complex64") for f in fused]
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype="complex64")
    C.functionals.initialize_basis_z(sv)
    progs[0].run(sv, use_graph=True)
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    e_s = []
    for p in progs:
        C.functionals.initialize_basis_z(sv)
        p.run(sv)
        e_s.append(C.functionals.expectation(sv, *obs)[0])
    dt_s = time.perf_counter() - t0
    st = pb.program.stats()
    log(f"parameter batch: {dt_b * 1e3:.1f} ms per {nsets} sets ({dt_s * 1e3:.1f} ms one by one)")
    return {
        "config": f"{nsets} parameter sets of a {n}-qubit hardware-efficient ansatz, {layers} layers "
                  f"({2 * n * layers} parameters), complex64, fused into <= 5-qubit blocks; energies "
                  f"of {nterms} random weight-{weight} Pauli strings",
        "ms_per_batch": dt_b * 1e3,
        "sets_per_s": nsets / dt_b,
        "ms_sequential": dt_s * 1e3,
        "sets_per_s_sequential": nsets / dt_s,
        "speedup": dt_s / dt_b,
        "launches_per_batch": st["launches"],
        "gate_GBps": st["algorithmic_bytes"] / dt_b / 1e9,
        "max_abs_energy_diff_batch_vs_sequential": float(np.abs(e_b - np.array(e_s)).max()),
    }


def run_cbatch_bench(C, ncirc=64, sizes=tuple(range(12, 23)), depth=10, reps=3):
    """Circuits of uneven qubit counts: 64 random layered circuits (a seeded
    pairing with a Haar 4x4 Unitary per pair and RotZ on every qubit per
    layer), depth 10, qubits uniform over 12..22, complex64, each fused into
    <= 5-qubit blocks.  One batch = every state set to |0..0> and every
    circuit applied, launch s applying block s of each circuit (hipGraph);
    against the same fused circuits run one after another as Programs
    (programs built before the clock, each run through its own hipGraph)."""
    from snuqs_amd.cbatch import circuit_batch
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    rng = np.random.default_rng(77)
    nq = [int(rng.choice(sizes)) for _ in range(ncirc)]
    circs = [(n, random_layered(n, depth, seed=1000 + j)) for j, n in enumerate(nq)]
    cb = circuit_batch(circs, dtype="complex64")
    cb.run(use_graph=True)  # capture
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        cb.run(use_graph=True)
    C.core.cuda.device_synchronize()
    dt_b = (time.perf_counter() - t0) / reps
    progs = [C.functionals.Program(n, fuse_operations(ops, 5), dtype="complex64") for n, ops in circs]
    svs = [C.result_types.StateVector(C.DeviceType.CUDA, n, dtype="complex64") for n in nq]
    for p, sv in zip(progs, svs):  # capture each program's graph
        C.functionals.initialize_basis_z(sv)
        p.run(sv, use_graph=True)
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p, sv in zip(progs, svs):
            C.functionals.initialize_basis_z(sv)
            p.run(sv, use_graph=True)
    C.core.cuda.device_synchronize()
    dt_s = (time.perf_counter() - t0) / reps
    err = max(float(np.abs(cb.state(j) - np.array(svs[j].cpu())).max()) for j in range(ncirc))
    st = cb.stats()
    log(f"circuit batch: {dt_b * 1e3:.1f} ms per {ncirc} circuits ({dt_s * 1e3:.1f} ms one by one)")
    return {
        "config": f"{ncirc} random layered circuits of depth {depth}, qubits uniform over "
                  f"{sizes[0]}..{sizes[-1]}, complex64, fused into <= 5-qubit blocks",
        "ms_per_batch": dt_b * 1e3,
        "circuits_per_s": ncirc / dt_b,
        "ms_sequential": dt_s * 1e3,
        "speedup": dt_s / dt_b,
        "launches_per_batch": st["launches"],
        "launches_sequential": int(sum(p.stats()["launches"] for p in progs)),
        "hbm_GBps": st["traffic_bytes"] / dt_b / 1e9,
        "state_GiB": st["state_bytes"] / 2**30,
        "max_abs_state_diff_batch_vs_sequential": err,
    }


def host_memory_gib():
    info = {}
    with open("/proc/meminfo") as f:
        for line in f:
            k, v = line.split(":", 1)
            info[k] = int(v.split()[0]) * 1024 / 2**30
    return info.get("MemTotal"), info.get("MemAvailable")


def run_host_tier_bench(C, n=32, m=30, k=22, depth=10):
    """States larger than the device slice: a 32-qubit complex64 state (32 GiB)
    in pinned host memory, evolved through 30-qubit device slices (three
    8 GiB buffers) by the random layered circuit of depth 10, pass by pass
    over PCIe (each pass reads and writes the whole host state)."""
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.host_tier import HostStateSimulation

    total, avail = host_memory_gib()
    need = (1 << n) * 8 / 2**30
    if avail is None or avail < need * 2 + 16:
        return f"skipped: {avail:.0f} GiB of host memory available, {need:.0f} GiB state"
    ops = random_layered(n, depth, seed=4242)
    t0 = time.perf_counter()
    sim = HostStateSimulation(n, m, dtype="complex64", contiguous_bits=k, buffers=3)
    t_alloc = time.perf_counter() - t0
    plan = sim.plan(ops)
    t0 = time.perf_counter()
    sim.evolve(ops)
    dt = time.perf_counter() - t0
    moved = 2.0 * need * 2**30 * (len(plan) - 1) + 2.0 * (1 << m) * 8  # first pass: slice 0 only
    log(f"host tier: {n} q c64 through {m} q slices, {len(plan)} passes in {dt:.1f} s")
    return {
        "config": f"random layered circuit, {n} qubits, depth {depth}, complex64 ({need:.0f} GiB "
                  f"in pinned host memory), {m}-qubit device slices gathered in chunks of 2^{k} "
                  f"amplitudes, 3 device buffers",
        "passes": len(plan), "gates": len(ops), "seconds": dt, "alloc_init_seconds": t_alloc,
        "pcie_GBps": moved / dt / 1e9,
        "host_mem_total_GiB": total, "host_mem_available_GiB": avail,
    }


def run_qaoa_bench(C, reps=2):
    """Config 4: 256 QAOA MaxCut instances, qubits uniform over the even values
    12..26 (3-regular graphs need an even count), p = 4, complex64 with FP64
    reductions; energies + adjoint gradients of the whole batch per run."""
    from snuqs_amd.qaoa import make_batch, qaoa_batch

    nq, graphs, gam, bet = make_batch(256, p=4, seed=7)
    batch = qaoa_batch(nq, graphs, 4, dtype="complex64")
    t0 = time.perf_counter()
    out = batch.run(gam, bet, use_graph=True)  # captures the graph
    first = time.perf_counter() - t0
    C.core.cuda.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = batch.run(gam, bet, use_graph=True)
    dt = (time.perf_counter() - t0) / reps
    st = batch.stats()
    log(f"qaoa batch: {dt * 1e3:.0f} ms per run ({first * 1e3:.0f} ms first run incl. capture)")
    res = {
        "config": "256 QAOA MaxCut instances, qubits uniform over {12,14,...,26}, random "
                  "3-regular graphs, p=4, complex64, energies + 8-parameter adjoint gradients",
        "ms_per_batch": dt * 1e3,
        "instances_per_s": 256 / dt,
        "launches_per_batch": st["launches"],
        "hbm_GBps": st["traffic_bytes"] / dt / 1e9,
        "state_GiB": st["state_bytes"] / 2**30,
        "energy_mean": float(out[:, 0].mean()),
    }
    del batch
    return res


def cpu_baseline(qubits, depth, seconds):
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--qubits", str(qubits), "--depth",
           str(depth), "--seconds", str(seconds)]
    log("cpu baseline:", " ".join(cmd))
    p = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        log("cpu baseline failed:", p.stderr[-2000:])
        return None
    return json.loads(p.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tune-reps", type=int, default=2,
                    help="timed program runs per candidate shape when tuning")
    ap.add_argument("--no-tune", action="store_true",
                    help="run with the default launch shapes (no per-gate tuning)")
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=30)
    ap.add_argument("--dtype", default="complex64")
    ap.add_argument("--skip-unfused", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-qaoa", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-cbatch", action="store_true")
    ap.add_argument("--skip-host-tier", action="store_true")
    ap.add_argument("--skip-configs", action="store_true",
                    help="skip the config 1, config 2 complex128 and config 5 legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--gate-times", default=None, metavar="PATH",
                    help="write the fused circuit's per-gate times (HIP events, mean over the "
                         "timed steps) with each gate's target bits and class to PATH as JSON")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the state-vector engine runs on a single MI355X "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from snuqs_amd import _C as C
    from snuqs_amd.circuits import random_layered
    from snuqs_amd.transpile import fuse_operations

    n, dtype = args.qubits, args.dtype
    item = 8 if dtype == "complex64" else 16
    log(f"building circuit: {n} qubits, depth {args.depth}")
    ops = random_layered(n, args.depth, seed=1234)
    fused = fuse_operations(ops, 5)
    log(f"{len(ops)} gates fused into {len(fused)} blocks")
    prog = C.functionals.Program(n, fused, dtype=dtype)
    sv = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
    tuning = None
    if not args.no_tune:
        # per-gate launch shapes of the 5-target MFMA kernels, chosen by timing
        # on this state before the timed region (the state is reset after)
        t0 = time.perf_counter()
        ntuned = prog.tune(sv, reps=args.tune_reps)
        tuning = {"tuned_gates": int(ntuned), "seconds": time.perf_counter() - t0}
        log(f"tuned {ntuned} gates in {tuning['seconds']:.1f} s")
    C.functionals.initialize_basis_z(sv)

    wall, acc_ms, info = sweep_stats(C, prog, sv, args.steps, args.warmup)
    step_bytes = sum(b for _, _, b in info)
    value = step_bytes * args.steps / wall / 1e9
    kernel_ms = float(acc_ms.sum())
    tab = class_table(acc_ms, info, args.steps)
    dom = max(tab.items(), key=lambda kv: kv[1]["ms"])
    (dcls, dkk), d = dom
    per_launch_bytes = d["bytes"] / d["launches"]
    avg_ms = d["ms"] / d["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    # MFMA utilisation of the dominant fused kernel: 8 * 2^k real flops per
    # amplitude for the complex 2^k x 2^k matrix-vector product; the kernels
    # run it as three real products (Re, Im, Re + Im), 6 * 2^k flops executed
    flops = 8.0 * (1 << dkk) * (1 << n) if dcls == 5 else 0.0
    flops_exec = 0.75 * flops
    norm = C.functionals.norm2(sv)
    log(f"fused: {wall / args.steps * 1e3:.1f} ms/step, {value:.0f} GB/s, norm^2 {norm:.12f}")
    if args.gate_times:
        rows = []
        for g, (op, (cls, kk, b)) in enumerate(zip(fused, info)):
            ms = float(acc_ms[g]) / args.steps
            rows.append({"gate": g, "bits": sorted(n - 1 - q for q in op.targets), "class": int(cls),
                         "kk": int(kk), "ms": ms, "GBps": b / (ms * 1e-3) / 1e9 if ms > 0 else None})
        with open(args.gate_times, "w") as f:
            json.dump({"qubits": n, "dtype": dtype, "steps": args.steps, "gates": rows}, f)
        log(f"per-gate times written to {args.gate_times}")

    # copy-kernel ceiling on the same state size
    tmp = C.result_types.StateVector(C.DeviceType.CUDA, n, dtype=dtype)
    C.functionals.copy_kernel(tmp, sv)
    C.core.cuda.device_synchronize()
    ev0, ev1 = C.event.Event.create(), C.event.Event.create()
    ev0.record()
    for _ in range(5):
        C.functionals.copy_kernel(tmp, sv)
    ev1.record()
    ev1.synchronize()
    copy_gbs = 5 * 2 * (1 << n) * item / (ev0.elapsed_ms(ev1) * 1e-3) / 1e9
    del tmp

    unfused = None
    if not args.skip_unfused:
        prog_u = C.functionals.Program(n, ops, dtype=dtype)
        C.functionals.initialize_basis_z(sv)
        wall_u, acc_u, info_u = sweep_stats(C, prog_u, sv, 1, 0)
        tab_u = class_table(acc_u, info_u, 1)
        bytes_u = sum(b for _, _, b in info_u)
        unfused = {
            "ms_per_step": wall_u * 1e3,
            "GBps": bytes_u / wall_u / 1e9,
            # full-sweep convention (2 N s per launched kernel, touched or not)
            "GBps_full_sweep": sum(2.0 * (1 << n) * item for c, _, _ in info_u if c) / wall_u / 1e9,
            "launches": int(sum(1 for c, _, _ in info_u if c)),
            "kernels": {f"{CLASS_NAMES[c]}_k{k}": {
                "launches": t["launches"], "avg_ms": t["ms"] / t["launches"],
                "GBps": t["bytes"] / (t["ms"] * 1e-3) / 1e9} for (c, k), t in tab_u.items()},
        }
        log(f"unfused: {wall_u * 1e3:.0f} ms/step, {unfused['GBps']:.0f} GB/s")
        del prog_u
    del sv, prog

    adj = None if args.skip_adjoint else run_adjoint_bench(C)
    others = None if args.skip_configs else run_other_configs(C, args.depth)
    qaoa = None if args.skip_qaoa else run_qaoa_bench(C)
    batch = None if args.skip_batch else run_batch_bench(C)
    cbatch = None if args.skip_cbatch else run_cbatch_bench(C)
    host_tier = None if args.skip_host_tier else run_host_tier_bench(C)
    cpu = None if args.skip_cpu else cpu_baseline(n, args.depth, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if dtype == "complex64" else "c128",
        "data": "synthetic",
        "config": {
            "workload": f"random layered circuit, {n} qubits, depth {args.depth}, {dtype}: "
                        f"seeded qubit pairing, Haar-random 4x4 Unitary per pair, RotZ U[0,2pi) "
                        f"on every qubit; {len(ops)} gates fused into {len(fused)} <=5-qubit blocks",
            "qubits": n, "depth": args.depth, "gates": len(ops), "kernels_per_step": len(fused),
            "algorithmic_bytes_per_step": step_bytes, "parallelism": "single GPU",
            "launch_shapes": tuning if tuning else "defaults",
        },
        "roofline": {
            "bound": "hbm", "kernel": kernel_name(dcls, dkk, dtype),
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (pmc_traffic(kernel_name(dcls, dkk, dtype).split("/")[0])
                        if n == 30 else None),
            "algorithmic_bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_ms,
            "launches": d["launches"],
            "mfma_tflops": flops / (avg_ms * 1e-3) / 1e12 if flops else None,
            "mfma_frac_fp32": flops / (avg_ms * 1e-3) / 1e12 / FP32_MFMA_PEAK_TFLOPS if flops else None,
            "mfma_tflops_executed": flops_exec / (avg_ms * 1e-3) / 1e12 if flops else None,
            "mfma_frac_fp32_executed": (flops_exec / (avg_ms * 1e-3) / 1e12 / FP32_MFMA_PEAK_TFLOPS
                                        if flops else None),
        },
        "kernel_ms_per_step": kernel_ms / args.steps,
        # the same bytes over the summed kernel times (HIP events), without launch gaps
        "GBps_kernel_only": step_bytes * args.steps / (kernel_ms * 1e-3) / 1e9,
        "copy_kernel_GBps": copy_gbs,
        "frac_of_copy": value / copy_gbs,
        "kernels": {f"{CLASS_NAMES[c]}_k{k}": {
            "launches": t["launches"], "avg_ms": t["ms"] / t["launches"],
            "GBps": t["bytes"] / (t["ms"] * 1e-3) / 1e9} for (c, k), t in tab.items()},
        "unfused": unfused,
        "adjoint": adj,
        "qaoa_batch": qaoa,
        "parameter_batch": batch,
        "circuit_batch": cbatch,
        "host_tier": host_tier,
        "other_configs": others,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
