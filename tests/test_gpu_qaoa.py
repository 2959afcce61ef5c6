"""Batched QAOA (uneven qubit counts, one launch sequence for the batch) vs the
FP64 CPU oracle: energies, and gradients vs central finite differences of the
oracle energy; hipGraph replays vs eager runs."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_energy(n, edges, gam, bet):
    from snuqs_amd.circuits import qaoa_maxcut

    ops, _ = qaoa_maxcut(n, edges, gam, bet)
    st = O.basis_state(n)
    for op in ops:
        O.apply(st, np.array(op.matrix()), op.targets, n)
    e = 0.0
    for a, b in edges:
        e += 0.5 * (1.0 - O.expect_pauli(st, "ZZ", [a, b], n).real)
    return e


def _graph(n, rng):
    """Seeded 3-regular graph for even n; a ring plus n // 2 random chords for
    odd n (no 3-regular graph exists on an odd vertex count)."""
    from snuqs_amd.circuits import random_3_regular

    if n % 2 == 0:
        return random_3_regular(n, rng)
    edges = {(i, (i + 1) % n) if i < (i + 1) % n else ((i + 1) % n, i) for i in range(n)}
    while len(edges) < n + n // 2:
        a, b = sorted(int(x) for x in rng.choice(n, 2, replace=False))
        edges.add((a, b))
    return sorted(edges)


def _batch(sizes, p, seed):
    rng = np.random.default_rng(seed)
    graphs = [_graph(n, rng) for n in sizes]
    gam = rng.uniform(0, np.pi, (len(sizes), p))
    bet = rng.uniform(0, np.pi / 2, (len(sizes), p))
    return graphs, gam, bet


@pytest.mark.parametrize("dtype,etol,gtol", [("complex128", 1e-10, 1e-7),
                                             ("complex64", 2e-5, 2e-4)])
def test_qaoa_batch_energies_and_gradients(C, dtype, etol, gtol):
    from snuqs_amd.qaoa import qaoa_batch

    sizes = [4, 6, 12, 8, 14, 6, 16, 13, 15, 17]  # below, at and above the 2^12 chunk
    p = 3
    graphs, gam, bet = _batch(sizes, p, 1)
    batch = qaoa_batch(sizes, graphs, p, dtype=dtype)
    out = batch.run(gam, bet, use_graph=False)
    assert out.shape == (len(sizes), 1 + 2 * p)
    h = 1e-5
    for j, n in enumerate(sizes):
        e_ref = oracle_energy(n, graphs[j], gam[j], bet[j])
        assert abs(out[j, 0] - e_ref) < etol * max(1.0, abs(e_ref)), (j, n)
        if n > 12:
            continue  # finite differences of the CPU oracle: keep to small n
        fd = np.empty(2 * p)
        for k in range(2 * p):
            gp, bp = gam[j].copy(), bet[j].copy()
            gm, bm = gam[j].copy(), bet[j].copy()
            if k < p:
                gp[k] += h
                gm[k] -= h
            else:
                bp[k - p] += h
                bm[k - p] -= h
            fd[k] = (oracle_energy(n, graphs[j], gp, bp)
                     - oracle_energy(n, graphs[j], gm, bm)) / (2 * h)
        assert np.abs(out[j, 1:] - fd).max() < gtol, (j, n, out[j, 1:], fd)


def test_qaoa_gradients_match_general_adjoint(C):
    """The batch kernels against the general engine's adjoint on the same
    circuit (ZZ(2 gamma): dE/dgamma = 2 * sum of the ZZ-angle derivatives).
    Instances of n >= 12 qubits are swept in 7-bit windows above bits 0-4
    (bits 5-11, 12-18, 19-25, 26-29): 12 has one window, 13-19 two, 20-26
    three and 27 four, and the sizes give the last window every width 1-7;
    with p = 3 every kind of pass (first, window-only, cost boundary, turn)
    runs at every window count; 26 is the largest size of the benchmark
    batch."""
    from snuqs_amd.circuits import qaoa_maxcut
    from snuqs_amd.qaoa import qaoa_batch

    sizes = [14, 10, 17, 19, 13, 15, 20, 16, 24, 26, 12, 18, 27]
    p = 3
    graphs, gam, bet = _batch(sizes, p, 5)
    out = qaoa_batch(sizes, graphs, p, dtype="complex128").run(gam, bet, use_graph=False)
    for j, n in enumerate(sizes):
        ops, pmap = qaoa_maxcut(n, graphs[j], gam[j], bet[j])
        prog = C.functionals.Program(n, ops, dtype="complex128")
        st = [C.result_types.StateVector(C.DeviceType.CUDA, n) for _ in range(3)]
        paulis = ["ZZ"] * len(graphs[j])
        qubits = [[a, b] for a, b in graphs[j]]
        coeffs = [-0.5] * len(graphs[j])
        e, g = prog.adjoint(st[0], st[1], st[2], [], pmap, 2 * p, paulis, qubits, coeffs)
        e += 0.5 * len(graphs[j])
        assert abs(out[j, 0] - e) < 1e-10
        assert np.abs(out[j, 1:] - 2.0 * g).max() < 1e-9


def test_qaoa_complex64_batch_matches_complex128_at_benchmark_sizes(C):
    """complex64 states with FP64 reductions against the complex128 batch on
    3-regular graphs of the benchmark's sizes (p = 4): energies within 1e-5
    relative; gradients, sums of products of lambda = C psi (|C| up to |E|)
    and S psi (|S| up to |E|), within 1e-5 of their term scale |E|^2."""
    from snuqs_amd.circuits import random_3_regular
    from snuqs_amd.qaoa import qaoa_batch

    rng = np.random.default_rng(21)
    sizes = [26, 22, 12, 24, 18]
    p = 4
    graphs = [random_3_regular(n, rng) for n in sizes]
    gam = rng.uniform(0, np.pi, (len(sizes), p))
    bet = rng.uniform(0, np.pi / 2, (len(sizes), p))
    ref = qaoa_batch(sizes, graphs, p, dtype="complex128").run(gam, bet, use_graph=False)
    got = qaoa_batch(sizes, graphs, p, dtype="complex64").run(gam, bet, use_graph=True)
    assert np.all(np.abs(got[:, 0] - ref[:, 0]) <= 1e-5 * np.abs(ref[:, 0])), (got[:, 0], ref[:, 0])
    scale = np.array([len(g) ** 2 for g in graphs], dtype=float)[:, None]
    gerr = np.abs(got[:, 1:] - ref[:, 1:]) / scale
    print("complex64 energy rel err", np.abs(got[:, 0] / ref[:, 0] - 1).max(), "grad err / |E|^2", gerr.max())
    assert gerr.max() < 1e-5, gerr.max()


def test_qaoa_graph_replays_match_eager(C):
    from snuqs_amd.qaoa import qaoa_batch

    sizes = [12, 14, 6, 16, 8]
    p = 2
    graphs, gam, bet = _batch(sizes, p, 9)
    batch = qaoa_batch(sizes, graphs, p)
    rng = np.random.default_rng(0)
    for r in range(4):
        g = gam + rng.uniform(-0.1, 0.1, gam.shape)
        b = bet + rng.uniform(-0.1, 0.1, bet.shape)
        eager = batch.run(g, b, use_graph=False)
        C.core.cuda.device_synchronize()
        replay = batch.run(g, b, use_graph=True)
        assert np.array_equal(eager, replay), r
    st = batch.stats()
    assert st["launches"] > 1 and st["state_bytes"] > 0


def test_qaoa_rejects_repeated_edges(C):
    """MaxCut graphs are simple: an edge given twice (either order) is an
    invalid argument, reported before anything is allocated."""
    from snuqs_amd.qaoa import qaoa_batch

    with pytest.raises(Exception, match="repeated edge"):
        qaoa_batch([14, 6], [[(0, 1), (2, 3)], [(0, 1), (1, 0)]], 1)
    qaoa_batch([14, 6], [[(0, 1), (2, 3)], [(0, 1), (1, 2)]], 1)


def test_qaoa_largest_instances_vs_general_adjoint(C):
    """The largest qubit counts a batch accepts: 30 qubits (four windows, the
    last 4 bits wide, amplitude offsets up to 2^30 per state) and 28 (last
    window 2 bits) beside a small instance, complex128 against the general
    adjoint on the same circuit, and complex64 against complex128."""
    import gc

    from snuqs_amd.circuits import qaoa_maxcut, random_3_regular
    from snuqs_amd.qaoa import qaoa_batch

    if C.core.cuda.mem_info()[0] / 2**30 < 80:
        pytest.skip("needs 80 GiB of free device memory")
    rng = np.random.default_rng(30)
    sizes = [30, 14, 28]
    p = 2
    graphs = [random_3_regular(n, rng) for n in sizes]
    gam = rng.uniform(0, np.pi, (len(sizes), p))
    bet = rng.uniform(0, np.pi / 2, (len(sizes), p))
    out = {}
    for dtype in ("complex128", "complex64"):
        batch = qaoa_batch(sizes, graphs, p, dtype=dtype)
        out[dtype] = batch.run(gam, bet, use_graph=False)
        del batch
        gc.collect()
    o128, o64 = out["complex128"], out["complex64"]
    for j, n in enumerate(sizes):
        ops, pmap = qaoa_maxcut(n, graphs[j], gam[j], bet[j])
        prog = C.functionals.Program(n, ops, dtype="complex128")
        st = [C.result_types.StateVector(C.DeviceType.CUDA, n) for _ in range(3)]
        paulis = ["ZZ"] * len(graphs[j])
        qubits = [[a, b] for a, b in graphs[j]]
        coeffs = [-0.5] * len(graphs[j])
        e, g = prog.adjoint(st[0], st[1], st[2], [], pmap, 2 * p, paulis, qubits, coeffs)
        e += 0.5 * len(graphs[j])
        assert abs(o128[j, 0] - e) < 1e-10, (j, n)
        assert np.abs(o128[j, 1:] - 2.0 * g).max() < 1e-8, (j, n)
        del prog, st
        gc.collect()
    scale = np.array([len(g) ** 2 for g in graphs], dtype=float)[:, None]
    assert np.all(np.abs(o64[:, 0] - o128[:, 0]) <= 1e-5 * np.abs(o128[:, 0]))
    assert (np.abs(o64[:, 1:] - o128[:, 1:]) / scale).max() < 1e-5
