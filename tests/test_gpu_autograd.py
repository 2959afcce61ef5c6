"""torch.autograd wrappers over the adjoint gradients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_circuit_energy_backward_matches_parameter_shift(C):
    import torch

    from snuqs_amd.autograd import circuit_energy
    from snuqs_amd.circuits import hea_ansatz, random_pauli_sum
    from oracle import oracle as O

    n, layers = 6, 2
    rng = np.random.default_rng(0)
    obs = random_pauli_sum(n, 8, 2, rng)
    builder = lambda th: hea_ansatz(n, layers, th)  # noqa: E731
    th0 = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    theta = torch.tensor(th0, requires_grad=True)
    e = circuit_energy(theta, builder, obs)
    e.backward()

    def energy(th):
        st = O.basis_state(n)
        for op in builder(th)[0]:
            O.apply(st, np.array(op.matrix()), op.targets, n)
        return sum(c * O.expect_pauli(st, p, q, n).real for p, q, c in zip(*obs))

    assert abs(e.item() - energy(th0)) < 1e-12
    ps = np.array([0.5 * (energy(th0 + np.pi / 2 * np.eye(th0.size)[j])
                          - energy(th0 - np.pi / 2 * np.eye(th0.size)[j]))
                   for j in range(th0.size)])
    assert np.abs(theta.grad.numpy() - ps).max() < 1e-9


def test_qaoa_energies_backward(C):
    import torch

    from snuqs_amd.autograd import qaoa_energies
    from snuqs_amd.qaoa import make_batch, qaoa_batch

    nq, graphs, gam, bet = make_batch(6, sizes=(6, 8, 12, 14), p=2, seed=3)
    batch = qaoa_batch(nq, graphs, 2, dtype="complex128")
    g = torch.tensor(gam, requires_grad=True)
    b = torch.tensor(bet, requires_grad=True)
    w = torch.arange(1.0, 7.0, dtype=torch.float64)
    e = qaoa_energies(g, b, batch)
    (e * w).sum().backward()
    out = batch.run(gam, bet)
    assert np.allclose(e.detach().numpy(), out[:, 0])
    assert np.allclose(g.grad.numpy(), w.numpy()[:, None] * out[:, 1:3])
    assert np.allclose(b.grad.numpy(), w.numpy()[:, None] * out[:, 3:5])
