"""Pins the learned-preconditioner oracle (oracle/mpnn.py).

The reference's training code (gnn_blueprint, jax/optax) is not runnable
here, so the oracle is pinned by calculus and libraries: the backward pass
against central finite differences in FP64, the bfloat16 rounding against
torch's round-to-nearest-even conversion, Adam against torch.optim.Adam, the
learned V-cycle's gradient against finite differences, and the V-cycle
reducing to the multigrid V-cycle structure when the network is zero.
"""
import numpy as np
import pytest

from oracle import chorin as C
from oracle import mpnn as M


def _case(n=16, seed=0):
    rng = np.random.default_rng(seed)
    dx = 0.0625
    X, Y = np.meshgrid(np.arange(n) * dx, np.arange(n) * dx, indexing="ij")
    h = 1 + 0.2 * np.sin(2 * X) * np.cos(Y) + 0.05 * rng.random((n, n))
    b = 0.1 * np.cos(X + 2 * Y)
    coef = C.coefficients(h, b, dx, dx)
    return h, b, dx, coef, rng


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(1).standard_normal(10000).astype(np.float32)
    x[:4] = [1.00390625, 1.01171875, -2.0 ** -130, 3.0e38]  # ties, subnormal
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16(x), ref)


def test_backward_matches_finite_differences():
    h, b, dx, coef, rng = _case(8)
    flat = M.init_params(3, zero_out=False).astype(np.float64)
    flat += 0.05 * rng.standard_normal(flat.size)
    phi = M.features(*coef)
    r = rng.standard_normal(h.shape)
    w = rng.standard_normal(h.shape)
    z, cache = M.forward(flat, r, coef[0], phi, bf=False)
    G, dr = M.backward(flat, cache, w, coef[0], bf=False)
    f = lambda p: np.sum(w * M.forward(p, r, coef[0], phi, bf=False)[0])
    idx = rng.choice(flat.size, 60, replace=False)
    idx = np.concatenate([idx, [0, 5, M.N_PARAMS - 1, M.N_PARAMS - 7]])
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = eps
        fd = (f(flat + e) - f(flat - e)) / (2 * eps)
        assert abs(fd - G[i]) <= 1e-6 * max(1.0, abs(fd)), i
    # input gradient (linear map: exact adjoint)
    v = rng.standard_normal(h.shape)
    zv, _ = M.forward(flat, v, coef[0], phi, bf=False)
    assert abs(np.sum(w * zv) - np.sum(dr * v)) < 1e-9 * abs(np.sum(w * zv))


def test_network_is_linear_in_the_residual():
    h, b, dx, coef, rng = _case(8, 2)
    flat = M.init_params(4, zero_out=False).astype(np.float64)
    phi = M.features(*coef)
    r1, r2 = rng.standard_normal((2,) + h.shape)
    z = lambda r: M.forward(flat, r, coef[0], phi, bf=False)[0]
    assert np.abs(z(2 * r1 - 3 * r2) - (2 * z(r1) - 3 * z(r2))).max() < 1e-12


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    p = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([p], lr=3e-3)
    flat, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 4):
        g = rng.standard_normal(50)
        p.grad = torch.tensor(g)
        opt.step()
        flat, m, v = M.adam(flat, g, m, v, t, lr=3e-3)
    np.testing.assert_allclose(flat, p.detach().numpy(), rtol=0, atol=1e-12)


def test_learned_vcycle_gradient_matches_finite_differences():
    h, b, dx, coef, rng = _case(16, 6)
    H = C.Hierarchy(h, b, dx, dx)
    flat = M.init_params(7, zero_out=False).astype(np.float64)
    r = rng.standard_normal(h.shape)
    g = rng.standard_normal(h.shape)
    z, caches = M.learned_vcycle(flat, H, r, bf=False)
    G = M.learned_vcycle_backward(flat, H, caches, g, bf=False)
    f = lambda p: np.sum(g * M.learned_vcycle(p, H, r, bf=False)[0])
    eps = 1e-6
    for i in rng.choice(flat.size, 25, replace=False):
        e = np.zeros_like(flat)
        e[i] = eps
        fd = (f(flat + e) - f(flat - e)) / (2 * eps)
        assert abs(fd - G[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_zero_network_vcycle_is_the_plain_vcycle():
    h, b, dx, coef, rng = _case(16, 8)
    H = C.Hierarchy(h, b, dx, dx)
    flat = M.init_params(0, zero_out=True)
    r = rng.standard_normal(h.shape)
    z, _ = M.learned_vcycle(flat, H, r, bf=True)
    np.testing.assert_array_equal(z, C.vcycle(H, r, level=0))
