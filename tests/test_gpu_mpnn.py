"""GPU parity of the learned preconditioner (BF16 MFMA kernels through the
C ABI) with the numpy oracle, and its use in the Chorin pressure solve.

The oracle rounds to bfloat16 at the same points as the kernels; what is
left is FP32-vs-FP64 accumulation order, which now and then moves a stored
bfloat16 by one unit in the last place (2^-8 relative).  Bars: forward and
input gradient within 1e-3 relative L2 (and 1e-2 max-norm), parameter
gradients within 5e-3 relative L2 per tensor, Adam within 1e-6 of the
oracle's step applied to the same gradient; learned V-cycle 1e-3, its
parameter gradient 1e-2.  Pressure solves with the learned preconditioner
reach the 1e-8 FP64 relative residual.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import chorin as C  # noqa: E402
from oracle import mpnn as M  # noqa: E402


def _operator(nx, ny, seed=0, dx=0.0625):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx, indexing="ij")
    h = 1 + 0.2 * np.sin(2 * X) * np.cos(Y) + 0.05 * rng.random((nx, ny))
    b = 0.1 * np.cos(X + 2 * Y)
    # FP32 stencil as the device holds it
    coef = tuple(np.asarray(a, np.float32).astype(np.float64)
                 for a in C.coefficients(h, b, dx, dx))
    return h, b, coef, rng


def _t32(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32,
                           device="cuda")


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _net(seed=3, zero_out=False):
    from zoomy_hip.learned import MPPreconditioner
    flat = M.init_params(seed, zero_out=zero_out)
    return MPPreconditioner(params=flat), flat.astype(np.float64)


@pytest.mark.parametrize("nx,ny,B", [(32, 32, 1), (16, 24, 2)])
def test_forward_backward_parity(nx, ny, B):
    net, flat = _net()
    zs, grads, drs = [], [], []
    ops = [_operator(nx, ny, seed=k) for k in range(B)]
    rs = [op[3].standard_normal((nx, ny)) for op in ops]
    ws = [op[3].standard_normal((nx, ny)) for op in ops]
    G = np.zeros(M.N_PARAMS)
    for (h, b, coef, _), r, w in zip(ops, rs, ws):
        phi = M.features(*coef)
        rr = np.asarray(r, np.float32).astype(np.float64)
        z, cache = M.forward(flat, rr, coef[0], phi)
        Gk, dr = M.backward(flat, cache, np.asarray(w, np.float32), coef[0])
        zs.append(z)
        drs.append(dr)
        G += Gk
    stack = lambda arrs: _t32(np.stack(arrs))  # noqa: E731
    diag = stack([op[2][0] for op in ops])
    cE = stack([op[2][1] for op in ops])
    cN = stack([op[2][2] for op in ops])
    net.zero_grad()
    z = net.forward(stack(rs), diag, cE, cN, slot=0).cpu().numpy()
    dr = net.backward(stack(ws), diag, slot=0).cpu().numpy()
    Gg = net.get_grad().astype(np.float64)
    zr = np.stack(zs)
    assert _rel(z, zr) < 1e-3
    assert np.abs(z - zr).max() < 1e-2 * np.abs(zr).max()
    assert _rel(dr, np.stack(drs)) < 1e-3
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        assert _rel(Gg[o:o + k], G[o:o + k]) < 5e-3, name
        o += k


@pytest.mark.parametrize("nx,ny,B", [(16, 24, 2), (96, 80, 1), (64, 64, 1)])
def test_fused_inference_equals_layer_by_layer(nx, ny, B):
    net, flat = _net(11)
    ops = [_operator(nx, ny, seed=k + 3) for k in range(B)]
    stack = lambda arrs: _t32(np.stack(arrs))  # noqa: E731
    r = stack([op[3].standard_normal((nx, ny)) for op in ops])
    diag = stack([op[2][0] for op in ops])
    cE = stack([op[2][1] for op in ops])
    cN = stack([op[2][2] for op in ops])
    z_layers = net.forward(r, diag, cE, cN, slot=0).cpu().numpy()
    net.set_fused(True)
    z_fused = net.forward(r, diag, cE, cN, slot=2).cpu().numpy()
    np.testing.assert_array_equal(z_fused, z_layers)
    base = torch.ones_like(r)
    out = net.forward(r, diag, cE, cN, slot=2, z=base, add=True).cpu().numpy()
    np.testing.assert_array_equal(out, z_layers + 1.0)


def test_adam_step_parity():
    net, flat = _net(5)
    h, b, coef, rng = _operator(16, 16, 1)
    r = rng.standard_normal((16, 16))
    w = rng.standard_normal((16, 16))
    d, e, n = (_t32(a) for a in coef)
    net.zero_grad()
    net.forward(_t32(r), d, e, n, slot=0)
    net.backward(_t32(w), d, slot=0, want_dr=False)
    g = net.get_grad().astype(np.float64)
    p0 = net.get_params().astype(np.float64)
    net.adam(lr=1e-3)
    p1 = net.get_params().astype(np.float64)
    ref, _, _ = M.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, lr=1e-3)
    assert np.abs(p1 - ref).max() < 1e-6


def _chorin(n, precond="learned", **kw):
    from zoomy_hip.fvm import ChorinSplitVAMSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    box = (0.0, 0.0625 * n, 0.0, 0.0625 * n)
    mesh = Mesh.quad(n, n, box).set_bc("wall")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll")
    s = ChorinSplitVAMSolver(model, precond=precond, **kw)
    s.setup_simulation(mesh, n, n)
    Q, _ = C.solitary_wave_state(n, n, box)
    return s, Q, box[1] / n


def test_learned_vcycle_and_gradient_parity():
    n = 32
    s, Q, dx = _chorin(n)
    net, flat = _net(7)
    s.set_preconditioner("learned", net)
    Qd = torch.as_tensor(Q, device="cuda")
    s.setup_operator(Qd[0], Qd[1])
    h, b = Q[1].reshape(n, n), Q[0].reshape(n, n)
    H = C.Hierarchy(h, b, dx, dx)
    for L in H.levels:  # the FP32 hierarchy of the device
        L["coef"] = tuple(np.asarray(a, np.float32).astype(np.float64)
                          for a in L["coef"])
    rng = np.random.default_rng(2)
    r = rng.standard_normal(n * n).astype(np.float32)
    g = rng.standard_normal(n * n).astype(np.float32)
    z_ref, caches = M.learned_vcycle(flat, H, r.reshape(n, n).astype(float))
    G_ref = M.learned_vcycle_backward(flat, H, caches,
                                      g.reshape(n, n).astype(float))
    net.zero_grad()
    z = s.learned_vcycle(_t32(r), train=True).cpu().numpy()
    s.learned_vcycle_backward(_t32(g))
    G = net.get_grad().astype(np.float64)
    assert _rel(z, z_ref.ravel()) < 1e-3
    assert _rel(G, G_ref) < 1e-2


def test_learned_solve_reaches_target_and_falls_back():
    import scipy.sparse.linalg as spla
    n = 128
    s, Q, dx = _chorin(n)
    net, _ = _net(1, zero_out=False)
    s.set_preconditioner("learned", net)
    Qd = torch.as_tensor(Q, device="cuda")
    s.setup_operator(Qd[0], Qd[1])
    f = np.random.default_rng(3).standard_normal(n * n)
    p = s.solve(torch.as_tensor(f, device="cuda")).cpu().numpy()
    st = s.stats()
    assert st["fallbacks"] == 0 and st["relres"] <= 1e-8
    assert st["lp_solves"] == 1 and st["lp_iterations"] > 0
    coef = C.coefficients(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    res = f - C.apply(coef, p.reshape(n, n)).ravel()
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f)
    pd = spla.spsolve(C.to_sparse(coef).tocsc(), f)
    assert np.linalg.norm(p - pd) <= 1e-6 * np.linalg.norm(pd)
    # a broken network (its products overflow to inf and NaN): the fresh
    # FP64 residual misses the target and the solve is redone plainly
    from zoomy_hip.learned import MPPreconditioner
    bad = MPPreconditioner(params=np.full(M.N_PARAMS, 1e30, np.float32))
    s.set_preconditioner("learned", bad)
    s.setup_operator(Qd[0], Qd[1])
    p2 = s.solve(torch.as_tensor(f, device="cuda")).cpu().numpy()
    st2 = s.stats()
    assert st2["fallbacks"] == 1 and st2["relres"] <= 1e-8
    assert not (st2["lp_relres"] <= 1e-8)
    assert np.linalg.norm(p2 - pd) <= 1e-6 * np.linalg.norm(pd)


def test_learned_guess_solve_and_retry():
    """The learned V-cycle as the initial guess of the FP64 V-cycle CG:
    the solve reaches 1e-8 (checked against a direct solve) in no more
    iterations than the plain CG from zero, and a broken network's guess
    is retried from zero."""
    import scipy.sparse.linalg as spla
    from zoomy_hip.learned import MPPreconditioner
    n = 128
    s, Q, dx = _chorin(n, precond="mg")
    Qd = torch.as_tensor(Q, device="cuda")
    f = np.random.default_rng(4).standard_normal(n * n)
    fd = torch.as_tensor(f, device="cuda")
    s.setup_operator(Qd[0], Qd[1])
    s.solve(fd)
    plain_its = s.stats()["iterations"]
    s.set_preconditioner("learned_guess", MPPreconditioner.load())
    s.setup_operator(Qd[0], Qd[1])
    p = s.solve(fd).cpu().numpy()
    st = s.stats()
    print("plain", plain_its, "learned guess", st)
    assert st["fallbacks"] == 0 and st["relres"] <= 1e-8
    assert st["lp_solves"] == 1 and 0 < st["iterations"] <= plain_its
    coef = C.coefficients(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    pd = spla.spsolve(C.to_sparse(coef).tocsc(), f)
    assert np.linalg.norm(p - pd) <= 1e-6 * np.linalg.norm(pd)
    bad = MPPreconditioner(params=np.full(M.N_PARAMS, 1e30, np.float32))
    s.set_preconditioner("learned_guess", bad)
    s.setup_operator(Qd[0], Qd[1])
    p2 = s.solve(fd).cpu().numpy()
    st2 = s.stats()
    assert st2["fallbacks"] == 1 and st2["relres"] <= 1e-8
    assert np.linalg.norm(p2 - pd) <= 1e-6 * np.linalg.norm(pd)


def test_training_reduces_vcycle_loss():
    from zoomy_hip.learned import MPPreconditioner, train, vcycle_loss
    n = 64
    s, Q, dx = _chorin(n)
    Qd = torch.as_tensor(Q, device="cuda")
    b, h = Qd[0].contiguous(), Qd[1].contiguous()
    s.set_preconditioner("learned", MPPreconditioner(seed=0))
    s.setup_operator(b, h)
    gen = torch.Generator(device="cpu").manual_seed(0)
    pairs = []
    for _ in range(8):
        f = torch.randn(n * n, generator=gen, dtype=torch.float64).cuda()
        s.set_preconditioner("mg")
        s.setup_operator(b, h)
        pairs.append((b, h, f, s.solve(f)))
    net = MPPreconditioner(seed=0)
    s.set_preconditioner("learned", net)
    before = np.mean([vcycle_loss(s, net, *pr, train=False) for pr in pairs])
    hist = train(net, pairs, s, epochs=8, batch=2, lr=3e-3)
    after = np.mean([vcycle_loss(s, net, *pr, train=False) for pr in pairs])
    assert np.isfinite(hist).all() and after < 0.8 * before


def test_trained_network_gain_over_the_vcycle_without_it():
    """What the shipped trained network gains, on a workload it never saw
    (512^2, the solitary wave over the seeded bed of the Chorin tests; the
    training used 256^2 grids and other beds): against the same learned
    V-cycle with the network switched off (output weights zero, so z = V r,
    the FP32 V-cycle of the mixed-precision solve), the energy-norm error of
    one V-cycle and the flexible-CG iterations of a 1e-8 solve."""
    from zoomy_hip.learned import MPPreconditioner, vcycle_loss
    n = 512
    s, Q, dx = _chorin(n)
    Qd = torch.as_tensor(Q, device="cuda")
    b, h = Qd[0].contiguous(), Qd[1].contiguous()
    gen = torch.Generator(device="cpu").manual_seed(11)
    f = torch.randn(n * n, generator=gen, dtype=torch.float64).cuda()
    s.set_preconditioner("mg")
    s.setup_operator(b, h)
    p = s.solve(f)
    nets = {"trained": MPPreconditioner.load(),
            "without": MPPreconditioner(params=M.init_params(3, zero_out=True))}
    loss, inner = {}, {}
    for k, net in nets.items():
        s.set_preconditioner("learned", net)
        loss[k] = vcycle_loss(s, net, b, h, f, p, train=False)
        s.setup_operator(b, h)
        s.solve(f)
        st = s.stats()
        assert st["relres"] <= 1e-8
        # a solve that needed the plain redo did not converge on its own
        inner[k] = st["lp_iterations"] if st["fallbacks"] == 0 else 10 ** 6
    assert inner["trained"] < 10 ** 6
    print(f"one V-cycle energy error: trained {loss['trained']:.4f}, "
          f"without the network {loss['without']:.4f}; FCG iterations "
          f"to 1e-8: trained {inner['trained']}, without {inner['without']}")
    assert loss["trained"] < 0.5 * loss["without"]
    assert inner["trained"] <= inner["without"]
