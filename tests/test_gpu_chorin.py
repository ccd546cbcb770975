"""GPU parity of the Chorin projection (through the C ABI) with the oracle.

Bars: FP64 operator, right-hand side, corrector and predictor within 1e-12
relative (max-norm) of the oracle; one FP64 V-cycle within 1e-12 (the same
sweeps in the same order), FP32 within 1e-5; pressure solves reach the 1e-8
relative residual measured by the oracle's own operator and agree with a
sparse direct solve to 1e-6; whole stages and cycles agree with the oracle's
direct-solve cycle to 1e-6 (the solve tolerance bounds the difference).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import chorin as C  # noqa: E402


def _case(n, seed=4, L_per_cell=0.0625):
    box = (0.0, L_per_cell * n, 0.0, L_per_cell * n)
    Q, hw = C.solitary_wave_state(n, n, box, seed=seed)
    return box, Q, hw


def _solver(n, box, precond="mg", fv_order="morton", **kw):
    from zoomy_hip.fvm import ChorinSplitVAMSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(n, n, box).set_bc("wall")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll")
    s = ChorinSplitVAMSolver(model, precond=precond, **kw)
    s.setup_simulation(mesh, n, n, fv_order=fv_order)
    return s, mesh


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _state5(Q, hw):
    return _dev(np.vstack([Q, hw[None]]))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_operator_rhs_correct_parity():
    n = 64
    box, Q, hw = _case(n)
    rng = np.random.default_rng(0)
    Q[2] += 0.05 * rng.standard_normal(n * n)
    Q[3] += 0.05 * rng.standard_normal(n * n)
    hw = 0.02 * rng.standard_normal(n * n)
    s, _ = _solver(n, box)
    dx = box[1] / n
    Qd = _state5(Q, hw)
    s.setup_operator(Qd[0], Qd[1])
    x = rng.standard_normal(n * n)
    y = s.apply(_dev(x)).cpu().numpy()
    coef = C.coefficients(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    ref = C.apply(coef, x.reshape(n, n)).ravel()
    assert _rel(y, ref) < 1e-12
    dt = 1e-3
    rhs = s.pressure_rhs(Qd, dt).cpu().numpy()
    ref = C.pressure_rhs(Q, hw, n, n, dx, dx, dt).ravel()
    assert _rel(rhs, ref) < 1e-12
    p = rng.standard_normal(n * n)
    out = s._step_corrector_pure(Qd, _dev(p), dt).cpu().numpy()
    Qn, hwn = C.correct(Q, hw, p.reshape(n, n), n, n, dx, dx, dt)
    assert _rel(out[2], Qn[2]) < 1e-12 and _rel(out[3], Qn[3]) < 1e-12
    assert _rel(out[4], hwn) < 1e-12
    np.testing.assert_array_equal(out[:2], Q[:2])
    Q0 = rng.standard_normal((5, n * n))
    out = s._step_corrector_pure(Qd, _dev(p), dt, _dev(Q0)).cpu().numpy()
    ref = 0.5 * (Q0 + np.vstack([Qn, hwn[None]]))
    assert _rel(out, ref) < 1e-12


@pytest.mark.parametrize("n", [32, 128])
def test_vcycle_parity(n):
    box, Q, _ = _case(n)
    s, _ = _solver(n, box)
    dx = box[1] / n
    Qd = _dev(Q)
    s.setup_operator(Qd[0], Qd[1])
    assert s.levels()[-1][0] == 4 and len(s.levels()) == int(np.log2(n)) - 1
    r = np.random.default_rng(n).standard_normal(n * n)
    z = s.vcycle(_dev(r)).cpu().numpy()
    H = C.Hierarchy(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    ref = C.vcycle(H, r.reshape(n, n)).ravel()
    assert _rel(z, ref) < 1e-12


@pytest.mark.parametrize("n,nu", [(64, 1), (256, 1), (96, 1), (64, 2),
                                  (1024, 1)])
def test_fused_vcycle_bit_identical_to_separate_kernels(n, nu):
    box, Q, _ = _case(n)
    s, _ = _solver(n, box, nu=nu)
    dx = box[1] / n
    Qd = _dev(Q)
    s.setup_operator(Qd[0], Qd[1])
    r = np.random.default_rng(n + nu).standard_normal(n * n)
    s.set_mg_fused(True)
    z1 = s.vcycle(_dev(r)).cpu().numpy()
    s.set_mg_fused(False)
    z0 = s.vcycle(_dev(r)).cpu().numpy()
    np.testing.assert_array_equal(z1, z0)
    H = C.Hierarchy(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    ref = C.vcycle(H, r.reshape(n, n), nu=nu).ravel()
    assert _rel(z1, ref) < 1e-12


def test_vcycle_fp32_parity():
    # the FP32 hierarchy (inner solves of the learned path) is built when
    # the solver is configured for the learned preconditioner
    n = 64
    box, Q, _ = _case(n)
    s, _ = _solver(n, box, precond="learned")
    dx = box[1] / n
    Qd = _dev(Q)
    s.setup_operator(Qd[0], Qd[1])
    r = np.random.default_rng(1).standard_normal(n * n)
    z32 = s.vcycle(torch.as_tensor(r, dtype=torch.float32, device="cuda"))
    H = C.Hierarchy(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    ref = C.vcycle(H, r.reshape(n, n)).ravel()
    assert _rel(z32.cpu().numpy().astype(np.float64), ref) < 1e-5


@pytest.mark.parametrize("n", [64, 256])
def test_mg_cg_solve(n):
    import scipy.sparse.linalg as spla
    box, Q, _ = _case(n)
    s, _ = _solver(n, box)
    dx = box[1] / n
    Qd = _dev(Q)
    s.setup_operator(Qd[0], Qd[1])
    f = np.random.default_rng(7).standard_normal(n * n)
    p = s.solve(_dev(f)).cpu().numpy()
    st = s.stats()
    coef = C.coefficients(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    res = f - C.apply(coef, p.reshape(n, n)).ravel()
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f)
    assert st["relres"] <= 1e-8 and st["fallbacks"] == 0
    _, it_ref, _ = C.mg_solve(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx,
                              f.reshape(n, n), tol=1e-8)
    assert abs(st["iterations"] - it_ref) <= 1
    pd = spla.spsolve(C.to_sparse(coef).tocsc(), f)
    assert np.linalg.norm(p - pd) <= 1e-6 * np.linalg.norm(pd)
    # zero right-hand side: zero iterations, zero pressure
    p0 = s.solve(torch.zeros(n * n, dtype=torch.float64, device="cuda"))
    assert s.stats()["iterations"] == 0 and float(p0.abs().max()) == 0.0


@pytest.mark.parametrize("n", [64, 256])
def test_mixed_precision_solve(n):
    """FP32 V-cycle inside FP64 flexible CG: the fresh FP64 residual reaches
    1e-8, the iteration count stays within one of the FP64 MG-CG's, the
    pressure agrees with a direct solve."""
    import scipy.sparse.linalg as spla
    box, Q, _ = _case(n)
    s, _ = _solver(n, box)
    dx = box[1] / n
    Qd = _dev(Q)
    f = np.random.default_rng(7).standard_normal(n * n)
    s.setup_operator(Qd[0], Qd[1])
    s.solve(_dev(f))
    it_plain = s.stats()["iterations"]
    s.set_preconditioner("mixed")
    s.setup_operator(Qd[0], Qd[1])
    p = s.solve(_dev(f)).cpu().numpy()
    st = s.stats()
    assert st["fallbacks"] == 0 and st["lp_solves"] == 1
    assert st["relres"] <= 1e-8 and st["lp_relres"] == st["relres"]
    assert abs(st["iterations"] - it_plain) <= 1
    coef = C.coefficients(Q[1].reshape(n, n), Q[0].reshape(n, n), dx, dx)
    res = f - C.apply(coef, p.reshape(n, n)).ravel()
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f)
    pd = spla.spsolve(C.to_sparse(coef).tocsc(), f)
    assert np.linalg.norm(p - pd) <= 1e-6 * np.linalg.norm(pd)


def test_elliptic_solve_error_when_capped():
    from zoomy_hip.fvm import EllipticSolveError
    n = 64
    box, Q, _ = _case(n)
    s, _ = _solver(n, box, maxiter=1)
    Qd = _dev(Q)
    s.setup_operator(Qd[0], Qd[1])
    f = np.random.default_rng(8).standard_normal(n * n)
    with pytest.raises(EllipticSolveError):
        s.solve(_dev(f))


def _oracle_chorin(n, box):
    from oracle.oracle import Oracle
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(n, n, box).set_bc("wall")
    fv = Oracle(mesh, order=2, limiter="minmod", riemann="hll")
    return C.ChorinOracle(fv, n, n, box)


def test_predictor_parity():
    n = 64
    box, Q, hw = _case(n)
    hw = 0.02 * np.random.default_rng(9).standard_normal(n * n)
    s, _ = _solver(n, box)
    co = _oracle_chorin(n, box)
    dt = 0.45 * (box[1] / n) / (np.sqrt(9.81) + 1)
    out = s._hyperbolic_stage(dt, 0.0, _state5(Q, hw)).cpu().numpy()
    Qs, hws = co.predictor(Q, hw, dt)
    for k in range(4):
        assert _rel(out[k], Qs[k]) < 1e-12, k
    assert _rel(out[4], hws) < 1e-12


def test_mixed_precision_predictor():
    # FP32 fluxes, FP64 state: within the FP32 flux bar of the oracle
    n = 64
    box, Q, hw = _case(n)
    hw = 0.02 * np.random.default_rng(12).standard_normal(n * n)
    s, _ = _solver(n, box, flux_dtype=torch.float32)
    co = _oracle_chorin(n, box)
    dt = 0.45 * (box[1] / n) / (np.sqrt(9.81) + 1)
    out = s._hyperbolic_stage(dt, 0.0, _state5(Q, hw)).cpu().numpy()
    Qs, hws = co.predictor(Q, hw, dt)
    for k in (1, 2):
        inc = out[k] - Q[k]
        ref = Qs[k] - Q[k]
        assert np.abs(inc - ref).max() <= 1e-5 * np.abs(ref).max(), k
    assert _rel(out[4], hws) < 1e-12
    st = s.chorin_stage(dt, 0.0, _state5(Q, hw))
    assert s.stats()["relres"] <= 1e-8 and bool(torch.isfinite(st).all())


def test_predictor_same_in_grid_and_morton_order():
    n = 64
    box, Q, hw = _case(n)
    hw = 0.02 * np.random.default_rng(10).standard_normal(n * n)
    dt = 0.45 * (box[1] / n) / (np.sqrt(9.81) + 1)
    outs = []
    for order in ("grid", "morton"):
        s, _ = _solver(n, box, fv_order=order)
        outs.append(s._hyperbolic_stage(dt, 0.0, _state5(Q, hw)).cpu().numpy())
    # same arithmetic per cell; the vertex numbering of the reordered mesh
    # can change the last bit of a face geometry term
    scale = np.abs(outs[0][1:4]).max()
    assert np.abs(outs[1] - outs[0]).max() <= 1e-13 * scale


def test_stage_and_cycle_parity():
    n = 64
    box, Q, hw = _case(n)
    s, _ = _solver(n, box)
    co = _oracle_chorin(n, box)
    dt = 0.45 * (box[1] / n) / (np.sqrt(9.81) + 1)
    Qd = _state5(Q, hw)
    out = s.chorin_stage(dt, 0.0, Qd).cpu().numpy()
    Qn, hwn, p = co.stage(Q, hw, dt)
    ref = np.vstack([Qn, hwn[None]])
    for k in range(5):
        assert _rel(out[k], ref[k]) < 1e-6, k
    Q2 = Qd
    Qo, hwo = Q, hw
    for _ in range(3):
        Q2 = s.chorin_cycle(dt, 0.0, Q2)
        assert s.stats()["solves"] == 2
        Qo, hwo, _ = co.cycle(Qo, hwo, dt)
    out = Q2.cpu().numpy()
    ref = np.vstack([Qo, hwo[None]])
    for k in range(5):
        assert _rel(out[k], ref[k]) < 1e-6, k
    assert abs(out[1].sum() - Q[1].sum()) < 1e-12 * Q[1].sum()


def test_run_jit_steps_fixed_dt():
    from zoomy_hip.fvm.chorin import fixed_dt
    n = 128
    box, Q, hw = _case(n)
    s, _ = _solver(n, box)
    dt = fixed_dt(box[1] / n, 1.0)
    Qd, t = s.run_jit_steps(dt, 5, _state5(Q, hw))
    assert t == pytest.approx(5 * dt)
    out = Qd.cpu().numpy()
    assert np.isfinite(out).all()
    assert abs(out[1].sum() - Q[1].sum()) < 1e-12 * Q[1].sum()
    # the solitary wave moves in +x: momentum centre advances
    x = (np.arange(n) + 0.5) * box[1] / n
    X = np.repeat(x, n)
    c0 = (X * Q[2]).sum() / Q[2].sum()
    c1 = (X * out[2]).sum() / out[2].sum()
    assert c1 > c0
