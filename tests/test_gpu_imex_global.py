"""GPU parity of the global implicit source stage (Newton-GMRES, the
reference's _build_implicit_global) against the oracle's Newton with a dense
central-difference Jacobian, and of the IMEX step that uses it."""
import numpy as np
import pytest

from helpers import circular_dam_state, max_rel_rows

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(kind):
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    if kind == "quad":
        mesh = Mesh.quad(14, 12).set_bc("wall")
    elif kind == "tri":
        mesh = delaunay_mesh(jittered_points(12, 10, (0, 1, 0, 1), 0.3, 4),
                             (0, 1, 0, 1)).set_bc("wall")
    else:
        mesh = Mesh.line(60, 0.0, 1.0).set_bc("wall")
    Q = circular_dam_state(mesh, seed=6) if mesh.dim == 2 else None
    if Q is None:
        x = mesh.cell_centers()[:, 0]
        Q = np.zeros((4, mesh.n_cells))
        Q[1] = 1.0 + 0.3 * np.sin(2 * np.pi * x)
    rng = np.random.default_rng(8)
    Q[2] = rng.normal(0.0, 0.4, mesh.n_cells)
    if mesh.dim == 2:
        Q[3] = rng.normal(0.0, 0.4, mesh.n_cells)
    return mesh, Q


@pytest.mark.parametrize("kind", ["quad", "tri", "line"])
def test_global_stage_matches_oracle_newton(kind):
    """zj_imex_global_solve through the IMEX solver's global mode: the
    Newton-GMRES root equals the oracle's dense Newton root (relative
    1e-10), the residual is below the tolerance, and Newton converges in a
    few steps (exact Jacobian-vector products)."""
    from oracle import imex_global as IG
    from oracle.oracle import Oracle
    from zoomy_hip import _lib as L
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    import ctypes
    mesh, Q = _case(kind)
    alpha, n, dt = 0.5, 0.05, 0.02
    model = ShallowWater(dimension=mesh.dim, order=2, riemann="hll",
                         manning_n=n, depth_gradient_source=alpha)
    s = IMEXSourceSolver(time_end=np.inf, cfl=0.9)
    s.implicit_tol = 1e-11
    s.gmres_tol = 1e-10
    s.implicit_maxiter = 10
    s.setup_simulation(mesh, model)
    assert s.resolved_source_mode() == "global"
    o = Oracle(mesh, order=2, riemann="hll", manning_n=n)
    Qo, norms = IG.newton(o, Q, dt, alpha, n=n)
    Qe = torch.as_tensor(Q, device="cuda")
    out = torch.empty_like(Qe)
    rc = L.lib().zj_imex_global_solve(
        s._global, dt, ctypes.c_void_p(Qe.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), 10, 1e-11, 1e-10, 30,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    st = s.implicit_stats()
    print(kind, st, "oracle norms", norms, flush=True)
    assert st["rnorm"] <= 1e-11 and st["newton"] <= 5 and st["fallbacks"] == 0
    Qg = out.cpu().numpy()
    rows = (1, 2, 3) if mesh.dim == 2 else (1, 2)
    assert max_rel_rows(Qg, Qo, rows) < 1e-10
    # the depth-gradient term matters: the local stage (Manning only) gives
    # another state
    Ql = o.imex_local(dt, Q, maxiter=30)
    assert max_rel_rows(Ql, Qo, rows) > 1e-4


def test_imex_step_with_global_source_against_oracle():
    """Two IMEX steps (explicit MUSCL-HLL stage, then the global implicit
    source stage at the default tolerances) against the oracle's explicit
    step followed by its Newton root."""
    from oracle import imex_global as IG
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    mesh, Q = _case("quad")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll", manning_n=0.03,
                         depth_gradient_source=0.3)
    s = IMEXSourceSolver(time_end=np.inf, cfl=0.9)
    s.setup_simulation(mesh, model)
    oh = Oracle(mesh, order=2, limiter="minmod", riemann="hll", cfl=0.9)
    om = Oracle(mesh, order=2, riemann="hll", manning_n=0.03)
    Qo, Qg = Q.copy(), torch.as_tensor(Q, device="cuda")
    for _ in range(2):
        dt = oh.dt(Qo)
        Qe, _ = oh.step(dt, Qo)
        Qo, _ = IG.newton(om, Qe, dt, 0.3, n=0.03)
        Qg = s.step(dt, 0.0, Qg)
    # Newton stops at ||R|| <= 1e-8: the two roots agree to that level
    assert max_rel_rows(Qg.cpu().numpy(), Qo, (1, 2, 3)) < 1e-8
    with pytest.raises(NotImplementedError):
        from zoomy_hip.fvm import HyperbolicSolver
        HyperbolicSolver().setup_simulation(mesh, model)


def _global_solver(mesh, restart, alpha=0.5, n=0.05):
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    model = ShallowWater(dimension=mesh.dim, order=2, riemann="hll",
                         manning_n=n, depth_gradient_source=alpha)
    s = IMEXSourceSolver(time_end=np.inf, cfl=0.9)
    s.gmres_restart = restart
    s.setup_simulation(mesh, model)
    assert s.resolved_source_mode() == "global"
    return s


def _solve(s, Q, dt, maxiter, tol, gmres_tol, gmres_maxiter):
    import ctypes
    from zoomy_hip import _lib as L
    Qe = torch.as_tensor(Q, device="cuda")
    out = torch.empty_like(Qe)
    rc = L.lib().zj_imex_global_solve(
        s._global, dt, ctypes.c_void_p(Qe.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), maxiter, tol, gmres_tol,
        gmres_maxiter, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    return out.cpu().numpy(), s.implicit_stats()


@pytest.mark.parametrize("restart", [1, 2])
def test_restarted_gmres_matches_oracle_newton(restart):
    """GMRES(1) and GMRES(2) (a Newton step here takes 3 Krylov iterations
    unrestarted): every Newton step needs several restart
    cycles (the device restarts from the true residual -R - J dx and adds
    each cycle's correction to dx), and the root still equals the oracle's
    dense Newton root."""
    from oracle import imex_global as IG
    from oracle.oracle import Oracle
    mesh, Q = _case("quad")
    dt, alpha, n = 0.02, 0.5, 0.05
    s = _global_solver(mesh, restart, alpha, n)
    Qg, st = _solve(s, Q, dt, 10, 1e-11, 1e-10, 60)
    print("restart", restart, st, flush=True)
    assert st["fallbacks"] == 0 and st["rnorm"] <= 1e-11
    # more Krylov iterations than one cycle holds, in some Newton step
    assert st["gmres"] > restart * st["newton"]
    o = Oracle(mesh, order=2, riemann="hll", manning_n=n)
    Qo, _ = IG.newton(o, Q, dt, alpha, n=n)
    assert max_rel_rows(Qg, Qo, (1, 2, 3)) < 1e-10


def test_gmres_miss_falls_back_to_minus_residual():
    """One GMRES(1) cycle against a tolerance it cannot reach: every Newton
    step takes delta = -R instead (the reference's jnp.where(info == 0,
    delta, b)); the iteration Q <- Q - R(Q) still converges here (dt dS/dQ
    is small), to the oracle's root."""
    from oracle import imex_global as IG
    from oracle.oracle import Oracle
    mesh, Q = _case("tri")
    dt, alpha, n = 0.02, 0.5, 0.05
    s = _global_solver(mesh, 1, alpha, n)
    Qg, st = _solve(s, Q, dt, 40, 1e-11, 1e-15, 1)
    print(st, flush=True)
    assert st["fallbacks"] == st["newton"] >= 1
    assert st["gmres"] == st["newton"]
    assert st["rnorm"] <= 1e-11
    o = Oracle(mesh, order=2, riemann="hll", manning_n=n)
    Qo, _ = IG.newton(o, Q, dt, alpha, n=n)
    assert max_rel_rows(Qg, Qo, (1, 2, 3)) < 1e-9
