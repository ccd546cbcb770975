"""GPU parity of the implicit (Crank-Nicolson) diffusion (zj_diffusion_*)
against the oracle (oracle/diffusion.py, a sparse direct solve of the
reference's DiffusionOperatorJAX system).

Bars: L u to 1e-13 relative; the CN step, solved by CG to a relative
residual of 1e-12, to 1e-10 relative L2 (the CG error is at most the
condition number of the volume-weighted system times its residual); the
IMEX step with diffusion to 1e-10.
"""
import numpy as np
import pytest
import torch

from helpers import circular_dam_state, max_rel_rows, rel_l2

pytestmark = pytest.mark.gpu


def _mesh(kind):
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    if kind == "quad":
        return Mesh.quad(40, 30, (0.0, 4.0, 0.0, 3.0)).set_bc("wall")
    if kind == "tri":
        pts = jittered_points(31, 23, (0.0, 4.0, 0.0, 3.0), 0.3, 7)
        return delaunay_mesh(pts, (0.0, 4.0, 0.0, 3.0)).set_bc("wall")
    return Mesh.line(200, 0.0, 4.0).set_bc("wall")


def _field(mesh, seed=3):
    c = mesh.cell_centers()
    rng = np.random.default_rng(seed)
    return (np.sin(1.3 * c[:, 0]) * np.cos(0.7 * c[:, 1])
            + 0.1 * rng.normal(size=mesh.n_cells))


@pytest.mark.parametrize("kind", ["quad", "tri", "line"])
def test_explicit_laplacian(kind):
    from oracle import diffusion as D
    from zoomy_hip.fvm import DiffusionOperator
    m = _mesh(kind)
    u = _field(m)
    op = DiffusionOperator(m, m.dim, nu=0.35)
    got = op.explicit(torch.as_tensor(u, device="cuda")).cpu().numpy()
    assert rel_l2(got, D.explicit(m, u, 0.35)) < 1e-13


@pytest.mark.parametrize("kind", ["quad", "tri", "line"])
@pytest.mark.parametrize("dt", [1e-3, 0.05])
def test_crank_nicolson_step(kind, dt):
    """Mildly and strongly stiff steps (dt nu / dx^2 ~ 0.04 .. 30)."""
    from oracle import diffusion as D
    from zoomy_hip.fvm import DiffusionOperator
    m = _mesh(kind)
    u = _field(m)
    op = DiffusionOperator(m, m.dim, nu=0.8)
    got = op.implicit_solve(torch.as_tensor(u, device="cuda"), dt)
    st = op.stats()
    print(kind, dt, st)
    assert st["relres"] <= 1e-12
    assert rel_l2(got.cpu().numpy(), D.cn_solve(m, u, dt, 0.8)) < 1e-10


def test_cosine_mode_amplification():
    """Closed form on the device: a product cosine mode on uniform quads is
    multiplied by (1 + dt lam / 2) / (1 - dt lam / 2)."""
    from zoomy_hip.fvm import DiffusionOperator
    from zoomy_hip.mesh import Mesh
    nx, ny, Lx, Ly, nu, dt = 64, 48, 2.0, 1.5, 0.5, 0.02
    m = Mesh.quad(nx, ny, (0.0, Lx, 0.0, Ly))
    c = m.cell_centers()
    u = np.cos(3 * np.pi * c[:, 0] / Lx) * np.cos(2 * np.pi * c[:, 1] / Ly)
    lam = nu * (-(4 * nx ** 2 / Lx ** 2) * np.sin(3 * np.pi / (2 * nx)) ** 2
                - (4 * ny ** 2 / Ly ** 2) * np.sin(2 * np.pi / (2 * ny)) ** 2)
    amp = (1 + 0.5 * dt * lam) / (1 - 0.5 * dt * lam)
    op = DiffusionOperator(m, 2, nu=nu)
    got = op.implicit_solve(torch.as_tensor(u, device="cuda"), dt).cpu().numpy()
    assert np.abs(got - amp * u).max() < 1e-11


def test_zero_field_and_errors():
    from zoomy_hip import _lib
    from zoomy_hip.fvm import DiffusionOperator
    m = _mesh("quad")
    op = DiffusionOperator(m, 2, nu=1.0)
    z = torch.zeros(m.n_cells, dtype=torch.float64, device="cuda")
    assert torch.count_nonzero(op.implicit_solve(z, 0.1)) == 0
    c = torch.full((m.n_cells,), 2.5, dtype=torch.float64, device="cuda")
    assert torch.allclose(op.implicit_solve(c, 0.1), c, rtol=0, atol=1e-14)
    with pytest.raises(ValueError):
        DiffusionOperator(m, 2, nu=-1.0)
    op2 = DiffusionOperator(m, 2, nu=1.0, tol=1e-30, maxiter=3)
    with pytest.raises(_lib.ZoomyHipError):
        op2.implicit_solve(torch.as_tensor(_field(m), device="cuda"), 0.1)


def test_imex_step_with_diffusion():
    """IMEX step = hyperbolic stage (no source) -> CN on hu and hv ->
    local implicit Manning, against the same composition of oracles."""
    from oracle import diffusion as D
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(32, 28).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=5)
    nu = 2e-3
    model = ShallowWater(dimension=2, order=2, riemann="hll",
                         manning_n=0.03, diffusion={"hu": nu, "hv": nu})
    s = IMEXSourceSolver(time_end=np.inf, cfl=0.9)
    s.setup_simulation(mesh, model)
    oh = Oracle(mesh, order=2, riemann="hll", cfl=0.9)
    om = Oracle(mesh, order=2, riemann="hll", cfl=0.9, manning_n=0.03)
    Qo, Qg = Q0.copy(), torch.as_tensor(Q0, device="cuda")
    for _ in range(3):
        dt = oh.dt(Qo)
        Qe, _ = oh.step(dt, Qo)
        for r in (2, 3):
            Qe[r] = D.cn_solve(mesh, Qe[r], dt, nu)
        Qo = om.imex_local(dt, Qe, maxiter=6)
        Qg = s.step(dt, 0.0, Qg)
    assert max_rel_rows(Qg.cpu().numpy(), Qo, (1, 2, 3)) < 1e-10
