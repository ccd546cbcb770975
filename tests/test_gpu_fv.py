"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bars: FP64 kernels within relative L2 1e-12 on h, hu, hv; FP32 within 1e-5;
integer/index outputs (troubled counts) exact.  The reference's own golden
end states are also reproduced directly on the GPU.
"""
import numpy as np
import pytest

from helpers import (acoustic_ic, bump_lake_ic, circular_dam_state,
                     dam1d_ic, ic_rows, inflow_bc,
                     lake_ic, max_rel_rows, pulse_ic, rel_l2, ritter_ic,
                     smooth_dirichlet_ic, stoker_ic, tilted_ic)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _solver(mesh, dtype=None, recon="conservative", cfl=0.9, **kw):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    dtype = dtype or torch.float64
    model = ShallowWater(dimension=mesh.dim, **kw)
    s = HyperbolicSolver(time_end=np.inf, cfl=cfl, dtype=dtype,
                         reconstruction_variables=recon)
    s.setup_simulation(mesh, model)
    return s


def _oracle(mesh, recon="conservative", cfl=0.9, **kw):
    from oracle.oracle import Oracle
    return Oracle(mesh, recon=recon, cfl=cfl, **kw)


def _to_dev(Q, dtype=None):
    return torch.as_tensor(Q, dtype=dtype or torch.float64, device="cuda")


def test_library_loads_on_gpu():
    from zoomy_hip import _lib
    assert _lib.lib().zj_version().startswith(b"zoomy_hip")
    assert torch.cuda.is_available()


@pytest.mark.parametrize("name,n,L,t_end,bc,ic,riemann", [
    ("parity", 50, 10.0, 0.5, "extrapolation", stoker_ic, "rusanov_nc"),
    ("stoker_wet", 100, 10.0, 1.0, "extrapolation", stoker_ic, "rusanov_nc"),
    ("ritter_dry", 100, 10.0, 1.0, "extrapolation", ritter_ic, "rusanov_nc"),
    ("wb_lake", 100, 10.0, 1.0, "wall", lake_ic, "rusanov"),
    ("bc_wall", 40, 1.0, 0.2, "wall", tilted_ic, "rusanov_nc"),
    ("bc_periodic", 40, 1.0, 0.2, "periodic", tilted_ic, "rusanov_nc"),
    ("bc_inflow", 40, 1.0, 0.2, "inflow", tilted_ic, "rusanov_nc"),
])
def test_reference_goldens_1d(golden, name, n, L, t_end, bc, ic, riemann):
    """The reference's golden end states (state and every aux row)
    reproduced by the HIP march."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.line(n, 0.0, L).set_bc(inflow_bc() if bc == "inflow" else bc)
    model = ShallowWater(dimension=1, order=1, riemann=riemann,
                         initial_conditions=ic, boundary_conditions=bc)
    s = HyperbolicSolver(time_end=t_end, cfl=0.9)
    Q, Qaux = s.solve(mesh, model)
    ref = golden(name)
    Qn = Q.cpu().numpy()
    assert Qn.shape == ref["Q"].shape
    assert np.abs(Qn - ref["Q"]).max() < 1e-12
    A = Qaux.cpu().numpy()
    assert np.abs(A - ref["Qaux"]).max() <= 1e-12 * np.abs(ref["Qaux"]).max()


@pytest.mark.parametrize("name,L,bc,ic,recon,mood", [
    ("stoker_wet_o2_small", 10.0, "extrapolation", stoker_ic,
     "conservative", False),
    ("ritter_dry_o2_small", 10.0, "extrapolation", ritter_ic, "eta_hu", True),
    ("boundary_o2_small", 1.0, "inflow", smooth_dirichlet_ic,
     "conservative", False),
])
def test_reference_o2_goldens(golden, name, L, bc, ic, recon, mood):
    """The order-2 goldens (20 cells, two adaptive steps driven through
    compute_timestep / step / post_step as the reference's tests do) with
    the frozen stage-2 aux they were recorded with."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.line(20, 0.0, L).set_bc(inflow_bc() if bc == "inflow" else bc)
    model = ShallowWater(dimension=1, order=2, limiter="minmod",
                         riemann="rusanov_nc", initial_conditions=ic)
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, mood=mood,
                         reconstruction_variables=recon, stage_aux="frozen")
    Q, Qaux = s.setup_simulation(mesh, model)
    t = 0.0
    for _ in range(2):
        dt = float(s.compute_timestep(Q, Qaux))
        Qn = s.step(dt, t, Q, Qaux)
        Q, Qaux = s.post_step(t + dt, dt, Qn, Q, Qaux)
        t += dt
    ref = golden(name)
    assert np.abs(Q.cpu().numpy()[:3] - ref["Q"]).max() < 1e-13
    A = Qaux.cpu().numpy()
    assert np.abs(A - ref["Qaux"]).max() <= 1e-12 * np.abs(ref["Qaux"]).max()


@pytest.mark.parametrize("name,ic,order", [
    ("swashes_stoker_wet_o1", stoker_ic, 1),
    ("swashes_stoker_wet_o2", stoker_ic, 2),
    ("swashes_ritter_dry_o1", ritter_ic, 1)])
def test_reference_swashes_goldens(golden, name, ic, order):
    """The 400-cell SWASHES end states (t = 6) by the HIP march: the wet
    Stoker marches to round-off, the dry Ritter front to 1e-9 absolute (its
    last wet cells, h < 1e-7, part; the same bound as the oracle's)."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.line(400, 0.0, 10.0).set_bc("extrapolation")
    model = ShallowWater(dimension=1, order=order, limiter="minmod",
                         riemann="rusanov_nc", initial_conditions=ic)
    s = HyperbolicSolver(time_end=6.0, cfl=0.9, stage_aux="frozen")
    Q, _ = s.solve(mesh, model)
    ref = golden(name)["Q"]
    err = np.abs(Q.cpu().numpy() - ref).max()
    assert err < (2e-9 if "ritter" in name else 1e-13)


@pytest.mark.parametrize("name,n,L,t_end,ic,riemann", [
    ("boundary_order2", 640, 1.0, 2.0 / np.sqrt(9.81), acoustic_ic,
     "rusanov_nc"),
    ("wb_long", 200, 10.0, 50.0, bump_lake_ic, "rusanov")])
def test_reference_long_o2_goldens(golden, name, n, L, t_end, ic, riemann):
    """Long order-2 marches between walls (acoustic standing wave, lake at
    rest over a bump) by the HIP march, state and aux."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.line(n, 0.0, L).set_bc("wall")
    model = ShallowWater(dimension=1, order=2, limiter="minmod",
                         riemann=riemann, initial_conditions=ic)
    s = HyperbolicSolver(time_end=t_end, cfl=0.9)
    Q, Qaux = s.solve(mesh, model)
    ref = golden(name)
    assert np.abs(Q.cpu().numpy() - ref["Q"]).max() < 2e-14
    assert np.abs(Qaux.cpu().numpy() - ref["Qaux"]).max() < 5e-12


@pytest.mark.parametrize("recon", ["conservative", "xz", "eta", "eta_hu"])
@pytest.mark.parametrize("kind", ["quad", "tri", "line"])
def test_post_step_aux_every_reconstruction(recon, kind):
    """post_step's aux rows (the patch kernel's gradient pass, in
    conservative variables whatever the reconstruction) against the
    oracle."""
    from zoomy_hip.mesh import Mesh
    if kind == "quad":
        mesh = Mesh.quad(37, 29).set_bc({"left": "wall", "right": "extrapolation",
                                         "bottom": "periodic", "top": "periodic"})
    elif kind == "tri":
        mesh, _ = _wetdry_tri()
    else:
        mesh = Mesh.line(90, 0.0, 3.0).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=13) if kind != "line" else \
        np.stack([0.1 * np.sin(mesh.cell_centers()[:, 0]),
                  1.0 + 0.1 * np.cos(3 * mesh.cell_centers()[:, 0]),
                  0.05 * np.ones(mesh.n_cells), np.zeros(mesh.n_cells)])
    rng = np.random.default_rng(7)
    Q0[2] = rng.normal(0.0, 0.1, mesh.n_cells)
    if kind != "line":
        Q0[3] = rng.normal(0.0, 0.1, mesh.n_cells)
    o = _oracle(mesh, order=2, riemann="hll", recon=recon)
    s = _solver(mesh, order=2, riemann="hll", recon=recon)
    A = s.update_qaux(_to_dev(Q0), None).cpu().numpy()
    Ao = o.aux(Q0)
    assert np.abs(A - Ao).max() <= 1e-12 * np.abs(Ao).max()


@pytest.mark.parametrize("recon,riemann", [
    ("conservative", "hll"), ("conservative", "rusanov_nc"),
    ("eta", "rusanov"), ("xz", "hll")])
def test_2d_mixed_boundaries_step_parity(recon, riemann):
    """Quads with every boundary kind at once: a Dirichlet inflow on the
    left (h and hu prescribed), extrapolation on the right, periodic
    bottom/top; four order-2 steps against the oracle (FP64)."""
    from zoomy_hip.mesh import Dirichlet, Mesh
    mesh = Mesh.quad(40, 30, (0.0, 4.0, 0.0, 3.0)).set_bc(
        {"left": [Dirichlet("h", 1.3), Dirichlet("hu", 0.4)],
         "right": "extrapolation", "bottom": "periodic", "top": "periodic"})
    xc = mesh.cell_centers()
    Q0 = np.zeros((4, mesh.n_cells))
    Q0[0] = 0.1 * np.sin(2.0 * xc[:, 0]) * np.cos(2.0 * np.pi * xc[:, 1] / 3.0)
    Q0[1] = 1.0 + 0.2 * np.exp(-((xc[:, 0] - 2.0) ** 2 + (xc[:, 1] - 1.5) ** 2))
    Q0[2] = 0.1
    o = _oracle(mesh, order=2, limiter="minmod", riemann=riemann, recon=recon)
    s = _solver(mesh, order=2, limiter="minmod", riemann=riemann, recon=recon)
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(4):
        dt = o.dt(Qo)
        Qo, _ = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
    assert max_rel_rows(Qg.cpu().numpy(), Qo, (1, 2, 3)) < 1e-12


@pytest.mark.parametrize("recon,stage_aux,ftol", [
    ("conservative", "frozen", None), ("eta", "frozen", 0.5),
    ("eta_hu", "frozen", None), ("eta_hu", "refresh", None),
    ("eta_hu", "frozen", 0.5)])
@pytest.mark.parametrize("kind", ["quad", "tri"])
def test_stage_aux_and_eta_hu_step_parity(recon, stage_aux, ftol, kind):
    """Frozen stage-2 aux and the conservative-momentum eta reconstruction
    against the oracle on the wet/dry cases, three HLL steps at CFL 1/6
    (FP64).  (The Rusanov speed reads 1/h of each face state, which is
    ill-conditioned where a reconstructed face depth cancels to round-off
    at a shoreline: there the device's contracted multiply-adds and the
    oracle's separate roundings part at O(1e-6), so the Rusanov variants
    are compared on the wet dam break below.)"""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    if kind == "quad":
        from helpers import wetdry_quads
        mesh, Q0 = wetdry_quads(24)
    else:
        mesh, Q0 = _wetdry_tri()
    cfl = 1.0 / 6.0
    o = _oracle(mesh, order=2, limiter="minmod", riemann="hll", recon=recon,
                stage_aux=stage_aux, front_tol=ftol, cfl=cfl)
    s = HyperbolicSolver(time_end=np.inf, cfl=cfl,
                         reconstruction_variables=recon, stage_aux=stage_aux,
                         front_theta_tol=ftol)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          limiter="minmod", riemann="hll"))
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(3):
        dt = o.dt(Qo)
        Qo, _ = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
    Qg = Qg.cpu().numpy()
    assert Qo[1].min() >= 0.0
    assert rel_l2(Qg[1], Qo[1]) < 1e-12
    qs = max(np.linalg.norm(Qo[2]), np.linalg.norm(Qo[3]))
    for r in (2, 3):
        assert np.linalg.norm(Qg[r] - Qo[r]) <= 1e-12 * qs


@pytest.mark.parametrize("riemann", ["rusanov_nc", "rusanov", "hll"])
@pytest.mark.parametrize("recon", ["conservative", "eta"])
def test_frozen_stage_aux_wet_dam_break(riemann, recon):
    """Frozen stage-2 aux on the wet circular dam break (quads, four
    order-2 steps) against the oracle, every Riemann solver."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(48, 40).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=7)
    o = _oracle(mesh, order=2, limiter="minmod", riemann=riemann,
                recon=recon, stage_aux="frozen")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables=recon, stage_aux="frozen")
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          limiter="minmod", riemann=riemann))
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(4):
        dt = o.dt(Qo)
        Qo, _ = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
    assert max_rel_rows(Qg.cpu().numpy(), Qo, (0, 1, 2, 3)) < 1e-12


def test_reference_golden_swe_2d(golden):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(32, 32, (-1.0, 1.0, -1.0, 1.0)).set_bc("extrapolation")
    model = ShallowWater(dimension=2, order=1, riemann="rusanov_nc",
                         initial_conditions=pulse_ic)
    s = HyperbolicSolver(time_end=0.1, cfl=0.9)
    Q, Qaux = s.solve(mesh, model)
    ref = golden("swe_2d")
    assert np.abs(Q.cpu().numpy() - ref["Q"]).max() < 1e-12
    assert np.abs(Qaux.cpu().numpy() - ref["Qaux"]).max() < 1e-11


def test_config1_dam_break_vs_oracle():
    """1000 cells, Rusanov o1, forward Euler, CFL 0.45, to t = 0.1, FP64."""
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.line(1000, 0.0, 1.0).set_bc("extrapolation")
    Q0 = ic_rows(mesh, dam1d_ic)
    o = _oracle(mesh, order=1, riemann="rusanov_nc", cfl=0.45)
    Qo, to, ko = o.march(Q0, t_end=0.1)
    s = _solver(mesh, order=1, riemann="rusanov_nc", cfl=0.45)
    s.time_end = 0.1
    Q, _ = s.run_simulation(_to_dev(Q0)[:3], None)
    assert s.n_steps_taken == ko
    Qg = Q.cpu().numpy()
    for r in (1, 2):
        assert rel_l2(Qg[r], Qo[r]) < 1e-12
    mass = float(s.compute_mass(Q))
    assert abs(mass - o.mass(Qo)) < 1e-14


@pytest.mark.parametrize("riemann", ["rusanov_nc", "rusanov", "hll"])
@pytest.mark.parametrize("limiter", ["minmod", "venkatakrishnan", "vanleer",
                                     "none"])
def test_quad_o2_step_fp64(riemann, limiter):
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(48, 40).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=7)
    o = _oracle(mesh, order=2, limiter=limiter, riemann=riemann, cfl=0.9)
    s = _solver(mesh, order=2, limiter=limiter, riemann=riemann, cfl=0.9)
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(4):
        dto = o.dt(Qo)
        dtg = float(s.compute_timestep(Qg))
        assert abs(dtg - dto) <= 1e-13 * dto
        Qo, bo = o.step(dto, Qo)
        Qg = s.step(dto, 0.0, Qg)
        assert s.troubled_count() == bo
    Qgn = Qg.cpu().numpy()
    assert max_rel_rows(Qgn, Qo, (0, 1, 2, 3)) < 1e-12


def test_flux_operator_and_aux_fp64():
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(33, 29).set_bc({"left": "wall", "right": "extrapolation",
                                     "bottom": "wall", "top": "wall"})
    Q0 = circular_dam_state(mesh, seed=11)
    rng = np.random.default_rng(5)
    Q0[2:] = rng.normal(0, 0.1, (2, mesh.n_cells))
    for order in (1, 2):
        o = _oracle(mesh, order=order, riemann="hll")
        s = _solver(mesh, order=order, riemann="hll")
        dQo = o.flux(Q0)
        dQg = s.get_flux_operator()(0.0, 0.0, _to_dev(Q0), None).cpu().numpy()
        assert max_rel_rows(dQg, dQo) < 1e-12
    Ao = o.aux(Q0)
    Ag = s.update_qaux(_to_dev(Q0), None).cpu().numpy()
    for r in range(Ao.shape[0]):
        assert rel_l2(Ag[r], Ao[r]) < 1e-12


def test_fp32_step_within_1e5():
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(64, 64).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=3)
    o = _oracle(mesh, order=2, limiter="minmod", riemann="hll")
    s = _solver(mesh, order=2, limiter="minmod", riemann="hll",
                dtype=torch.float32)
    dt = o.dt(Q0)
    Qo, _ = o.step(dt, Q0)
    Qg = s.step(dt, 0.0, _to_dev(Q0, torch.float32)).double().cpu().numpy()
    # FP32 state vs the FP64 oracle, relative L2 on the updated state and on
    # the increment of h (the increment isolates the flux arithmetic)
    assert rel_l2(Qg[1], Qo[1]) < 1e-5
    assert rel_l2(Qg[1] - Q0[1], Qo[1] - Q0[1]) < 1e-4
    for r in (2, 3):
        assert np.linalg.norm(Qg[r] - Qo[r]) / np.linalg.norm(Qo[1]) < 1e-5


def test_triangles_eta_rcm_fp64():
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    from zoomy_hip import _lib
    import ctypes
    pts = jittered_points(31, 17, (0.0, 10.0, 0.0, 5.0), 0.3, 3)
    mesh = delaunay_mesh(pts, (0.0, 10.0, 0.0, 5.0))
    ip, ix = mesh.adjacency_csr()
    perm = np.zeros(mesh.n_cells, np.int32)
    assert _lib.lib().zj_rcm(mesh.n_cells, ip.ctypes.data, ix.ctypes.data,
                             perm.ctypes.data) == 0
    mesh = mesh.reorder(perm).set_bc("wall")
    xc = mesh.cell_centers()
    b = 0.3 * np.sin(xc[:, 0]) ** 2 + 0.05 * xc[:, 1]
    Q0 = np.zeros((4, mesh.n_cells))
    Q0[0] = b
    Q0[1] = np.where(xc[:, 0] < 2.0, np.maximum(3.0 - b, 0.0), 0.0)
    o = _oracle(mesh, order=2, limiter="vanleer", riemann="hll", recon="eta")
    s = _solver(mesh, order=2, limiter="vanleer", riemann="hll", recon="eta")
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(3):
        dt = o.dt(Qo)
        Qo, bo = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
        # h < 0 flags: identical except for cells whose h is zero to
        # round-off (dry cells the front has not reached)
        near = int(np.count_nonzero(np.abs(Qo[1]) < 1e-13))
        assert abs(s.troubled_count() - bo) <= near
    assert max_rel_rows(Qg.cpu().numpy(), Qo) < 1e-12


def _wetdry_tri():
    from zoomy_hip.mesh import delaunay_mesh, jittered_points
    pts = jittered_points(31, 17, (0.0, 10.0, 0.0, 5.0), 0.3, 3)
    mesh = delaunay_mesh(pts, (0.0, 10.0, 0.0, 5.0)).set_bc("wall")
    xc = mesh.cell_centers()
    b = 0.3 * np.sin(xc[:, 0]) ** 2 + 0.05 * xc[:, 1]
    Q0 = np.zeros((4, mesh.n_cells))
    Q0[0] = b
    Q0[1] = np.where(xc[:, 0] < 2.0, np.maximum(3.0 - b, 0.0), 0.0)
    return mesh, Q0


@pytest.mark.parametrize("recon,pos,ftol,lim", [
    ("xz", "", None, "none"), ("xz", "", None, "venkatakrishnan"),
    ("eta", "zhang_shu", None, "minmod"), ("eta", "", 0.5, "minmod"),
    ("eta", "zhang_shu", 0.5, "vanleer")])
@pytest.mark.parametrize("kind", ["quad", "tri"])
def test_positivity_guards_step_parity(recon, pos, ftol, lim, kind):
    """The Xing-Zhang scaling of the conservative reconstruction, the
    Zhang-Shu cap and the front pre-detector of the eta reconstruction
    against the oracle on wet/dry cases, three steps at CFL 1/6 (FP64)."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    if kind == "quad":
        from helpers import wetdry_quads
        mesh, Q0 = wetdry_quads(24)
    else:
        mesh, Q0 = _wetdry_tri()
    cfl = 1.0 / 6.0
    o = _oracle(mesh, order=2, limiter=lim, riemann="hll", recon=recon,
                positivity=pos, front_tol=ftol, cfl=cfl)
    s = HyperbolicSolver(time_end=np.inf, cfl=cfl,
                         reconstruction_variables=recon,
                         positivity_method=pos, front_theta_tol=ftol)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2, limiter=lim,
                                          riemann="hll"))
    Qo, Qg = Q0.copy(), _to_dev(Q0)
    for k in range(3):
        dt = o.dt(Qo)
        Qo, _ = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
    Qg = Qg.cpu().numpy()
    assert rel_l2(Qg[1], Qo[1]) < 1e-12
    # the reservoir is a lake at rest in y: hv is round-off noise there, so
    # both momentum rows are measured against the momentum scale
    qs = max(np.linalg.norm(Qo[2]), np.linalg.norm(Qo[3]))
    for r in (2, 3):
        assert np.linalg.norm(Qg[r] - Qo[r]) <= 1e-12 * qs


def test_positivity_options_validated():
    from zoomy_hip.fvm import HyperbolicSolver
    with pytest.raises(ValueError):
        HyperbolicSolver(positivity_method="clip")
    with pytest.raises(ValueError):
        HyperbolicSolver(front_theta_tol=1.5)
    with pytest.raises(ValueError):
        HyperbolicSolver(reconstruction_variables="primitive")


@pytest.mark.parametrize("morton", [False, True])
def test_device_march_and_mass(morton):
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(40, 40)
    if morton:
        mesh = mesh.reorder(mesh.morton_permutation())
    mesh.set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=2)
    o = _oracle(mesh, order=2, riemann="hll")
    s = _solver(mesh, order=2, riemann="hll")
    Qo, to, ko = o.march(Q0, n_steps=6)
    Qaux = torch.zeros((7, mesh.n_cells), dtype=torch.float64, device="cuda")
    Qg, t_dev, dt_dev = s.march_device(_to_dev(Q0), 6, Qaux=Qaux)
    assert abs(float(t_dev) - to) <= 1e-12 * to
    assert max_rel_rows(Qg.cpu().numpy(), Qo) < 1e-12
    Ao = o.aux(Qo)
    Ag = Qaux.cpu().numpy()
    for r in range(7):
        assert rel_l2(Ag[r], Ao[r]) < 1e-12
    m0 = o.mass(Q0)
    mg = float(s.compute_mass(Qg))
    assert abs(mg - o.mass(Qo)) <= 1e-14 * m0
    assert abs(mg - m0) <= 1e-12 * m0          # conservation with walls


def test_imex_local_manning():
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(24, 24).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=4)
    rng = np.random.default_rng(9)
    Q0[2:] = rng.normal(0, 0.5, (2, mesh.n_cells))
    o = _oracle(mesh, manning_n=0.05)
    s = _solver(mesh, manning_n=0.05)
    Qo = o.imex_local(0.01, Q0, maxiter=6)
    from zoomy_hip import _lib
    import ctypes
    out = torch.empty((4, mesh.n_cells), dtype=torch.float64, device="cuda")
    dt = torch.full((), 0.01, dtype=torch.float64, device="cuda")
    Qe = _to_dev(Q0)
    rc = _lib.lib().zj_imex_local(
        s.runtime.handle, ctypes.c_void_p(dt.data_ptr()), 6,
        ctypes.c_void_p(Qe.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    assert max_rel_rows(out.cpu().numpy(), Qo) < 1e-12


def test_dt_guard_raises_on_zero_dt():
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(8, 8).set_bc("wall")
    s = _solver(mesh)
    with pytest.raises(FloatingPointError):
        s.assert_dt_admissible(0.0, 0.0, 3)
    # an all-dry domain is wave-free: dt falls back to dt_max
    s2 = _solver(mesh, dt_max=0.125)
    Q = torch.zeros((4, mesh.n_cells), dtype=torch.float64, device="cuda")
    assert float(s2.compute_timestep(Q)) == 0.125


def _mood_case():
    from helpers import wetdry_quads
    mesh, Q = wetdry_quads()
    o = _oracle(mesh, order=2, limiter="vanleer", riemann="hll", recon="eta")
    return mesh, Q, o


def test_mood_step_parity():
    """One MOOD step from a state whose candidate has only clearly negative
    depths (so the device and the oracle flag the same cells)."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh, Q, o = _mood_case()
    for _ in range(2):
        Q, _, _ = o.step_mood(o.dt(Q), Q)
    dt = o.dt(Q)
    Qc, bc = o.step(dt, Q)
    neg = Qc[1][Qc[1] < 0]
    assert bc > 0 and neg.max() < -1e-9
    Qo, bad, nd = o.step_mood(dt, Q)
    model = ShallowWater(dimension=2, order=2, limiter="vanleer", riemann="hll")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables="eta", mood=True)
    s.setup_simulation(mesh, model)
    Qg = s.step(dt, 0.0, _to_dev(Q)).cpu().numpy()
    assert s.mood_stats() == (1, nd) and s.troubled_count() == bad == 0
    assert max_rel_rows(Qg, Qo) < 1e-12


@pytest.mark.parametrize("n", [7, 8])
def test_mood_march_matches_oracle_mood_steps(n):
    """The device MOOD march (the state and the candidate buffer alternate
    between steps; the head kernel picks the redo's dt when a redo ran and
    keeps the statistics) against the oracle's MOOD steps: the same state
    to round-off after an odd and an even number of steps, and the same
    count of redone steps and demoted cells."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh, Q, o = _mood_case()
    model = ShallowWater(dimension=2, order=2, limiter="vanleer", riemann="hll")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables="eta", mood=True)
    s.setup_simulation(mesh, model)
    Qd = _to_dev(Q)
    s.march_device(Qd, n)
    Qo = Q.copy()
    redone = demoted = 0
    for _ in range(n):
        Qo, bad, nd = o.step_mood(o.dt(Qo), Qo)
        if nd:
            redone += 1
            demoted += nd
    assert redone >= 2
    assert s.mood_stats() == (redone, demoted)
    assert max_rel_rows(Qd.cpu().numpy(), Qo) < 1e-11


def test_mood_march_keeps_depth_nonnegative_and_mass():
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh, Q, o = _mood_case()
    model = ShallowWater(dimension=2, order=2, limiter="vanleer", riemann="hll")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables="eta", mood=True)
    s.setup_simulation(mesh, model)
    Qd = _to_dev(Q)
    m0 = float(s.compute_mass(Qd))
    s.march_device(Qd, 10)
    steps, cells = s.mood_stats()
    assert steps >= 3 and cells > 0
    assert float(Qd[1].min()) >= 0.0 and s.troubled_count() == 0
    assert abs(float(s.compute_mass(Qd)) - m0) < 1e-12 * m0
    # MOOD on a step without troubled cells is the plain step, bit for bit
    from helpers import circular_dam_state
    from zoomy_hip.mesh import Mesh
    m2 = Mesh.quad(32, 32).set_bc("wall")
    Q2 = circular_dam_state(m2, seed=1)
    outs = []
    for mood in (False, True):
        s2 = HyperbolicSolver(time_end=np.inf, cfl=0.9, mood=mood)
        s2.setup_simulation(m2, ShallowWater(dimension=2, order=2,
                                             riemann="hll"))
        outs.append(s2.step(1e-4, 0.0, _to_dev(Q2)).cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
