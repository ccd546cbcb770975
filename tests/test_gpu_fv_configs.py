"""GPU parity of the workload recipes and options that the plain step tests
leave out: the wet/dry triangle recipe at FP32 in RCM order, the explicit
Manning source, the in-place device march with the frozen stage aux, and the
IMEX explicit stage with MOOD and the positivity guards.  Every case runs the
HIP path through the C ABI against the FP64 oracle."""
import math

import numpy as np
import pytest

from helpers import (circular_dam_state, ic_rows, max_rel_rows, rel_l2,
                     rcm_wetdry_triangles, stoker_ic, wetdry_quads)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(Q, dtype=torch.float64):
    return torch.as_tensor(np.asarray(Q), dtype=dtype, device="cuda")


def test_wetdry_triangles_fp32_recipe_vs_fp64_oracle():
    """The wet/dry triangle recipe (Delaunay of jittered points, RCM order,
    fractal valley bed, reservoir h = 3 in x < 2; van Leer on the eta
    reconstruction, HLL, front pre-detector 0.5, MOOD) with an FP32 state,
    one step at a time from FP32-representable states along the FP64
    oracle's own march: the FP32 step against the FP64 oracle's step within
    relative L2 1e-5 on the wet cells (h > 1e-3) for h and both momentum
    rows; the FP64 compensated mass of the FP32 state equal to the exactly
    rounded sum; mass conserved by the FP32 march."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh, Q0 = rcm_wetdry_triangles(100, 50)
    assert 9000 < mesh.n_cells < 11000
    kw = dict(order=2, limiter="vanleer", riemann="hll")
    o = Oracle(mesh, recon="eta", front_tol=0.5, cfl=0.9, **kw)
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=torch.float32,
                         reconstruction_variables="eta", mood=True,
                         front_theta_tol=0.5)
    s.setup_simulation(mesh, ShallowWater(dimension=2, **kw))
    vol = mesh.cell_volumes()
    Qo = Q0.astype(np.float32).astype(np.float64)
    worst = {}
    for k in range(60):
        Q32 = Qo.astype(np.float32).astype(np.float64)
        dt = o.dt(Q32)
        Qn, bad, nd = o.step_mood(dt, Q32)
        if k % 10 == 9:
            print(f"step {k}: dt {dt:.3e}, oracle demoted {nd}", flush=True)
            Qg = s.step(dt, 0.0, _dev(Q32, torch.float32)).double()
            Qg = Qg.cpu().numpy()
            wet = Qn[1] > 1e-3
            eh = rel_l2(Qg[1][wet], Qn[1][wet])
            qs = max(np.linalg.norm(Qn[2][wet]), np.linalg.norm(Qn[3][wet]))
            eq = max(np.linalg.norm(Qg[r][wet] - Qn[r][wet]) for r in (2, 3))
            worst[k] = (eh, eq / qs)
            mg = float(s.compute_mass(_dev(Q32, torch.float32)))
            exact = math.fsum((Q32[1] * vol).tolist())
            assert abs(mg - exact) <= 1e-15 * exact
        Qo = Qn
    print("FP32 step vs FP64 oracle (h, momentum) on wet cells:", worst)
    for k, (eh, eq) in worst.items():
        assert eh < 1e-5 and eq < 1e-5, (k, eh, eq)
    # the FP32 march conserves the FP64-accumulated mass to its round-off
    Qd = _dev(Q0, torch.float32)
    m0 = float(s.compute_mass(Qd))
    s.march_device(Qd, 30)
    assert float(Qd[1].min()) >= 0.0 and bool(torch.isfinite(Qd).all())
    assert abs(float(s.compute_mass(Qd)) - m0) <= 1e-6 * m0


@pytest.mark.parametrize("kind,order", [("quad", 1), ("quad", 2),
                                        ("tri", 2), ("line", 2)])
def test_explicit_manning_source_step(kind, order):
    """The explicit Manning friction of the step (the source stage after
    the hyperbolic update, solver_jax.py:1129) with n > 0 against the
    oracle, three steps, FP64."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    if kind == "quad":
        mesh = Mesh.quad(40, 36).set_bc("wall")
        Q0 = circular_dam_state(mesh, seed=13)
    elif kind == "tri":
        mesh = delaunay_mesh(jittered_points(30, 24, (0, 1, 0, 1), 0.3, 5),
                             (0, 1, 0, 1)).set_bc("wall")
        Q0 = circular_dam_state(mesh, seed=13)
    else:
        mesh = Mesh.line(200, 0.0, 10.0).set_bc("extrapolation")
        Q0 = ic_rows(mesh, stoker_ic)
    rng = np.random.default_rng(2)
    Q0[2] += rng.normal(0.0, 0.05, mesh.n_cells) * Q0[1]
    if mesh.dim == 2:
        Q0[3] += rng.normal(0.0, 0.05, mesh.n_cells) * Q0[1]
    kw = dict(order=order, limiter="minmod", riemann="hll", manning_n=0.04)
    o = Oracle(mesh, cfl=0.9, **kw)
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9)
    s.setup_simulation(mesh, ShallowWater(dimension=mesh.dim, **kw))
    Qo, Qg = Q0.copy(), _dev(Q0)
    for _ in range(3):
        dt = o.dt(Qo)
        Qo, _ = o.step(dt, Qo)
        Qg = s.step(dt, 0.0, Qg)
    Qg = Qg.cpu().numpy()
    rows = (1, 2, 3) if mesh.dim == 2 else (1, 2)
    assert max_rel_rows(Qg, Qo, rows) < 1e-12
    # the friction acted: the same steps without it differ
    o0 = Oracle(mesh, cfl=0.9, **{**kw, "manning_n": 0.0})
    Qf = Q0.copy()
    for _ in range(3):
        Qf, _ = o0.step(o0.dt(Qf), Qf)
    assert rel_l2(Qf[2], Qo[2]) > 1e-6


@pytest.mark.parametrize("case", ["stoker_o2", "wetdry_quads", "dam_quads"])
def test_frozen_stage_aux_march_in_place(case):
    """march_device with the frozen stage-2 aux updates Q in place; its
    second stage reads the neighbours' step-start depth, so the in-place
    step goes through scratch.  Against the oracle's march, several steps."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    if case == "stoker_o2":
        mesh = Mesh.line(20, 0.0, 10.0).set_bc("extrapolation")
        Q0 = ic_rows(mesh, stoker_ic)
        kw = dict(order=2, limiter="minmod", riemann="rusanov_nc")
        recon, steps, rows = "conservative", 12, (1, 2)
    elif case == "wetdry_quads":
        # CFL 1/6 keeps the depths non-negative without MOOD
        mesh, Q0 = wetdry_quads(32)
        kw = dict(order=2, limiter="minmod", riemann="hll")
        recon, steps, rows = "eta", 6, (1, 2, 3)
    else:
        mesh = Mesh.quad(64, 64).set_bc("wall")
        Q0 = circular_dam_state(mesh, seed=21)
        kw = dict(order=2, limiter="minmod", riemann="hll")
        recon, steps, rows = "conservative", 8, (1, 2, 3)
    wd = case == "wetdry_quads"
    cfl = 1.0 / 6.0 if wd else 0.9
    pos = "zhang_shu" if wd else ""
    o = Oracle(mesh, recon=recon, cfl=cfl, stage_aux="frozen",
               positivity=pos, **kw)
    Qo, to, ko = o.march(Q0, n_steps=steps)
    assert Qo[1].min() >= -1e-20
    s = HyperbolicSolver(time_end=np.inf, cfl=cfl,
                         reconstruction_variables=recon, stage_aux="frozen",
                         positivity_method=pos)
    s.setup_simulation(mesh, ShallowWater(dimension=mesh.dim, **kw))
    Qd = _dev(Q0)
    _, t_dev, _ = s.march_device(Qd, steps)
    assert abs(float(t_dev) - to) <= 1e-12 * to
    Qg = Qd.cpu().numpy()
    # the in-place march agrees with the out-of-place step loop (the two
    # differ only in how dt's minimum is formed: last-bit differences)
    Qs = _dev(Q0)
    for _ in range(steps):
        Qs = s.step(s.compute_timestep(Qs), 0.0, Qs)
    Qs = Qs.cpu().numpy()
    Qg = Qg[:Qs.shape[0]]
    assert np.abs(Qs - Qg).max() <= 1e-13 * np.abs(Qg).max()
    if case == "wetdry_quads":
        assert rel_l2(Qg[1], Qo[1]) < 1e-12
        qs = max(np.linalg.norm(Qo[2]), np.linalg.norm(Qo[3]))
        for r in (2, 3):
            assert np.linalg.norm(Qg[r] - Qo[r]) <= 1e-12 * qs
    else:
        assert max_rel_rows(Qg, Qo, rows) < 1e-12


def test_imex_explicit_stage_keeps_mood_and_positivity():
    """IMEXSourceSolver with MOOD and the Zhang-Shu cap on a wet/dry case:
    its explicit stage runs both (as the reference's IMEX explicit stage
    does, solver_imex_jax.py:604-611), then the local implicit Manning
    stage; against the same composition of oracles."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    mesh, Q0 = wetdry_quads(32)
    kw = dict(order=2, limiter="vanleer", riemann="hll")
    s = IMEXSourceSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables="eta", mood=True,
                         positivity_method="zhang_shu")
    s.setup_simulation(mesh, ShallowWater(dimension=2, manning_n=0.03, **kw))
    oh = Oracle(mesh, recon="eta", positivity="zhang_shu", cfl=0.9, **kw)
    om = Oracle(mesh, recon="eta", cfl=0.9, manning_n=0.03, **kw)
    Qo, Qg = Q0.copy(), _dev(Q0)
    for _ in range(4):
        dt = oh.dt(Qo)
        Qe, bad, nd = oh.step_mood(dt, Qo)
        Qo = om.imex_local(dt, Qe, maxiter=6)
        Qg = s.step(dt, 0.0, Qg)
    Qg = Qg.cpu().numpy()
    # MOOD leaves depths >= 0 up to round-off of the shoreline cells
    assert Qo[1].min() >= -1e-20 and Qg[1].min() >= -1e-20
    assert rel_l2(Qg[1], Qo[1]) < 1e-12
    qs = max(np.linalg.norm(Qo[2]), np.linalg.norm(Qo[3]))
    for r in (2, 3):
        assert np.linalg.norm(Qg[r] - Qo[r]) <= 1e-12 * qs


def test_qaux_is_validated():
    """A Qaux of the wrong dtype, shape or device is refused before any
    kernel writes through it."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(16, 16).set_bc("wall")
    s = HyperbolicSolver(time_end=np.inf)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          riemann="hll"))
    Q = _dev(circular_dam_state(mesh))
    for bad in (torch.zeros((7, mesh.n_cells), dtype=torch.float32,
                            device="cuda"),
                torch.zeros((6, mesh.n_cells), dtype=torch.float64,
                            device="cuda"),
                torch.zeros((7, mesh.n_cells), dtype=torch.float64)):
        with pytest.raises(ValueError):
            s.update_qaux(Q, bad)
        with pytest.raises(ValueError):
            s.march_device(Q.clone(), 1, Qaux=bad)
