"""The stage kernel's patch decomposition is a device layout only: patches of
consecutive ids, patches cut from the Morton order of the cell centres (own
cells listed explicitly) and smaller patches give bit-identical FP64
marches, and the clustered layout matches the FP64 oracle like the
consecutive one.  FP32 stages the centres as offsets from each patch's first
own cell, so its marches under different layouts agree to FP32 round-off."""
import numpy as np
import pytest

from helpers import circular_dam_state, rcm_wetdry_triangles

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _march(mesh, Q0, dtype, n, layout, solver_kw, model_kw):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.fvm.solver import patch_plan
    cells, order = layout
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=dtype,
                         patch_cells=cells, patch_order=order, **solver_kw)
    s.setup_simulation(mesh, ShallowWater(dimension=2, **model_kw))
    rc, st = patch_plan(mesh, cells, order)
    assert rc == 0
    print(f"layout {layout}: {st}", flush=True)
    Q = torch.as_tensor(Q0, dtype=dtype, device="cuda").contiguous()
    s.march_device(Q, n)
    torch.cuda.synchronize()
    return Q.cpu().numpy(), st


LAYOUTS = [(0, "auto"), (0, "consecutive"), (0, "clusters"), (64, "clusters"),
           (128, "consecutive")]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_wetdry_triangles_layouts_agree(dtype):
    """The wet/dry triangle recipe in RCM order (eta, van Leer, HLL, front
    detector, MOOD): 25 device steps under five patch layouts (FP32: the
    same to 2e-5 of the depth and momentum scales)."""
    mesh, Q0 = rcm_wetdry_triangles(100, 50)
    kw = dict(reconstruction_variables="eta", mood=True, front_theta_tol=0.5)
    mk = dict(order=2, limiter="vanleer", riemann="hll")
    ref, st = _march(mesh, Q0, dtype, 25, LAYOUTS[0], kw, mk)
    assert st["clustered"] == 1 and st["P"] == 256
    assert np.isfinite(ref).all() and ref[1].min() >= 0.0
    for lay in LAYOUTS[1:]:
        Q, _ = _march(mesh, Q0, dtype, 25, lay, kw, mk)
        if dtype == torch.float64:
            assert np.array_equal(Q, ref), lay
        else:
            assert np.abs(Q[1] - ref[1]).max() <= 2e-5 * ref[1].max(), lay
            qs = np.abs(ref[2:]).max()
            assert np.abs(Q[2:] - ref[2:]).max() <= 2e-5 * qs, lay


def test_strip_ordered_quads_clustered_layout_matches_oracle():
    """Quads numbered along lines (a patch of consecutive ids is one line of
    cells): the automatic layout clusters them; 20 FP64 device steps of the
    circular dam break (minmod + HLL) against the oracle's march."""
    from oracle.oracle import Oracle
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(48, 256).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=2)
    mk = dict(order=2, limiter="minmod", riemann="hll")
    Qc, st = _march(mesh, Q0, torch.float64, 20, (0, "auto"), {}, mk)
    assert st["clustered"] == 1
    Qs, _ = _march(mesh, Q0, torch.float64, 20, (0, "consecutive"), {}, mk)
    assert np.array_equal(Qc, Qs)
    o = Oracle(mesh, cfl=0.9, **mk)
    Qo, _, _ = o.march(Q0, n_steps=20)
    for r in (1, 2, 3):
        scale = max(np.abs(Qo[r]).max(), 1.0)
        assert np.abs(Qc[r] - Qo[r]).max() <= 1e-12 * scale, r


@pytest.mark.parametrize("recon,kw", [
    ("conservative", dict(order=2, limiter="minmod", riemann="hll")),
    ("eta", dict(order=2, limiter="vanleer", riemann="hll")),
])
def test_api_entry_points_agree_across_layouts(recon, kw):
    """Every entry point that runs the patch kernel -- step, flux_operator
    with a force_o1 mask, the reconstruction object with its limited
    gradient, post_step's aux rows, compute_timestep -- gives bit-identical
    FP64 results on the RCM-ordered wet/dry triangles with the clustered
    layout (own cells listed) and with consecutive-id patches."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh, Q0 = rcm_wetdry_triangles(100, 50)
    Q0 = Q0.copy()
    Q0[1] += 0.01            # wet everywhere: every code path has work
    rng = np.random.default_rng(3)
    Q0[2] = rng.normal(0.0, 0.05, mesh.n_cells) * Q0[1]
    Q0[3] = rng.normal(0.0, 0.05, mesh.n_cells) * Q0[1]
    mask = torch.zeros(mesh.n_cells, dtype=torch.uint8)
    mask[::5] = 1
    outs = {}
    for order in ("clusters", "consecutive"):
        s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                             reconstruction_variables=recon,
                             patch_order=order)
        s.setup_simulation(mesh, ShallowWater(dimension=2, **kw))
        Q = torch.as_tensor(Q0, dtype=torch.float64, device="cuda")
        dt = s.compute_timestep(Q)
        res = {"dt": dt.cpu().numpy()}
        res["step"] = s.step(float(dt), 0.0, Q).cpu().numpy()
        fo = s.get_flux_operator()
        res["dQ"] = fo(float(dt), 0.0, Q, None, None, None).cpu().numpy()
        if s.supports_force_o1:
            res["dQ_o1"] = fo(float(dt), 0.0, Q, None, None, None,
                              force_o1=mask.cuda()).cpu().numpy()
        QL, QR, lg = s.get_reconstruction().reconstruct_with_grad(Q)
        res.update(QL=QL.cpu().numpy(), QR=QR.cpu().numpy(),
                   lg=lg.cpu().numpy())
        res["aux"] = s.update_qaux(Q, None).cpu().numpy()
        outs[order] = res
        print(order, {k: v.shape for k, v in res.items()}, flush=True)
    a, b = outs["clusters"], outs["consecutive"]
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_renumbered_march_keeps_the_callers_order(dtype):
    """A strip-ordered mesh (RCM) with the automatic layout marches on its
    renumbered twin (cells in the patch clusters' order) and hands the state
    back in the caller's order: equal to the march on patches of
    consecutive ids -- state, aux rows of the final state, time, dt and the
    MOOD statistics -- bit for bit in FP64 (FP32: to 2e-5 of the scales)."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.fvm.solver import patch_plan
    mesh, Q0 = rcm_wetdry_triangles(100, 50)
    rc, st = patch_plan(mesh, 0, "auto")
    assert rc == 0 and st["clustered"] == 1
    kw = dict(reconstruction_variables="eta", mood=True, front_theta_tol=0.5)
    mk = dict(order=2, limiter="vanleer", riemann="hll")
    res = {}
    for order in ("auto", "consecutive"):
        s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=dtype,
                             patch_order=order, **kw)
        s.setup_simulation(mesh, ShallowWater(dimension=2, **mk))
        Q = torch.as_tensor(Q0, dtype=dtype, device="cuda").contiguous()
        Qaux = torch.zeros((7, mesh.n_cells), dtype=dtype, device="cuda")
        _, t, dt = s.march_device(Q, 13, Qaux=Qaux)
        res[order] = (Q.double().cpu().numpy(), Qaux.double().cpu().numpy(),
                      float(t), float(dt), tuple(s.mood_stats()))
        print(order, res[order][2:], flush=True)
    a, b = res["auto"], res["consecutive"]
    assert np.isfinite(a[0]).all() and a[0][1].min() >= 0.0
    if dtype == torch.float64:
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[2:] == b[2:]
    else:
        for r in (1, 2, 3):
            scale = max(np.abs(b[0][r]).max(), 1e-30)
            assert np.abs(a[0][r] - b[0][r]).max() <= 2e-5 * scale, r
        assert abs(a[2] - b[2]) <= 1e-5 * b[2]
