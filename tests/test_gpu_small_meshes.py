"""Meshes far below and just off the patch size (256 cells) and the wave
size (64): a few cells where every cell is a boundary cell, a single
partial patch, and full patches plus a ragged last one.  The device march
matches the FP64 oracle march (FP64: relative 1e-12 and the simulated time;
FP32: relative 1e-5 on the state after the same steps, each with its own
dt, which agree to 1e-5)."""
import numpy as np
import pytest

from helpers import circular_dam_state, max_rel_rows

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _mesh(kind):
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    if kind == "line3":
        return Mesh.line(3, 0.0, 1.0).set_bc("wall")
    if kind == "line65":
        return Mesh.line(65, 0.0, 1.0).set_bc("wall")
    if kind == "quad3x2":
        return Mesh.quad(3, 2).set_bc("wall")
    if kind == "quad17x15":
        return Mesh.quad(17, 15).set_bc("wall")
    if kind == "tri_small":
        return delaunay_mesh(jittered_points(7, 5, (0, 1, 0, 1), 0.3, 2),
                             (0, 1, 0, 1)).set_bc("wall")
    pts = jittered_points(23, 13, (0, 1, 0, 1), 0.3, 3)
    return delaunay_mesh(pts, (0, 1, 0, 1)).set_bc("wall")


def _state(mesh):
    if mesh.dim == 2:
        return circular_dam_state(mesh, seed=12)
    x = mesh.cell_centers()[:, 0]
    Q = np.zeros((4, mesh.n_cells))
    Q[1] = np.where(x < 0.5, 2.0, 1.0)
    return Q


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["line3", "line65", "quad3x2", "quad17x15",
                                  "tri_small", "tri_ragged"])
def test_small_and_ragged_meshes_march_matches_oracle(kind, dtype):
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    mesh = _mesh(kind)
    N = mesh.n_cells
    print(kind, "cells", N, flush=True)
    Q0 = _state(mesh)
    o = Oracle(mesh, order=2, limiter="minmod", riemann="hll", cfl=0.9)
    Qo, to, _ = o.march(Q0, n_steps=5)
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=td)
    s.setup_simulation(mesh, ShallowWater(dimension=mesh.dim, order=2,
                                          limiter="minmod", riemann="hll"))
    Q = torch.as_tensor(Q0, dtype=td, device="cuda").contiguous()
    Qg, t_dev, _ = s.march_device(Q, 5)
    rows = (1, 2, 3) if mesh.dim == 2 else (1, 2)
    assert abs(float(t_dev) - to) <= tol * to
    assert max_rel_rows(Qg.double().cpu().numpy(), Qo, rows) < tol
    assert abs(float(s.compute_mass(Qg)) - o.mass(Q0)) <= tol * o.mass(Q0)
