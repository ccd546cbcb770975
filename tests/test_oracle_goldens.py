"""The oracle against the reference's own golden end states.

tests/golden/refs_*.npz are copies of the reference's tests/refs/*.npz (whole
marches of zoomy_jax on CPU in x64).  Cases, meshes, BCs and CFL follow the
reference's tests/cases.py and tests/runtime/*.py.
"""
import numpy as np
import pytest

from helpers import (ic_rows, lake_ic, pulse_ic, ritter_ic, stoker_ic,
                     tilted_ic)
from oracle.oracle import Oracle
from zoomy_hip.mesh import Mesh


@pytest.mark.parametrize("name,n,L,t_end,bc,ic,riemann", [
    # test_parity_numpy_jax.py:19-47, test_swe_stoker_wet.py:20
    ("parity", 50, 10.0, 0.5, "extrapolation", stoker_ic, "rusanov_nc"),
    ("stoker_wet", 100, 10.0, 1.0, "extrapolation", stoker_ic, "rusanov_nc"),
    ("ritter_dry", 100, 10.0, 1.0, "extrapolation", ritter_ic, "rusanov_nc"),
    # test_wb_lake_at_rest.py:19-41 (PositiveNonconservativeRusanov, walls)
    ("wb_lake", 100, 10.0, 1.0, "wall", lake_ic, "rusanov"),
    # test_bc_runtime.py: wall kind
    ("bc_wall", 40, 1.0, 0.2, "wall", tilted_ic, "rusanov_nc"),
    # test_bc_runtime.py: periodic kind (the seam pairs the partner cell's
    # state with the inner cell's aux; s_M from each face state's own 1/h)
    ("bc_periodic", 40, 1.0, 0.2, "periodic", tilted_ic, "rusanov_nc"),
])
def test_1d_goldens(golden, name, n, L, t_end, bc, ic, riemann):
    m = Mesh.line(n, 0.0, L).set_bc(bc)
    o = Oracle(m, order=1, riemann=riemann, cfl=0.9)
    Q, t, k = o.march(ic_rows(m, ic), t_end=t_end)
    ref = golden(name)
    assert abs(t - t_end) < 1e-12
    assert np.abs(Q[:3] - ref["Q"]).max() < 1e-15
    A = o.aux(Q)
    if name in ("bc_wall", "bc_periodic"):
        # boundary-cell derivative aux rows are not reproduced (the boundary
        # row weights of zoomy_core's LSQMesh are unknown); interior exact
        assert np.abs(A[:, 1:-1] - ref["Qaux"][:, 1:-1]).max() < 1e-12
        assert np.abs(A[3] - ref["Qaux"][3]).max() < 1e-12
    else:
        assert np.abs(A - ref["Qaux"]).max() < 1e-12


def test_wb_lake_is_well_balanced(golden):
    m = Mesh.line(100, 0.0, 10.0).set_bc("wall")
    o = Oracle(m, order=1, riemann="rusanov", cfl=0.9)
    Q, _, _ = o.march(ic_rows(m, lake_ic), t_end=1.0)
    eta = Q[0] + Q[1]
    assert np.abs(eta - eta[0]).max() < 1e-12
    assert np.abs(Q[2] / Q[1]).max() < 1e-12
    # the plain NC Rusanov diffuses the bed and is NOT well balanced
    o2 = Oracle(m, order=1, riemann="rusanov_nc", cfl=0.9)
    Q2, _, _ = o2.march(ic_rows(m, lake_ic), t_end=1.0)
    assert np.abs(Q2[:3] - golden("wb_lake")["Q"]).max() > 1e-3


def test_swe_2d_golden(golden):
    """32x32 quads on [-1,1]^2, Gaussian pulse, o1, t = 0.1
    (test_swe_2d.py:19-39): cell order, 2-D dt law and the 3x3 LSQ stencil
    with per-direction boundary ghosts of the derivative aux rows."""
    m = Mesh.quad(32, 32, (-1.0, 1.0, -1.0, 1.0)).set_bc("extrapolation")
    o = Oracle(m, order=1, riemann="rusanov_nc", cfl=0.9)
    Q, t, k = o.march(ic_rows(m, pulse_ic), t_end=0.1)
    ref = golden("swe_2d")
    assert k == 9
    assert np.abs(Q - ref["Q"]).max() < 1e-15
    assert np.abs(o.aux(Q) - ref["Qaux"]).max() < 1e-13


@pytest.mark.parametrize("name,ic", [("stoker_wet_o2_small", stoker_ic),
                                     ("ritter_dry_o2_small", ritter_ic)])
def test_o2_goldens_not_reproduced(golden, name, ic):
    """Order-2 LSQ-MUSCL goldens (20 cells, 2 steps): the composition of
    zoomy_core's LSQ stencil/limiter/flux at order 2 is not reproduced; the
    oracle stays within a few percent.  Recorded so that a change that
    closes (or widens) the gap is noticed."""
    m = Mesh.line(20, 0.0, 10.0).set_bc("extrapolation")
    o = Oracle(m, order=2, limiter="minmod", riemann="rusanov_nc", cfl=0.9)
    Q, _, k = o.march(ic_rows(m, ic), n_steps=2)
    ref = golden(name)["Q"]
    err = np.abs(Q[:3] - ref).max()
    assert 1e-6 < err < 5e-4
    # mass identical (both conservative)
    assert abs(Q[1].sum() - ref[1].sum()) < 1e-15
