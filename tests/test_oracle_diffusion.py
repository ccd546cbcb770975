"""Pins the implicit-diffusion oracle (oracle/diffusion.py) to mathematics.

The reference's scalar diffusion operator (DiffusionOperatorJAX,
zoomy_jax/fvm/reconstruction_jax.py:1423-1520) has no golden output in the
reference's tests, so it is pinned by what the discretisation fixes: on
uniform grids with zero-flux boundaries the cosine modes are its
eigenvectors with the closed-form eigenvalues of the cell-centred
Neumann Laplacian, and Crank-Nicolson multiplies each mode by
(1 + dt lambda / 2) / (1 - dt lambda / 2); on any mesh it conserves
sum |A| u, annihilates constants, is symmetric in the |A|-weighted inner
product and has non-negative off-diagonals.
"""
import numpy as np
import pytest

from oracle import diffusion as D
from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points


def _lam(k, n, L):
    dx = L / n
    return -(4.0 / dx ** 2) * np.sin(k * np.pi * dx / (2.0 * L)) ** 2


@pytest.mark.parametrize("k", [1, 3, 7])
def test_1d_cosine_modes(k):
    n, L, nu, dt = 64, 2.5, 0.3, 0.01
    m = Mesh.line(n, 0.0, L)
    x = m.cell_centers()[:, 0]
    u = np.cos(k * np.pi * x / L)
    lam = nu * _lam(k, n, L)
    assert np.abs(D.explicit(m, u, nu) - lam * u).max() < 1e-10 * abs(lam)
    amp = (1 + 0.5 * dt * lam) / (1 - 0.5 * dt * lam)
    assert np.abs(D.cn_solve(m, u, dt, nu) - amp * u).max() < 1e-12


@pytest.mark.parametrize("k,j", [(1, 0), (2, 3), (5, 1)])
def test_2d_quad_product_modes(k, j):
    nx, ny, Lx, Ly, nu, dt = 24, 18, 3.0, 2.0, 0.7, 0.05
    m = Mesh.quad(nx, ny, (0.0, Lx, 0.0, Ly))
    c = m.cell_centers()
    u = np.cos(k * np.pi * c[:, 0] / Lx) * np.cos(j * np.pi * c[:, 1] / Ly)
    lam = nu * (_lam(k, nx, Lx) + _lam(j, ny, Ly))
    assert np.abs(D.explicit(m, u, nu) - lam * u).max() < 1e-9 * abs(lam)
    amp = (1 + 0.5 * dt * lam) / (1 - 0.5 * dt * lam)
    assert np.abs(D.cn_solve(m, u, dt, nu) - amp * u).max() < 1e-12


def _tri():
    pts = jittered_points(21, 15, (0.0, 4.0, 0.0, 3.0), 0.3, 5)
    return delaunay_mesh(pts, (0.0, 4.0, 0.0, 3.0))


@pytest.mark.parametrize("kind", ["tri", "quad", "line"])
def test_structure_on_any_mesh(kind):
    m = {"tri": _tri, "quad": lambda: Mesh.quad(13, 9),
         "line": lambda: Mesh.line(30, 0.0, 1.0)}[kind]()
    L = D.laplacian(m, 0.4).toarray()
    vol = m.cell_volumes()
    u = np.random.default_rng(3).normal(size=m.n_cells)
    # conservation and constants
    assert abs(vol @ (L @ u)) < 1e-12 * np.abs(vol * (L @ u)).sum()
    assert np.abs(L.sum(1)).max() < 1e-12 * np.abs(L).max()
    # |A|-weighted symmetry, non-negative couplings
    VL = vol[:, None] * L
    assert np.abs(VL - VL.T).max() < 1e-12 * np.abs(VL).max()
    off = L - np.diag(np.diag(L))
    assert off.min() >= 0.0
    # Crank-Nicolson conserves the weighted sum and is stable
    un = D.cn_solve(m, u, 0.5, 0.4)
    assert abs(vol @ un - vol @ u) < 1e-12 * np.abs(vol * u).sum()
    assert np.sqrt(vol @ un ** 2) <= np.sqrt(vol @ u ** 2) * (1 + 1e-12)
