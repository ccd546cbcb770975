"""FP64 oracle of the implicit (Crank-Nicolson) diffusion step -- TEST
INFRASTRUCTURE ONLY.

Only tests/ and bench.py's CPU leg may import this module, and only as the
checker.

It restates ``DiffusionOperatorJAX`` (``zoomy_jax/fvm/reconstruction_jax.py
:1423-1520``), the scalar ``nu`` Laplacian the reference's IMEX solver
applies per state row between its explicit hyperbolic stage and its implicit
source (``zoomy_jax/fvm/solver_imex_jax.py:396-460, 617-619``):

* two-point flux per INTERIOR face f = (a, b) (boundary faces carry no
  diffusive flux, :1467): with e = (x_b - x_a) / |x_b - x_a| and the face
  normal n, ``coeff = nu |f| / |x_b - x_a| / max(|n.e|, 0.1)`` (:1457-1465),
  ``L[a, b] += coeff / |A|``, ``L[a, a] -= coeff / |A|`` and the same for b
  (:1467-1471);
* ``explicit(u) = L u`` (:1481-1483);
* ``implicit_solve(u*, dt)``: Crank-Nicolson
  ``(I - dt/2 L) u = (I + dt/2 L) u*`` (:1485-1516), solved here with a
  sparse direct factorisation (the reference uses a dense solve).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def face_coefficients(mesh):
    """Interior faces ``(a, b, g)`` with ``g = |f| / dist / max(|n.e|,
    0.1)`` (the nu-free coefficient)."""
    fc, fs, _ = mesh.faces()
    nrm, length, _ = mesh.slot_geometry()
    cen = mesh.cell_centers()
    a, b = fc[0], fc[1]
    inner = b >= 0
    a, b, sa = a[inner], b[inner], fs[0][inner]
    dx = cen[b] - cen[a]
    dist = np.linalg.norm(dx, axis=1)
    keep = dist >= 1e-30
    a, b, sa, dx, dist = a[keep], b[keep], sa[keep], dx[keep], dist[keep]
    n = nrm[a, sa]
    ne = (n * dx).sum(1) / dist
    ne = np.maximum(np.abs(ne), 0.1) * np.sign(ne + 1e-30)
    g = length[a, sa] / dist / np.abs(ne)
    return a, b, g


def laplacian(mesh, nu=1.0):
    """The reference's L as a sparse (N, N) matrix."""
    N = mesh.n_cells
    a, b, g = face_coefficients(mesh)
    vol = mesh.cell_volumes()
    c = nu * g
    rows = np.concatenate([a, a, b, b])
    cols = np.concatenate([b, a, a, b])
    vals = np.concatenate([c / vol[a], -c / vol[a], c / vol[b], -c / vol[b]])
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def explicit(mesh, u, nu=1.0):
    return laplacian(mesh, nu) @ np.asarray(u, float)


def cn_solve(mesh, u_star, dt, nu=1.0):
    """One Crank-Nicolson step of du/dt = L u from u_star."""
    L = laplacian(mesh, nu)
    N = L.shape[0]
    eye = sp.identity(N, format="csr")
    rhs = (eye + 0.5 * dt * L) @ np.asarray(u_star, float)
    return spla.spsolve((eye - 0.5 * dt * L).tocsc(), rhs)
