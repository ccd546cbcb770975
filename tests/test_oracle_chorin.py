"""Pins the Chorin-projection oracle (oracle/chorin.py) to mathematics.

The reference's VAM split comes from zoomy_core, which is absent, and its
Chorin goldens (tests/refs/vam_*.npz) hold VAM(1) states of the reference's
own sub-models; so the oracle is pinned here by what the model fixes:
consistency with the continuum operator S = B^T h^-1 B + 4/h (sympy, second
order), the discrete adjoint/energy identity between the constraint D and
the pressure force B, symmetry/definiteness, the textbook 5-point Laplacian
special case, the lake at rest, the multigrid transfer identities and
mesh-independent MG-CG convergence against a sparse direct solve.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import sympy as sp

from oracle import chorin as C


def _fields(n, L=2.0):
    x, y = sp.symbols("x y")
    h = 1 + 0.3 * sp.sin(x) * sp.cos(2 * y)
    b = 0.2 * sp.cos(x + 0.5 * y)
    return x, y, h, b


def _grid(n, L=2.0):
    dx = L / n
    X, Y = np.meshgrid((np.arange(n) + .5) * dx, (np.arange(n) + .5) * dx,
                       indexing="ij")
    return dx, X, Y


def _orders(errs):
    return [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


def test_operator_second_order_consistent_with_continuum():
    x, y, h, b = _fields(0)
    p = sp.sin(1.3 * x) * sp.cos(0.7 * y) + 0.2
    E = h + 2 * b
    Gx, Gy = sp.diff(E, x), sp.diff(E, y)
    Bx = h * sp.diff(p, x) + p * Gx
    By = h * sp.diff(p, y) + p * Gy
    Ap = -(sp.diff(Bx, x) + sp.diff(By, y)) + (Gx * Bx + Gy * By) / h \
        + 4 * p / h
    fh, fb, fp, fA = (sp.lambdify((x, y), e, "numpy") for e in (h, b, p, Ap))
    errs = []
    for n in (32, 64, 128):
        dx, X, Y = _grid(n)
        coef = C.coefficients(fh(X, Y), fb(X, Y), dx, dx)
        Sp = C.apply(coef, fp(X, Y)) / dx ** 2
        errs.append(np.abs(Sp - fA(X, Y))[2:-2, 2:-2].max())
    assert min(_orders(errs)) > 1.9, errs


def test_constraint_and_correction_consistent_with_continuum():
    x, y, h, b = _fields(0)
    u = 0.3 * sp.cos(x) * sp.sin(y) + 0.1
    v = 0.2 * sp.sin(x + y)
    w = 0.05 * sp.cos(2 * x)
    p = sp.sin(1.3 * x) * sp.cos(0.7 * y)
    E = h + 2 * b
    Cc = sp.diff(h * u, x) + sp.diff(h * v, y) \
        - (u * sp.diff(E, x) + v * sp.diff(E, y)) + 2 * w
    Bx = h * sp.diff(p, x) + p * sp.diff(E, x)
    By = h * sp.diff(p, y) + p * sp.diff(E, y)
    f = {k: sp.lambdify((x, y), e, "numpy")
         for k, e in dict(h=h, b=b, u=u, v=v, w=w, p=p, C=Cc, Bx=Bx,
                          By=By).items()}
    eD, eB = [], []
    for n in (32, 64, 128):
        dx, X, Y = _grid(n)
        H = f["h"](X, Y)
        Q = np.stack([f["b"](X, Y) + 0 * X, H, H * f["u"](X, Y),
                      H * f["v"](X, Y)]).reshape(4, -1)
        hw = (H * f["w"](X, Y)).ravel()
        D = C.constraint(Q, hw, n, n, dx, dx) / dx ** 2
        eD.append(np.abs(D - f["C"](X, Y))[2:-2, 2:-2].max())
        dt = 1e-3
        Qn, _ = C.correct(Q, hw, f["p"](X, Y), n, n, dx, dx, dt)
        bx = ((Q[2] - Qn[2]) / dt).reshape(n, n)
        eB.append(np.abs(bx - f["Bx"](X, Y))[2:-2, 2:-2].max())
    assert min(_orders(eD)) > 1.9, eD
    assert min(_orders(eB)) > 1.9, eB


def test_adjoint_identity_between_constraint_and_pressure_force():
    # sum_c p_c D_c(v, w) = -sum_f V_f v_f (B p)_f + sum_c 2|A| w_c p_c
    rng = np.random.default_rng(1)
    nx, ny, dx, dy = 7, 5, 0.3, 0.2
    h = 1 + 0.5 * rng.random((nx, ny))
    b = 0.3 * rng.random((nx, ny))
    p = rng.standard_normal((nx, ny))
    vx = rng.standard_normal((nx - 1, ny))
    vy = rng.standard_normal((nx, ny - 1))
    w = rng.standard_normal((nx, ny))
    lhs = np.sum(p * C.constraint_faces(h, b, vx, vy, w, dx, dy))
    fx, fy = C.face_bp(h, b, p, dx, dy)
    rhs = -(dy * dx) * (np.sum(vx * fx) + np.sum(vy * fy)) \
        + 2 * dx * dy * np.sum(w * p)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_operator_is_constraint_of_projected_velocities():
    # S p = D(v_f = -hinv(h_f) (B p)_f, w = 2 hinv p)
    rng = np.random.default_rng(2)
    nx, ny, dx, dy = 6, 9, 0.25, 0.4
    h = 0.5 + rng.random((nx, ny))
    b = 0.2 * rng.random((nx, ny))
    p = rng.standard_normal((nx, ny))
    fx, fy = C.face_bp(h, b, p, dx, dy)
    hfx = C.hinv(0.5 * (h[:-1] + h[1:]))
    hfy = C.hinv(0.5 * (h[:, :-1] + h[:, 1:]))
    D = C.constraint_faces(h, b, -hfx * fx, -hfy * fy, 2 * C.hinv(h) * p,
                           dx, dy)
    Sp = C.apply(C.coefficients(h, b, dx, dy), p)
    assert np.abs(Sp - D).max() < 1e-12 * np.abs(Sp).max()


def test_operator_spd_and_flat_case_is_textbook_laplacian():
    rng = np.random.default_rng(3)
    nx, ny, dx, dy = 8, 6, 0.5, 0.5
    h = 1 + rng.random((nx, ny))
    b = 0.4 * rng.random((nx, ny))
    S = C.to_sparse(C.coefficients(h, b, dx, dy)).toarray()
    assert np.abs(S - S.T).max() < 1e-14
    assert np.linalg.eigvalsh(S).min() > 0
    # constant depth, flat bed: S = h * (5-point Neumann Laplacian) + 4|A|/h
    h0 = 1.7
    S = C.to_sparse(C.coefficients(np.full((nx, ny), h0), np.zeros((nx, ny)),
                                   dx, dy)).toarray()
    Lap = np.zeros((nx * ny, nx * ny))
    for i in range(nx):
        for j in range(ny):
            c = i * ny + j
            for ii, jj in ((i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)):
                if 0 <= ii < nx and 0 <= jj < ny:
                    Lap[c, c] += 1
                    Lap[c, ii * ny + jj] -= 1
    ref = h0 * Lap + 4 * dx * dy / h0 * np.eye(nx * ny)
    assert np.abs(S - ref).max() < 1e-12


def test_lake_at_rest_has_zero_pressure():
    n = 16
    rng = np.random.default_rng(4)
    b = 0.3 * rng.random((n, n))
    Q = np.zeros((4, n * n))
    Q[0] = b.ravel()
    Q[1] = (2.0 - b).ravel()
    rhs = C.pressure_rhs(Q, np.zeros(n * n), n, n, 0.1, 0.1, 1e-3)
    assert np.abs(rhs).max() == 0.0


def test_transfer_operators_are_transposes():
    n = 8
    P = C._w1(n)
    assert np.allclose(P.sum(axis=1), 1.0)
    # gather form of the restriction used by the device kernel
    nc = n // 2

    def rw(I, d):
        return {-1: 0.25 if I >= 1 else 0.0, 0: 1.0 if I == 0 else 0.75,
                1: 1.0 if I == nc - 1 else 0.75,
                2: 0.25 if I + 1 < nc else 0.0}[d]
    R = np.zeros((nc, n))
    for I in range(nc):
        for d in (-1, 0, 1, 2):
            if 0 <= 2 * I + d < n:
                R[I, 2 * I + d] = rw(I, d)
    assert np.abs(R - P.T).max() == 0.0


def test_vcycle_symmetric_positive():
    n = 32
    dx, X, Y = _grid(n, 4.0)
    h = 1 + 0.2 * np.sin(X) * np.cos(Y)
    b = 0.1 * np.cos(X + Y)
    H = C.Hierarchy(h, b, dx, dx)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, n, n))
    a1 = np.sum(v * C.vcycle(H, u))
    a2 = np.sum(u * C.vcycle(H, v))
    assert abs(a1 - a2) < 1e-12 * abs(a1)
    assert np.sum(u * C.vcycle(H, u)) > 0


@pytest.mark.parametrize("n", [32, 64, 128])
def test_mg_cg_mesh_independent_and_matches_direct(n):
    box = (0, 64.0 * n / 1024, 0, 64.0 * n / 1024)
    Q, _ = C.solitary_wave_state(n, n, box)
    dx = box[1] / n
    h, b = Q[1].reshape(n, n), Q[0].reshape(n, n)
    f = np.random.default_rng(n).standard_normal((n, n))
    p, it, hist = C.mg_solve(h, b, dx, dx, f, tol=1e-8)
    assert it <= 12 and hist[-1] <= 1e-8
    S = C.to_sparse(C.coefficients(h, b, dx, dx))
    pd = spla.spsolve(S.tocsc(), f.ravel())
    assert np.linalg.norm(p.ravel() - pd) <= 1e-6 * np.linalg.norm(pd)


def test_chorin_cycle_conserves_mass_and_w_advection_is_conservative():
    import sys
    sys.path.insert(0, "zoomy-jax_amd")
    from oracle.oracle import Oracle
    from zoomy_hip.mesh import Mesh
    n = 32
    box = (0, 8.0, 0, 8.0)
    Q, hw = C.solitary_wave_state(n, n, box)
    hw = 0.01 * np.random.default_rng(6).standard_normal(n * n)
    dx = box[1] / n
    hw1 = C.advect_w(Q, hw, n, n, dx, dx, 1e-3)
    assert abs(hw1.sum() - hw.sum()) < 1e-13 * np.abs(hw).sum()
    mesh = Mesh.quad(n, n, box).set_bc("wall")
    co = C.ChorinOracle(Oracle(mesh, order=2, limiter="minmod",
                               riemann="hll"), n, n, box)
    dt = 0.45 * dx / (np.sqrt(9.81) + 1)
    Q1, hw1, (p1, p2) = co.cycle(Q, hw, dt)
    assert abs(Q1[1].sum() - Q[1].sum()) < 1e-12 * Q[1].sum()
    assert np.isfinite(Q1).all() and np.abs(p1).max() > 0


def test_accept_gate():
    assert C.accept_tol(1e-8) == 1e-6
    assert C.accept_tol(1e-4) == pytest.approx(1e-3)
