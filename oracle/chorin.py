"""FP64 oracle of the Chorin projection step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.

It restates the projection-correction cycle of the reference's
``ChorinSplitVAMSolverJax`` (``zoomy_jax/fvm/solver_chorin_vam_jax.py``:
``_hyperbolic_stage`` :496-522, ``chorin_stage`` :524-572, ``chorin_cycle``
:574-615, pressure step ``_step_pressure_pure`` :673-883 with the accept gate
``max(1e-6, 10 tol)`` :72-91, corrector ``_step_corrector_pure`` :885-907)
for the depth-averaged non-hydrostatic shallow-water model.  The reference
takes the sub-models from zoomy_core's VAM ``chorin_split``, which is not
available, so the model is restated from the physics (linear vertical
profile of the non-hydrostatic pressure p with bottom value 2p):

    d(hu)/dt + ...            = -(h grad p + p grad(h + 2b))  =: -B p
    d(hw)/dt + div(hu w)      = 2 p
    C = div(hu) - u.grad(h+2b) + 2w = 0          (depth-integrated continuity
                                                  of the vertical velocity)

C is minus the adjoint of B (with weight 1/h), so the pressure operator
S = B^T h^-1 B + 4/h is symmetric positive definite.  Discretely, on a
uniform nx x ny quad grid with walls (cell id i*ny + j), with E = h + 2b and
for the face between cells c and e (|f| its length, d_f the centre
distance):

    (B p)_f  = (a_e p_e - a_c p_c) / d_f,   a_c = h_c + b_c - b_e
                                            a_e = h_e + b_e - b_c
    S_cc    += c_f a_c^2,   S_ce = -c_f a_c a_e,   c_f = |f|/d_f hinv(h_f)
    S_cc    += 4 |A_c| hinv(h_c)
    D_c      = sum_f |f| v_f a_c + 2 |A_c| w_c     (v_f: outward normal
                                                    velocity at the face)

so that S p = D(v = -hinv_f B p, w = 2 hinv p): S is exactly the operator
that maps a pressure to the constraint change of the face-projected
velocities.  Stage: predictor Q* (FV Euler step, hw advected upwind),
S(h*) p = -D(Q*)/dt with v_f = (u_c + u_e).n / 2, corrector
hu -= dt (cell average of the face B p), hw += 2 dt p.

The geometric multigrid (rediscretised coarse levels from 4-cell averages of
h and b, red-black Gauss-Seidel, cell-centred bilinear prolongation P and
restriction R = P^T) and the preconditioned CG mirror the device kernels
operation for operation.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

G = 9.81
EPS = 1e-8


def hinv(h, eps=EPS):
    m = np.maximum(h, eps)
    return np.sqrt(2.0) * h / np.sqrt(h ** 4 + m ** 4)


# ---------------------------------------------------------------- operator
def coefficients(h, b, dx, dy, eps=EPS):
    """5-point stencil of S on an (nx, ny) grid: (diag, cE, cN) with
    cE[i, j] = S[(i,j),(i+1,j)], cN[i, j] = S[(i,j),(i,j+1)] (0 past the
    last row/column)."""
    h = np.asarray(h, np.float64)
    b = np.asarray(b, np.float64)
    nx, ny = h.shape
    diag = 4.0 * dx * dy * hinv(h, eps)
    cE = np.zeros_like(h)
    cN = np.zeros_like(h)
    # x-faces
    hf = 0.5 * (h[:-1] + h[1:])
    cf = (dy / dx) * hinv(hf, eps)
    ac = h[:-1] + b[:-1] - b[1:]
    ae = h[1:] + b[1:] - b[:-1]
    diag[:-1] += cf * ac * ac
    diag[1:] += cf * ae * ae
    cE[:-1] = -cf * ac * ae
    # y-faces
    hf = 0.5 * (h[:, :-1] + h[:, 1:])
    cf = (dx / dy) * hinv(hf, eps)
    ac = h[:, :-1] + b[:, :-1] - b[:, 1:]
    ae = h[:, 1:] + b[:, 1:] - b[:, :-1]
    diag[:, :-1] += cf * ac * ac
    diag[:, 1:] += cf * ae * ae
    cN[:, :-1] = -cf * ac * ae
    return diag, cE, cN


def apply(coef, x):
    diag, cE, cN = coef
    y = diag * x
    y[:-1] += cE[:-1] * x[1:]
    y[1:] += cE[:-1] * x[:-1]
    y[:, :-1] += cN[:, :-1] * x[:, 1:]
    y[:, 1:] += cN[:, :-1] * x[:, :-1]
    return y


def to_sparse(coef):
    diag, cE, cN = coef
    nx, ny = diag.shape
    N = nx * ny
    idx = np.arange(N).reshape(nx, ny)
    r = [idx.ravel()]
    c = [idx.ravel()]
    v = [diag.ravel()]
    for a, bb, w in ((idx[:-1], idx[1:], cE[:-1]),
                     (idx[:, :-1], idx[:, 1:], cN[:, :-1])):
        r += [a.ravel(), bb.ravel()]
        c += [bb.ravel(), a.ravel()]
        v += [w.ravel(), w.ravel()]
    return sp.csr_matrix((np.concatenate(v), (np.concatenate(r),
                                              np.concatenate(c))),
                         shape=(N, N))


def face_bp(h, b, p, dx, dy):
    """(B p) on x-faces (oriented +x) and y-faces (oriented +y)."""
    ac = h[:-1] + b[:-1] - b[1:]
    ae = h[1:] + b[1:] - b[:-1]
    fx = (ae * p[1:] - ac * p[:-1]) / dx
    ac = h[:, :-1] + b[:, :-1] - b[:, 1:]
    ae = h[:, 1:] + b[:, 1:] - b[:, :-1]
    fy = (ae * p[:, 1:] - ac * p[:, :-1]) / dy
    return fx, fy


def constraint_faces(h, b, vx, vy, w, dx, dy):
    """D from face normal velocities vx (x-faces, +x) and vy (y-faces, +y)
    and cell w."""
    D = 2.0 * dx * dy * w
    ac = h[:-1] + b[:-1] - b[1:]
    ae = h[1:] + b[1:] - b[:-1]
    D[:-1] += dy * vx * ac
    D[1:] -= dy * vx * ae
    ac = h[:, :-1] + b[:, :-1] - b[:, 1:]
    ae = h[:, 1:] + b[:, 1:] - b[:, :-1]
    D[:, :-1] += dx * vy * ac
    D[:, 1:] -= dx * vy * ae
    return D


def constraint(Q, hw, nx, ny, dx, dy, eps=EPS):
    """D of a cell state: v_f = average of the two cells' velocities."""
    b, h, hu, hv = (Q[k].reshape(nx, ny) for k in range(4))
    hi = hinv(h, eps)
    u, v = hu * hi, hv * hi
    vx = 0.5 * (u[:-1] + u[1:])
    vy = 0.5 * (v[:, :-1] + v[:, 1:])
    return constraint_faces(h, b, vx, vy, hw.reshape(nx, ny) * hi, dx, dy)


def pressure_rhs(Q, hw, nx, ny, dx, dy, dt, eps=EPS):
    return -constraint(Q, hw, nx, ny, dx, dy, eps) / dt


def correct(Q, hw, p, nx, ny, dx, dy, dt):
    """hu -= dt * (average of the two x-face B p), same for hv; walls
    contribute 0; hw += 2 dt p.  ``p`` is (nx, ny)."""
    b, h = Q[0].reshape(nx, ny), Q[1].reshape(nx, ny)
    fx, fy = face_bp(h, b, p, dx, dy)
    cx = np.zeros((nx, ny))
    cy = np.zeros((nx, ny))
    cx[:-1] += fx
    cx[1:] += fx
    cy[:, :-1] += fy
    cy[:, 1:] += fy
    out = Q.copy()
    out[2] = Q[2] - dt * 0.5 * cx.ravel()
    out[3] = Q[3] - dt * 0.5 * cy.ravel()
    return out, hw + 2.0 * dt * p.ravel()


def advect_w(Q, hw, nx, ny, dx, dy, dt, eps=EPS):
    """hw - dt/|A| sum_f m_f w_upwind, m_f = |f| (q_c + q_e).n / 2 of the
    stage-start state; walls carry no flux."""
    h = Q[1].reshape(nx, ny)
    hu = Q[2].reshape(nx, ny)
    hv = Q[3].reshape(nx, ny)
    w = hw.reshape(nx, ny) * hinv(h, eps)
    acc = np.zeros((nx, ny))
    mx = 0.5 * (hu[:-1] + hu[1:]) * dy
    fx = mx * np.where(mx > 0, w[:-1], w[1:])
    acc[:-1] += fx
    acc[1:] -= fx
    my = 0.5 * (hv[:, :-1] + hv[:, 1:]) * dx
    fy = my * np.where(my > 0, w[:, :-1], w[:, 1:])
    acc[:, :-1] += fy
    acc[:, 1:] -= fy
    return hw - dt / (dx * dy) * acc.ravel()


# -------------------------------------------------------------- multigrid
def restrict_avg(a):
    return 0.25 * (a[0::2, 0::2] + a[1::2, 0::2] + a[0::2, 1::2]
                   + a[1::2, 1::2])


def _w1(n_fine):
    """Dense 1-D cell-centred bilinear prolongation (n_fine x n_fine/2):
    fine cell i takes 3/4 of parent i//2 and 1/4 of the parent on its other
    side, folded onto the parent at the two ends."""
    nc = n_fine // 2
    P = np.zeros((n_fine, nc))
    for i in range(n_fine):
        I = i >> 1
        J = I + (1 if i & 1 else -1)
        if 0 <= J < nc:
            P[i, I] += 0.75
            P[i, J] += 0.25
        else:
            P[i, I] += 1.0
    return P


class Hierarchy:
    """Levels of (h, b) averages and their stencils."""

    def __init__(self, h, b, dx, dy, min_size=4, eps=EPS):
        self.levels = []
        h = np.asarray(h, np.float64)
        b = np.asarray(b, np.float64)
        while True:
            self.levels.append(dict(coef=coefficients(h, b, dx, dy, eps),
                                    shape=h.shape))
            nx, ny = h.shape
            if nx % 2 or ny % 2 or min(nx, ny) // 2 < min_size:
                break
            h, b = restrict_avg(h), restrict_avg(b)
            dx, dy = 2 * dx, 2 * dy
        for L in self.levels[:-1]:
            nx, ny = L["shape"]
            L["Px"], L["Py"] = _w1(nx), _w1(ny)


def rbgs(coef, x, f, colour):
    diag, cE, cN = coef
    nx, ny = diag.shape
    s = np.zeros_like(x)
    s[:-1] += cE[:-1] * x[1:]
    s[1:] += cE[:-1] * x[:-1]
    s[:, :-1] += cN[:, :-1] * x[:, 1:]
    s[:, 1:] += cN[:, :-1] * x[:, :-1]
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    m = ((I + J) & 1) == colour
    x = x.copy()
    x[m] = (f[m] - s[m]) / diag[m]
    return x


def vcycle(H: Hierarchy, f, level=0, nu=1, coarse_sweeps=2):
    L = H.levels[level]
    coef = L["coef"]
    x = np.zeros_like(f)
    if level == len(H.levels) - 1:
        for _ in range(coarse_sweeps):
            for c in (0, 1, 1, 0):
                x = rbgs(coef, x, f, c)
        return x
    for _ in range(nu):
        for c in (0, 1):
            x = rbgs(coef, x, f, c)
    r = f - apply(coef, x)
    Px, Py = L["Px"], L["Py"]
    fc = Px.T @ r @ Py
    ec = vcycle(H, fc, level + 1, nu, coarse_sweeps)
    x = x + Px @ ec @ Py.T
    for _ in range(nu):
        for c in (1, 0):
            x = rbgs(coef, x, f, c)
    return x


def pcg(coef, rhs, precond=None, tol=1e-8, maxiter=500, x0=None):
    """Preconditioned CG on (nx, ny) fields; returns (x, iterations,
    relative residual history)."""
    b2 = float(np.sum(rhs * rhs))
    x = np.zeros_like(rhs) if x0 is None else x0.copy()
    r = rhs - apply(coef, x) if x0 is not None else rhs.copy()
    if b2 == 0.0:
        return x, 0, [0.0]
    z = precond(r) if precond else r.copy()
    p = z.copy()
    rz = float(np.sum(r * z))
    hist = [np.sqrt(float(np.sum(r * r)) / b2)]
    it = 0
    while it < maxiter and hist[-1] > tol:
        q = apply(coef, p)
        alpha = rz / float(np.sum(p * q))
        x += alpha * p
        r -= alpha * q
        it += 1
        hist.append(np.sqrt(float(np.sum(r * r)) / b2))
        if hist[-1] <= tol:
            break
        z = precond(r) if precond else r.copy()
        rz_new = float(np.sum(r * z))
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it, hist


def mg_solve(h, b, dx, dy, rhs, tol=1e-8, nu=1, coarse_sweeps=2,
             min_size=4, maxiter=200):
    H = Hierarchy(h, b, dx, dy, min_size=min_size)
    return pcg(H.levels[0]["coef"], rhs,
               lambda r: vcycle(H, r, 0, nu, coarse_sweeps), tol, maxiter)


def accept_tol(tol):
    """Accept gate of the pressure solve (solver_chorin_vam_jax.py:72-91)."""
    return max(1e-6, 10.0 * tol)


# ------------------------------------------------------------ Chorin cycle
class ChorinOracle:
    """Chorin projection on an nx x ny wall-bounded quad grid; ``fv`` is the
    FV oracle (``oracle.Oracle``) of the same mesh that supplies the
    hyperbolic predictor."""

    def __init__(self, fv, nx, ny, box, tol=1e-8):
        self.fv = fv
        self.nx, self.ny = nx, ny
        self.dx = (box[1] - box[0]) / nx
        self.dy = (box[3] - box[2]) / ny
        self.tol = tol

    def predictor(self, Q, hw, dt):
        Qs = Q + dt * self.fv.flux(Q)
        return Qs, advect_w(Q, hw, self.nx, self.ny, self.dx, self.dy, dt)

    def pressure(self, Qs, hws, dt, solver="direct"):
        nx, ny = self.nx, self.ny
        h = Qs[1].reshape(nx, ny)
        b = Qs[0].reshape(nx, ny)
        rhs = pressure_rhs(Qs, hws, nx, ny, self.dx, self.dy, dt)
        coef = coefficients(h, b, self.dx, self.dy)
        if solver == "direct":
            p = spla.spsolve(to_sparse(coef).tocsc(), rhs.ravel())
            return p.reshape(nx, ny), rhs
        p, _, _ = mg_solve(h, b, self.dx, self.dy, rhs, tol=self.tol)
        return p, rhs

    def stage(self, Q, hw, dt, solver="direct"):
        Qs, hws = self.predictor(Q, hw, dt)
        p, _ = self.pressure(Qs, hws, dt, solver)
        Qn, hwn = correct(Qs, hws, p, self.nx, self.ny, self.dx, self.dy, dt)
        return Qn, hwn, p

    def cycle(self, Q, hw, dt, time_order=2, solver="direct"):
        Q1, hw1, p1 = self.stage(Q, hw, dt, solver)
        if time_order == 1:
            return Q1, hw1, (p1,)
        Q2, hw2, p2 = self.stage(Q1, hw1, dt, solver)
        return 0.5 * (Q + Q2), 0.5 * (hw + hw2), (p1, p2)


def solitary_wave_state(nx, ny, box, amplitude=0.2, depth=1.0, seed=4,
                        x0_frac=0.25):
    """Solitary wave (sech^2 profile, Boussinesq speed) moving in +x over
    the seeded bed b = 0.1 sin(2 pi (kx x + ky y) + phi); returns (Q, hw)
    with Q (4, N) rows (b, h, hu, hv)."""
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = box
    Lx, Ly = x1 - x0, y1 - y0
    kx = rng.integers(1, 5) / Lx
    ky = rng.integers(1, 5) / Ly
    ph = rng.uniform(0.0, 2.0 * np.pi)
    dx, dy = Lx / nx, Ly / ny
    X, Y = np.meshgrid(x0 + (np.arange(nx) + 0.5) * dx,
                       y0 + (np.arange(ny) + 0.5) * dy, indexing="ij")
    k = np.sqrt(3.0 * amplitude / (4.0 * depth ** 3))
    c = np.sqrt(G * (depth + amplitude))
    eta = amplitude / np.cosh(k * (X - (x0 + x0_frac * Lx))) ** 2
    b = 0.1 * np.sin(2.0 * np.pi * (kx * X + ky * Y) + ph)
    h = depth + eta - b
    Q = np.zeros((4, nx * ny))
    Q[0] = b.ravel()
    Q[1] = h.ravel()
    Q[2] = (c * eta).ravel()
    return Q, np.zeros(nx * ny)
