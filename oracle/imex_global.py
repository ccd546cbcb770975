"""Global implicit source stage, restated for the oracle -- TEST
INFRASTRUCTURE ONLY (tests/ may import it, as the checker).

Newton's method on R(Q) = Q - Qexp - dt S(Q, aux(Q)) (the reference's
_build_implicit_global, zoomy_jax/fvm/solver_imex_jax.py:344-396) with a
dense Jacobian by central differences of R itself, so the root is fixed by R
alone and no derivative formula of the device path is reused.  aux(Q) is the
oracle's post-step aux refresh (LSQ derivative rows, zjo_update_aux); the
source rows q_d are the Manning friction and the depth-gradient term
-alpha g h dh/dx_d.
"""
import numpy as np


def hinv(h, eps):
    m = np.maximum(h, eps)
    return np.sqrt(2.0) * h / np.sqrt(h ** 4 + m ** 4)


def source(o, Q, aux, alpha, g=9.81, n=0.0, eps=1e-8):
    """S rows (q_x, q_y) of state Q (4, N) with aux rows (3 d + 1, N)."""
    d = o.rm.mesh.dim
    hi = hinv(Q[1], eps)
    qq = np.sqrt(Q[2] ** 2 + Q[3] ** 2)
    m = g * n * n * hi ** (7.0 / 3.0) * qq
    S = np.zeros((2, Q.shape[1]))
    S[0] = -alpha * g * Q[1] * aux[1] - m * Q[2]
    S[1] = (-alpha * g * Q[1] * aux[3] if d == 2 else 0.0) - m * Q[3]
    return S


def residual(o, Q, Qexp, dt, alpha, n=0.0, eps=1e-8):
    S = source(o, Q, o.aux(Q), alpha, n=n, eps=eps)
    R = Q - Qexp
    R[2] -= dt * S[0]
    R[3] -= dt * S[1]
    return R


def newton(o, Qexp, dt, alpha, n=0.0, tol=1e-12, maxiter=12, fd=1e-6):
    """Newton with the dense central-difference Jacobian of R over the
    unknowns (all four rows); returns (Q, residual norms per step)."""
    Q = Qexp.copy()
    N = Q.shape[1]
    norms = []
    for _ in range(maxiter):
        R = residual(o, Q, Qexp, dt, alpha, n)
        norms.append(float(np.linalg.norm(R)))
        if norms[-1] <= tol:
            break
        J = np.zeros((4 * N, 4 * N))
        for j in range(4 * N):
            E = np.zeros(4 * N)
            E[j] = fd * max(1.0, abs(Q.reshape(-1)[j]))
            Rp = residual(o, Q + E.reshape(4, N), Qexp, dt, alpha, n)
            Rm = residual(o, Q - E.reshape(4, N), Qexp, dt, alpha, n)
            J[:, j] = ((Rp - Rm) / (2.0 * E[j])).reshape(-1)
        Q = Q + np.linalg.solve(J, -R.reshape(-1)).reshape(4, N)
    return Q, norms
