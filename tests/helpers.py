"""Shared test helpers: the reference's test cases and comparison norms."""
import numpy as np

G = 9.81


def ic_rows(mesh, fn, nrow=4):
    xc = mesh.cell_centers()
    Q = np.zeros((4, mesh.n_cells))
    rows = np.array([np.asarray(fn(x), float) for x in xc]).T
    Q[:rows.shape[0]] = rows
    return Q


# initial conditions of the reference's tests/cases.py
def stoker_ic(x):
    return [0.0, 0.005 if x[0] < 5.0 else 0.001, 0.0]


def ritter_ic(x):
    return [0.0, 0.005 if x[0] < 5.0 else 0.0, 0.0]


def lake_ic(x):
    b = 0.1 * np.exp(-((x[0] - 5.0) ** 2) / 0.5)
    return [b, 0.3 - b, 0.0]


def tilted_ic(x):
    return [0.0, 0.5 + 0.05 * (x[0] - 0.5), 0.0]


def acoustic_ic(x):
    """Mode-1 standing wave of amplitude 1e-4 on h = 1 over [0, 1]."""
    return [0.0, 1.0 + 1e-4 * np.cos(np.pi * x[0]), 0.0]


def bump_lake_ic(x):
    """Lake at rest (eta = 0.3) over a Gaussian bump at x = 5."""
    b = 0.1 * np.exp(-((x[0] - 5.0) ** 2) / 0.5)
    return [b, 0.3 - b, 0.0]


def smooth_state(x):
    """(h, q) of the smooth Dirichlet profile (nonzero slope at both ends)."""
    return (1.0 + 0.1 * np.sin(2.0 * np.pi * x + 0.7),
            0.1 * np.cos(2.0 * np.pi * x))


def smooth_dirichlet_ic(x):
    return [0.0, *smooth_state(float(x[0]))]


def inflow_bc():
    """Dirichlet on h and q_0 at both ends of [0, 1], at the smooth
    profile's own boundary values (tests/cases.py:168-183)."""
    from zoomy_hip.mesh import Dirichlet
    (hl, ql), (hr, qr) = smooth_state(0.0), smooth_state(1.0)
    return {"left": [Dirichlet("h", hl), Dirichlet("q_0", ql)],
            "right": [Dirichlet("h", hr), Dirichlet("q_0", qr)]}


def pulse_ic(x):
    return [0.0, 0.5 + 0.05 * np.exp(-(x[0] ** 2 + x[1] ** 2) / 0.2), 0.0,
            0.0]


def dam1d_ic(x):
    """1-D dam break of the CPU configuration: h = 2 left of 0.5, else 1."""
    return [0.0, 2.0 if x[0] < 0.5 else 1.0, 0.0]


def bumps(n_bumps, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n_bumps, 2))
    a = rng.uniform(0.0, amp, n_bumps)
    w = rng.uniform(0.02, 0.1, n_bumps)

    def b(x):
        x = np.atleast_2d(x)
        r2 = ((x[:, None, :] - c[None]) ** 2).sum(-1)
        return (a[None] * np.exp(-r2 / (2 * w[None] ** 2))).sum(1)
    return b


def circular_dam_state(mesh, seed=2048, n_bumps=32):
    """Circular dam break over seeded Gaussian bumps (h=2 in r<0.25)."""
    xc = mesh.cell_centers()
    b = bumps(n_bumps, seed)(xc)
    r = np.sqrt(((xc - 0.5) ** 2).sum(1))
    Q = np.zeros((4, mesh.n_cells))
    Q[0] = b
    Q[1] = np.where(r < 0.25, 2.0, 1.0)
    return Q


def rel_l2(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def max_rel_rows(a, b, rows=(1, 2, 3)):
    return max(rel_l2(a[r], b[r]) for r in rows)


def wetdry_quads(n=48):
    """Wet/dry case on n x n quads over [0, 10]^2 (walls): reservoir
    h = 3 - b for x < 3, dry elsewhere, bed 0.3 sin(x)^2 + 0.05 y.  At order
    2 with the eta reconstruction and CFL 0.9 its candidate steps have cells
    with h < 0 (the MOOD test case)."""
    import numpy as np
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(n, n, (0, 10, 0, 10)).set_bc("wall")
    xc = mesh.cell_centers()
    Q = np.zeros((4, mesh.n_cells))
    Q[0] = 0.3 * np.sin(xc[:, 0]) ** 2 + 0.05 * xc[:, 1]
    Q[1] = np.where(xc[:, 0] < 3.0, np.maximum(3.0 - Q[0], 0.0), 0.0)
    return mesh, Q
