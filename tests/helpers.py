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


def swe_exact_riemann(hL, uL, hR, uR, g=G):
    """Exact solution of the shallow-water Riemann problem with both sides
    wet (Toro, "Shock-capturing methods for free-surface shallow flows",
    2001, ch. 5): the star depth from the bisection of
    f_L(h) + f_R(h) + u_R - u_L = 0 (shock branch (h - h_k) sqrt(g (h + h_k)
    / (2 h h_k)), rarefaction branch 2 (sqrt(g h) - c_k)).  Returns a sampler
    xi = x / t -> (h, u); ``sample.span`` = the slowest and fastest wave
    speeds, ``sample.us`` = the star (contact) speed."""
    cL, cR = np.sqrt(g * hL), np.sqrt(g * hR)

    def fk(h, hk, ck):
        if h > hk:
            return (h - hk) * np.sqrt(0.5 * g * (h + hk) / (h * hk))
        return 2.0 * (np.sqrt(g * h) - ck)

    def f(h):
        return fk(h, hL, cL) + fk(h, hR, cR) + uR - uL

    a, b = 1e-12, 10.0 * max(hL, hR) + (abs(uL) + abs(uR)) ** 2 / g
    for _ in range(200):
        m = 0.5 * (a + b)
        if f(a) * f(m) <= 0:
            b = m
        else:
            a = m
    hs = 0.5 * (a + b)
    cs = np.sqrt(g * hs)
    us = 0.5 * (uL + uR) + 0.5 * (fk(hs, hR, cR) - fk(hs, hL, cL))
    sl_shock = uL - cL * np.sqrt(0.5 * hs * (hs + hL)) / hL
    sr_shock = uR + cR * np.sqrt(0.5 * hs * (hs + hR)) / hR

    def sample(xi):
        xi = np.asarray(xi, float)
        h = np.empty_like(xi)
        u = np.empty_like(xi)
        left = xi <= us
        if hs > hL:                           # left shock
            m1 = left & (xi < sl_shock)
            m2 = left & (xi >= sl_shock)
            h[m1], u[m1] = hL, uL
            h[m2], u[m2] = hs, us
        else:                                 # left rarefaction
            head, tail = uL - cL, us - cs
            m1 = left & (xi < head)
            m3 = left & (xi > tail)
            m2 = left & ~m1 & ~m3
            h[m1], u[m1] = hL, uL
            h[m3], u[m3] = hs, us
            u[m2] = (uL + 2.0 * cL + 2.0 * xi[m2]) / 3.0
            h[m2] = ((uL + 2.0 * cL - xi[m2]) / 3.0) ** 2 / g
        right = ~left
        if hs > hR:                           # right shock
            m1 = right & (xi > sr_shock)
            m2 = right & (xi <= sr_shock)
            h[m1], u[m1] = hR, uR
            h[m2], u[m2] = hs, us
        else:                                 # right rarefaction
            head, tail = uR + cR, us + cs
            m1 = right & (xi > head)
            m3 = right & (xi < tail)
            m2 = right & ~m1 & ~m3
            h[m1], u[m1] = hR, uR
            h[m3], u[m3] = hs, us
            u[m2] = (uR - 2.0 * cR + 2.0 * xi[m2]) / 3.0
            h[m2] = ((-uR + 2.0 * cR + xi[m2]) / 3.0) ** 2 / g
        return h, u

    sample.span = (sl_shock if hs > hL else uL - cL,
                   sr_shock if hs > hR else uR + cR)
    sample.us = us
    return sample


def ritter_exact(x, t, hL, x0, g=G):
    """Ritter's dry dam break (the reference's tests/test_swe_dambreak_jax.py
    :106-125): depth at points x, time t."""
    cL = np.sqrt(g * hL)
    x = np.asarray(x, float)
    h = np.zeros_like(x)
    h[x < x0 - cL * t] = hL
    fan = (x >= x0 - cL * t) & (x <= x0 + 2.0 * cL * t)
    h[fan] = (2.0 * cL - (x[fan] - x0) / t) ** 2 / (9.0 * g)
    return h


def fractal_valley_bed(xc, seed=3):
    """Seeded fractal bed on [0, 10] x [0, 5] of the wet/dry triangle
    workload: four octaves of random-phase plane waves (amplitude halving,
    frequency doubling), scaled to a maximum of 1.5, plus a linear valley
    slope toward x = 10 and the centre line y = 2.5."""
    rng = np.random.default_rng(seed)
    x, y = xc[:, 0], xc[:, 1]
    f = np.zeros(len(xc))
    for o in range(4):
        for _ in range(6):
            th = rng.uniform(0.0, np.pi)
            k = 2.0 ** o * 0.6 * np.array([np.cos(th), np.sin(th)])
            f += 0.5 ** o * np.cos(k[0] * x + k[1] * y + rng.uniform(0, 6.283))
    f = 1.5 * (f - f.min()) / (f.max() - f.min())
    return f + 0.08 * (10.0 - x) + 0.1 * np.abs(y - 2.5)


def rcm_wetdry_triangles(nx=100, ny=50, seed=3):
    """The wet/dry triangle workload at a reduced size: Delaunay of nx x ny
    jittered points on [0, 10] x [0, 5] (about 2 nx ny triangles), cells in
    RCM order, walls; the fractal valley bed, reservoir h = 3 in x < 2 and
    dry elsewhere."""
    from zoomy_hip import _lib
    from zoomy_hip.mesh import delaunay_mesh, jittered_points
    base = delaunay_mesh(jittered_points(nx, ny, (0.0, 10.0, 0.0, 5.0), 0.3,
                                         seed), (0.0, 10.0, 0.0, 5.0))
    ip, ix = base.adjacency_csr()
    perm = np.zeros(base.n_cells, np.int32)
    _lib.check(_lib.lib().zj_rcm(base.n_cells, ip.ctypes.data, ix.ctypes.data,
                                 perm.ctypes.data), "zj_rcm")
    mesh = base.reorder(perm).set_bc("wall")
    xc = mesh.cell_centers()
    Q = np.zeros((4, mesh.n_cells))
    Q[0] = fractal_valley_bed(xc, seed)
    Q[1] = np.where(xc[:, 0] < 2.0, 3.0, 0.0)
    return mesh, Q
