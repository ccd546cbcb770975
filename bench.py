#!/usr/bin/env python3
"""Benchmark: MUSCL-HLL finite-volume steps on one MI355X.

Workload (the reference's 2-D configuration): shallow water on a 2048 x 2048
quad mesh stored unstructured (4,194,304 cells, 8,392,704 faces), circular
dam break (h = 2 in r < 0.25, else 1, u = v = 0) over 32 seeded Gaussian
bumps (amplitude U[0, 0.3], width U[0.02, 0.1], seed 2048), LSQ-MUSCL with the
minmod limiter, HLL flux with hydrostatic reconstruction, SSP-RK2, reflective
walls, CFL 0.9 (0.45 effective in 2-D); cells numbered along the Morton
curve of their centres.  A step is one full time step:
both RK stages, the dt reduction and the post_step aux refresh, all on the
device.  Inputs are resident in HBM when the timed region starts.

Prints ONE JSON line.  value = cell-updates/s in FP64 (the reference's
precision); the FP32 path is reported beside it under "fp32".
"""
import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "zoomy-jax_amd"))
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12           # MI355X HBM3E, bytes/s
NX = NY = 2048


def log(msg):
    """Progress of the legs, on stderr (stdout carries the one JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr,
          flush=True)


def algorithmic_bytes(w):
    """Algorithmic bytes per cell, quad mesh (k = 4 face neighbours, F/N = 2):
    slope pass 28w + 16 B, face pass 30w + 16 B (+ 3w re-read of the step
    state in stage 2).  One launch of the fused stage kernel does a slope and
    a face pass: stage 1 58w + 32 B, stage 2 61w + 32 B; a cell-update (two
    stages) is 119w + 64 B."""
    return {"stage1": 58 * w + 32, "stage2": 61 * w + 32,
            "cell_update": 119 * w + 64}


def build_state(mesh):
    import numpy as np
    rng = np.random.default_rng(2048)
    c = rng.uniform(0.0, 1.0, (32, 2))
    a = rng.uniform(0.0, 0.3, 32)
    wd = rng.uniform(0.02, 0.1, 32)
    xc = mesh.cell_centers()
    b = np.zeros(mesh.n_cells)
    for k in range(32):
        r2 = ((xc - c[k]) ** 2).sum(1)
        b += a[k] * np.exp(-r2 / (2.0 * wd[k] ** 2))
    r = np.sqrt(((xc - 0.5) ** 2).sum(1))
    Q = np.zeros((4, mesh.n_cells))
    Q[0] = b
    Q[1] = np.where(r < 0.25, 2.0, 1.0)
    return Q


def gpu_leg(mesh, Q0, dtype, steps, warmup):
    import torch
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll", boundary_conditions="wall")
    s = HyperbolicSolver(time_end=float("inf"), cfl=0.9, dtype=dtype)
    s.setup_simulation(mesh, model)
    Q = torch.as_tensor(Q0, dtype=dtype, device="cuda").contiguous()
    Qaux = torch.zeros((7, mesh.n_cells), dtype=dtype, device="cuda")
    s.march_device(Q, warmup, Qaux=Qaux)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, t_dev, dt_dev = s.march_device(Q, steps, Qaux=Qaux)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    troubled = s.troubled_count()
    sim_t = float(t_dev)
    # per-kernel durations with HIP events on the launch stream
    dt = float(dt_dev)
    ms = s.time_kernels(Q, dt, reps=20)
    mass = float(s.compute_mass(Q))
    finite = bool(torch.isfinite(Q).all())
    Qh = Q.double().cpu().numpy()
    del s
    return dict(elapsed=el, ms=ms, sim_t=sim_t, troubled=troubled, mass=mass,
                finite=finite, Q=Qh)


def rel_l2_rows(A, B, rows=(1, 2, 3)):
    """Relative L2 difference of rows h, hu, hv (and hw for Chorin) of A
    against B."""
    import numpy as np
    names = ("b", "h", "hu", "hv", "hw")
    return {names[r]: float(np.linalg.norm(A[r] - B[r]) /
                            max(np.linalg.norm(B[r]), 1e-300)) for r in rows}


def cpu_model():
    """CPU model string of the host (first 'model name' of /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_leg(runs=5):
    """The CPU baseline: the FP64 C oracle built with OpenMP (the same source
    as the serial checker, bit-identical results) on a bounded sample of the
    same workload -- the 256 x 256 mesh, a fixed number of steps per run
    (dt reduction, both RK stages, aux refresh) -- on every core of this
    process's affinity set and on one thread; one warm-up run, then the
    median of `runs` timed runs each."""
    import numpy as np
    from oracle import oracle as OO
    from zoomy_hip.mesh import Mesh
    m = Mesh.quad(256, 256).set_bc("wall")
    Q0 = build_state(m)
    cores = len(os.sched_getaffinity(0))
    out = {"cores": cores, "cpu_model": cpu_model(), "cells": m.n_cells}
    Qref = None
    for threads, steps in ((cores, 400), (1, 40)):
        OO.set_threads(threads)
        o = OO.Oracle(m, omp=True, order=2, limiter="minmod", riemann="hll",
                      cfl=0.9)
        times = []
        for r in range(runs + 1):
            Q = Q0
            t0 = time.perf_counter()
            for _ in range(steps):
                dt = o.dt(Q)
                Q, _ = o.step(dt, Q)
                o.aux(Q)
            if r > 0:
                times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        out[f"threads_{threads}" if threads == 1 else "threads_all"] = {
            "threads": threads, "steps": steps, "runs_s": times,
            "median_s": med, "value": steps * m.n_cells / med}
        if threads == cores:
            Qref, k = Q, steps
    # the FP64 device march over the same sample, checked against the oracle
    import torch
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll", boundary_conditions="wall")
    s = HyperbolicSolver(time_end=float("inf"), cfl=0.9)
    s.setup_simulation(m, model)
    Qd = torch.as_tensor(Q0, dtype=torch.float64, device="cuda").contiguous()
    s.march_device(Qd, k)
    out["gpu_vs_oracle"] = rel_l2_rows(Qd.cpu().numpy(), Qref)
    out["steps_checked"] = k
    del s
    return out


def measured_traffic(w, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated with a plain
    1 GiB copy), or None when no summary exists for it."""
    p = ROOT / "profiles" / "fv_stage_traffic.json"
    try:
        t = json.loads(p.read_text())
        e = t["f64" if w == 8 else "f32"][kernel]
        return (e["fetch_kib"] * t["fetch_scale"] +
                e["write_kib"] * t["write_scale"]) * 1024.0
    except (OSError, KeyError, ValueError):
        return None


def summarize_leg(r, w, N, steps):
    cups = N * steps / r["elapsed"]
    B = algorithmic_bytes(w)
    ms = r["ms"]
    dom = max(("stage1", "stage2"), key=lambda k: ms[k])
    ach = B[dom] * N / (ms[dom] * 1e-3)
    return cups, dict(
        bound="hbm", kernel=dom, achieved=ach / 1e9, peak=PEAK_HBM / 1e9,
        unit="GB/s", frac=ach / PEAK_HBM, traffic=measured_traffic(w, dom),
        bytes_per_launch=B[dom] * N, launch_ms=ms[dom],
        step_alg_bytes_per_cell=B["cell_update"],
        step_frac=cups * B["cell_update"] / PEAK_HBM)


CHORIN_N = 1024
CHORIN_L = 64.0
# algorithmic FLOPs of one network application per 2x2-block node: three
# layers of [h | agg h] (128) x 64 plus the 16 x 64 gate product, the 4 -> 64
# embedding and the 64 -> 4 output
NET_FLOP_PER_NODE = 3 * (2 * 128 * 64 + 2 * 16 * 64) + 2 * 4 * 64 + 2 * 64 * 4


def chorin_state(n=CHORIN_N, L=CHORIN_L):
    """Solitary wave (amplitude 0.2, depth 1) over b = 0.1 sin(2 pi k.x +
    phi), k and phi seeded (seed 4), on [0, L]^2 with walls."""
    import numpy as np
    rng = np.random.default_rng(4)
    kx = rng.integers(1, 5) / L
    ky = rng.integers(1, 5) / L
    ph = rng.uniform(0.0, 2.0 * np.pi)
    dx = L / n
    X, Y = np.meshgrid((np.arange(n) + 0.5) * dx, (np.arange(n) + 0.5) * dx,
                       indexing="ij")
    a, d = 0.2, 1.0
    k = np.sqrt(3.0 * a / (4.0 * d ** 3))
    c = np.sqrt(9.81 * (d + a))
    eta = a / np.cosh(k * (X - 0.25 * L)) ** 2
    b = 0.1 * np.sin(2.0 * np.pi * (kx * X + ky * Y) + ph)
    Q = np.zeros((5, n * n))
    Q[0] = b.ravel()
    Q[1] = (d + eta - b).ravel()
    Q[2] = (c * eta).ravel()
    return Q


def chorin_leg(precond, steps, warmup, net=None, n=CHORIN_N, L=CHORIN_L,
               flux_dtype=None):
    """Fixed-dt Chorin steps (two projection stages each, SSP-RK2) on the
    n x n solitary-wave case; pressure solves to 1e-8 relative residual."""
    import numpy as np
    import torch
    from zoomy_hip.fvm import ChorinSplitVAMSolver, ShallowWater
    from zoomy_hip.fvm.chorin import fixed_dt
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(n, n, (0.0, L, 0.0, L)).set_bc("wall")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll")
    s = ChorinSplitVAMSolver(model, tol=1e-8, precond=precond,
                             flux_dtype=flux_dtype or torch.float64)
    Q = s.setup_simulation(mesh, n, n, chorin_state(n, L))
    if net is not None:
        s.set_preconditioner(precond, net)
    dt = fixed_dt(L / n, 1.0)
    for _ in range(warmup):
        Q = s.chorin_cycle(dt, 0.0, Q)
    torch.cuda.synchronize()
    its, fb, worst = [], 0, 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        Q = s.chorin_cycle(dt, 0.0, Q)
        st = s.stats()
        its.append(st["iterations"] / max(st["solves"], 1))
        fb += st["fallbacks"]
        worst = max(worst, st["relres"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # solve-only timing on the last predicted state
    Qs = s._hyperbolic_stage(dt, 0.0, Q).contiguous()
    s.setup_operator(Qs[0], Qs[1])
    rhs = s.pressure_rhs(Qs, dt)
    s.solve(rhs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        s.solve(rhs)
    torch.cuda.synchronize()
    solve_ms = (time.perf_counter() - t1) * 1e3 / reps
    out = {"ms_per_step": el * 1e3 / steps, "solve_ms": solve_ms,
           "Q": Q.double().cpu().numpy()}
    N = n * n
    if net is not None:
        # one application of the network (fused kernel) on this grid, timed
        # with HIP events around back-to-back launches; the stencil has the
        # operator's shape (h-weighted 5-point plus a reaction term) -- the
        # kernel's work does not depend on its values
        lv = s.levels()
        diag = (Qs[1].float() * 4.0 + 0.5).view(n, n).contiguous()
        cE = (-Qs[1].float()).view(n, n).contiguous()
        cN = (-Qs[1].float()).view(n, n).contiguous()
        ms = net.time_forward(rhs.float().view(n, n), diag, cE, cN, reps=20)
        flop = (N // 4) * NET_FLOP_PER_NODE
        out["network"] = {"ms": ms, "flop_per_application": flop,
                          "tflops": flop / ms / 1e9,
                          "frac_of_bf16_dense_peak": flop / ms / 1e9 / 2500.0,
                          "levels": len(lv)}
    its_mean = float(np.mean(its))
    # FP64 Krylov vectors in every form: 17 words per unknown and iteration
    out.update({
            "cg_bytes_per_unknown_iteration": 17 * 8,
            "solve_effective_gbs": 17 * 8 * N * its_mean / (solve_ms * 1e-3) / 1e9,
            "iterations_per_solve": its_mean,
            "fallbacks": fb, "worst_relres": worst,
            "finite": bool(torch.isfinite(Q).all())})
    return out


DIF_BYTES = 152   # per unknown and CG iteration (see diffusion_leg)


def diffusion_leg(mesh, Q0, reps=5):
    """One Crank-Nicolson diffusion step of the depth field on the same
    2048 x 2048 mesh (IMEX's implicit diffusion, DiffusionOperatorJAX), with
    dt nu / dx^2 = 10: Jacobi-CG to a relative residual of 1e-12.
    Algorithmic bytes per unknown and iteration: the direction update
    fused into M p (z, p_old 16, slot tables 4 x (4 + 8) = 48, volume 8,
    p_new and q written 16 = 88) and the update (p, q, diag, x, r read,
    x, r, z written = 64): 152 B."""
    import torch
    from zoomy_hip.fvm import DiffusionOperator
    dx = 1.0 / NX
    dt = 1e-4
    nu = 10.0 * dx * dx / dt
    op = DiffusionOperator(mesh, 2, nu=nu)
    u = torch.as_tensor(Q0[1], dtype=torch.float64, device="cuda")
    op.implicit_solve(u, dt)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        out = op.implicit_solve(u, dt)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    it = op.stats()["iterations"]
    gbs = DIF_BYTES * mesh.n_cells * it / (ms * 1e-3) / 1e9
    return {"workload": "one CN step of h on the 2048x2048 dam-break mesh, "
                        "dt nu / dx^2 = 10, CG to 1e-12",
            "ms_per_solve": ms, "iterations": it,
            "us_per_iteration": 1e3 * ms / max(it, 1),
            "bytes_per_unknown_iteration": DIF_BYTES,
            "achieved_gbs": gbs, "frac_of_hbm_peak": gbs * 1e9 / PEAK_HBM,
            "relres": op.stats()["relres"],
            "finite": bool(torch.isfinite(out).all())}


def imex_global_leg(mesh, Q0, reps=3):
    """One IMEX step with the global implicit source stage on the 2048 x 2048
    dam-break mesh (the reference's _build_implicit_global,
    solver_imex_jax.py:344-396): explicit MUSCL-HLL stage, then Newton-GMRES
    on R(Q) = Q - Qexp - dt S(Q, aux(Q)) with the Manning friction (n =
    0.03) and the depth-gradient source -alpha g h dh/dx_d (alpha = 0.5),
    which reads the LSQ derivative aux rows; seeded velocities of 0.5 so
    that the friction acts; Newton to 1e-8, GMRES(20) to 1e-7, FP64.  A
    GMRES iteration is one exact Jacobian-vector product (the source
    linearisation on the momentum rows) plus the classical Gram-Schmidt
    work; us_per_gmres_iteration is the whole stage's time over its GMRES
    iterations (the aux refresh, the residuals and the Newton updates
    included)."""
    import numpy as np
    import torch
    from zoomy_hip.fvm import IMEXSourceSolver, ShallowWater
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll", manning_n=0.03,
                         depth_gradient_source=0.5,
                         boundary_conditions="wall")
    s = IMEXSourceSolver(time_end=float("inf"), cfl=0.9)
    s.setup_simulation(mesh, model)
    rng = np.random.default_rng(11)
    Q = Q0.copy()
    Q[2] = Q[1] * rng.uniform(-0.5, 0.5, mesh.n_cells)
    Q[3] = Q[1] * rng.uniform(-0.5, 0.5, mesh.n_cells)
    Qd = torch.as_tensor(Q, dtype=torch.float64, device="cuda").contiguous()
    dt = float(s.compute_timestep(Qd, None))
    out = s.step(dt, 0.0, Qd)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        out = s.step(dt, 0.0, Qd)
    e1.record(st)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / reps
    # the global stage alone, from the explicit stage's output
    Qe = s._hyperbolic_only(dt, s._full(Qd))
    buf = torch.empty_like(Qe)
    import ctypes
    from zoomy_hip import _lib as L
    args = (s._global, float(dt), ctypes.c_void_p(Qe.data_ptr()),
            ctypes.c_void_p(buf.data_ptr()), int(s.implicit_maxiter),
            float(s.implicit_tol), float(s.gmres_tol),
            int(s.gmres_maxiter), ctypes.c_void_p(st.cuda_stream))
    L.check(L.lib().zj_imex_global_solve(*args), "zj_imex_global_solve")
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        L.check(L.lib().zj_imex_global_solve(*args), "zj_imex_global_solve")
    e1.record(st)
    torch.cuda.synchronize()
    ms_stage = e0.elapsed_time(e1) / reps
    stats = s.implicit_stats()
    return {"workload": "one IMEX step on the 2048x2048 dam-break mesh: "
                        "explicit MUSCL-HLL stage + global Newton-GMRES "
                        "source stage (Manning 0.03, depth-gradient "
                        "source 0.5, seeded velocities), FP64",
            "dt": dt, "ms_per_step": ms_step,
            "ms_per_global_stage": ms_stage,
            "newton_steps": stats["newton"], "gmres_iterations": stats["gmres"],
            "us_per_gmres_iteration": 1e3 * ms_stage / max(stats["gmres"], 1),
            "rnorm": stats["rnorm"], "fallbacks": stats["fallbacks"],
            "finite": bool(torch.isfinite(out).all())}


def wetdry_tri_bed(xc, seed=3):
    """Seeded fractal bed on [0, 10] x [0, 5]: four octaves of random-phase
    plane waves (amplitude halving, frequency doubling), scaled to a
    maximum of 1.5, plus a linear valley slope down toward x = 10 and the
    centre line y = 2.5."""
    import numpy as np
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


# algorithmic bytes per cell-update on a triangle mesh (k = 3 face
# neighbours, F/N = 1.5): 104 words + 48 index bytes, FP32 words
WETDRY_BYTES_F32 = 104 * 4 + 48


def wetdry_tri_leg(steps, warmup):
    """The wet/dry triangle configuration: Delaunay of 2000 x 1000 jittered
    points on [0, 10] x [0, 5] (~4M triangles), the bed above, reservoir
    h = 3 in x < 2 and dry elsewhere, walls; order 2 with van Leer on the
    eta reconstruction (hydrostatic face clip), HLL, the a-priori front
    pre-detector (theta < 0.5: first order) and MOOD; FP32 state and
    fluxes, the mass summed in FP64 (compensated).  (MOOD alone does not
    keep this case positive: at CFL 0.9 the dt law 2 r / (d a) exceeds the
    first-order positivity bound |A| / (P a) of a triangle 1.8-fold, so the
    first-order redo itself goes negative.)  Run in the RCM cell order the configuration names and
    in the Morton order of the cell centres (the order the patch kernel is
    laid out for: RCM's level sets are long thin strips, so a patch of
    consecutive cells drags in 8-12x its own size of ring cells, against
    2.5x for Morton).  Returns cell-updates/s and the checks per order."""
    import numpy as np
    import torch
    from zoomy_hip import _lib
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import delaunay_mesh, jittered_points
    t_setup = time.perf_counter()
    pts = jittered_points(2000, 1000, (0.0, 10.0, 0.0, 5.0), 0.3, 11)
    base = delaunay_mesh(pts, (0.0, 10.0, 0.0, 5.0))
    ip, ix = base.adjacency_csr()
    rcm = np.zeros(base.n_cells, np.int32)
    _lib.check(_lib.lib().zj_rcm(base.n_cells, ip.ctypes.data,
                                 ix.ctypes.data, rcm.ctypes.data), "zj_rcm")
    delaunay_s = time.perf_counter() - t_setup
    # the RCM of this mesh against scipy's (a check, not a step of the leg)
    try:
        import numpy as np_
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        A = csr_matrix((np_.ones(len(ix)), ix, ip),
                       shape=(base.n_cells, base.n_cells))
        sp = reverse_cuthill_mckee(A, symmetric_mode=True).astype(np_.int32)
        if sp[-1] == rcm[-1]:
            rcm_vs_scipy = bool(np_.array_equal(sp, rcm))
        else:
            # another start among equal degrees: compare from scipy's start
            alt = np_.zeros(base.n_cells, np_.int32)
            _lib.check(_lib.lib().zj_rcm_from(base.n_cells, ip.ctypes.data,
                                              ix.ctypes.data, int(sp[-1]),
                                              alt.ctypes.data), "zj_rcm_from")
            rcm_vs_scipy = "same start: " + str(bool(np_.array_equal(sp, alt)))
    except ImportError:
        rcm_vs_scipy = None
    out = {"workload": "Delaunay of 2000x1000 jittered points on [0,10]x"
                       "[0,5], fractal bed + valley slope, reservoir h=3 in "
                       "x<2, dry elsewhere, walls; van Leer on the eta "
                       "reconstruction, HLL, front pre-detector 0.5 + MOOD, "
                       "FP32 state, FP64 mass",
           "cells": base.n_cells, "steps": steps,
           "delaunay_and_rcm_s": delaunay_s,
           "rcm_equals_scipy": rcm_vs_scipy}
    log(f"triangulated {base.n_cells} cells in {delaunay_s:.1f} s")
    for name, perm in (("rcm", rcm), ("morton", None)):
        log(f"wet/dry march in {name} order")
        t1 = time.perf_counter()
        mesh = base.reorder(base.morton_permutation() if perm is None
                            else perm).set_bc("wall")
        xc = mesh.cell_centers()
        Q0 = np.zeros((4, mesh.n_cells))
        Q0[0] = wetdry_tri_bed(xc)
        Q0[1] = np.where(xc[:, 0] < 2.0, 3.0, 0.0)
        model = ShallowWater(dimension=2, order=2, limiter="vanleer",
                             riemann="hll", boundary_conditions="wall")
        s = HyperbolicSolver(time_end=float("inf"), cfl=0.9,
                             dtype=torch.float32,
                             reconstruction_variables="eta", mood=True,
                             front_theta_tol=0.5)
        s.setup_simulation(mesh, model)
        setup_s = time.perf_counter() - t1
        Q = torch.as_tensor(Q0, dtype=torch.float32,
                            device="cuda").contiguous()
        m0 = float(s.compute_mass(Q))
        s.march_device(Q, warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, t_dev, _ = s.march_device(Q, steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        m1 = float(s.compute_mass(Q))
        redone, demoted = s.mood_stats()
        out[name] = {
            "value": mesh.n_cells * steps / el, "unit": "cell-updates/s",
            "ms_per_step": el * 1e3 / steps, "setup_s": setup_s,
            "checks": {"sim_time": float(t_dev),
                       "mass_rel_drift": abs(m1 - m0) / m0,
                       "min_h": float(Q[1].min()),
                       "finite": bool(torch.isfinite(Q).all()),
                       "mood_steps_redone": redone,
                       "mood_cells_demoted": demoted}}
        B = WETDRY_BYTES_F32
        out[name]["roofline"] = {
            "bound": "hbm", "alg_bytes_per_cell_update": B,
            "achieved": out[name]["value"] * B / 1e9, "unit": "GB/s",
            "peak": PEAK_HBM / 1e9,
            "frac": out[name]["value"] * B / PEAK_HBM}
        print(json.dumps({"wetdry_tri_" + name: out[name]}), file=sys.stderr,
              flush=True)
        del s, Q
    return out


def ensemble_leg(steps, warmup):
    """The batched ensemble: 256 independent dam breaks of uneven size (cell
    counts log-uniform in [16384, 1048576], ~62M triangles in all; seeded
    jittered-lattice triangle meshes in Morton order, beds of eight seeded
    Gaussian bumps, reservoir free surface h0 ~ U[1, 4] for x < 0.4 and 0.5
    beyond, walls), LSQ-MUSCL minmod + HLL + SSP-RK2, FP32, every instance
    with its own CFL dt; the union of the meshes is one work list of
    patches per stage launch, and `steps` steps are captured in one HIP
    graph and replayed.  Returns cell-updates/s over all instances."""
    import numpy as np
    import torch
    from zoomy_hip.ensemble import (EnsembleSolver, build_instances,
                                    ensemble_spec)
    spec = ensemble_spec(256)
    t0 = time.perf_counter()
    workers = min(16, len(os.sched_getaffinity(0)))
    meshes, states = build_instances(spec, workers=workers)
    t1 = time.perf_counter()
    log(f"ensemble: built {len(meshes)} instances "
        f"({sum(m.n_cells for m in meshes)} cells) in {t1 - t0:.1f} s")
    es = EnsembleSolver(dtype=torch.float32)
    es.setup(meshes, states)
    del meshes, states
    t2 = time.perf_counter()
    log(f"ensemble: device set-up {t2 - t1:.1f} s")
    N = int(es.offsets[-1])
    m0 = es.mass().cpu().numpy()
    es.march(warmup)
    torch.cuda.synchronize()
    te = time.perf_counter()
    es.march(steps)
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - te
    es.capture(steps)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    es.replay()
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - tg
    m1 = es.mass().cpu().numpy()
    t = es.t.cpu().numpy()
    dt = es.dt.cpu().numpy()
    value = N * steps / graph_s
    B = WETDRY_BYTES_F32
    out = {"workload": "256 dam breaks, cell counts log-uniform 16k-1M "
                       "(seeded jittered-lattice triangles, Morton order), "
                       "seeded beds, h0 ~ U[1,4], walls; minmod + HLL + "
                       "SSP-RK2, FP32, per-instance dt",
           "instances": len(spec), "cells": N, "steps": steps,
           "value": value, "unit": "cell-updates/s",
           "ms_per_step": graph_s * 1e3 / steps,
           "eager_ms_per_step": eager_s * 1e3 / steps,
           "host_setup_s": t1 - t0, "device_setup_s": t2 - t1,
           "roofline": {"bound": "hbm", "alg_bytes_per_cell_update": B,
                        "achieved": value * B / 1e9, "unit": "GB/s",
                        "peak": PEAK_HBM / 1e9,
                        "frac": value * B / PEAK_HBM},
           "chorin_per_instance": "not included: the Chorin projection and "
                                  "its multigrid are built on structured "
                                  "quad grids; these instances are "
                                  "unstructured triangle meshes",
           "checks": {"finite": bool(torch.isfinite(es.Q).all()),
                      "min_h": float(es.Q[1].min()),
                      "max_mass_rel_drift": float(np.max(np.abs(m1 - m0) /
                                                         m0)),
                      "sim_time_range": [float(t.min()), float(t.max())],
                      "dt_range": [float(dt.min()), float(dt.max())]}}
    del es
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--chorin-steps", type=int, default=10)
    ap.add_argument("--no-chorin", action="store_true")
    ap.add_argument("--chorin-only", action="store_true")
    ap.add_argument("--no-diffusion", action="store_true")
    ap.add_argument("--no-imex", action="store_true",
                    help="skip the global IMEX source-stage leg")
    ap.add_argument("--imex-only", action="store_true",
                    help="only the global IMEX source-stage leg")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched ensemble leg (256 instances, "
                         "~62M triangles; a minute or two of set-up)")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--no-wetdry", action="store_true",
                    help="skip the wet/dry triangle leg (~4M triangles, "
                         "about a minute of host-side mesh set-up)")
    ap.add_argument("--wetdry-tri", action="store_true",
                    help="only the wet/dry triangle configuration (~4M "
                         "triangles; about a minute of host-side mesh set-up)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 1 GiB (10 times): "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 8,
                          "launches": 10}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the multi-GPU halo exchange "
              "is not implemented; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from zoomy_hip.mesh import Mesh

    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(args.ensemble_steps,
                                                   args.warmup)}))
        return
    if args.wetdry_tri:
        print(json.dumps({"wetdry_tri": wetdry_tri_leg(args.steps,
                                                       args.warmup)}))
        return
    if args.imex_only:
        args.no_chorin = True
    chorin = None
    if not args.no_chorin:
        log("chorin legs (plain, mixed, learned, learned_guess)")
        chorin = {"workload": f"{CHORIN_N}x{CHORIN_N} solitary wave "
                              "(amplitude 0.2, depth 1) over a seeded "
                              "sinusoidal bed, fixed dt, SSP-RK2 projection, "
                              "pressure to 1e-8 rel. residual",
                  "plain": chorin_leg("mg", args.chorin_steps, 2)}
        from zoomy_hip.learned import MPPreconditioner
        # mixed precision without the network: FP32 fluxes, FP32 multigrid,
        # FP64 flexible CG (residual, Krylov vectors)
        chorin["mixed"] = chorin_leg("mixed", args.chorin_steps, 2,
                                     flux_dtype=torch.float32)
        # the learned form: as mixed, plus one BF16 network correction per
        # V-cycle
        chorin["learned"] = chorin_leg("learned", args.chorin_steps, 2,
                                       net=MPPreconditioner.load(),
                                       flux_dtype=torch.float32)
        # the learned V-cycle once per solve, as the initial guess of the
        # FP64 V-cycle CG (retried from zero if it misses the tolerance),
        # with the FP32 predictor fluxes
        chorin["learned_guess"] = chorin_leg("learned_guess",
                                             args.chorin_steps, 2,
                                             net=MPPreconditioner.load(),
                                             flux_dtype=torch.float32)
        chorin["learned_over_plain"] = (chorin["learned"]["ms_per_step"] /
                                        chorin["plain"]["ms_per_step"])
        chorin["learned_guess_over_plain"] = (
            chorin["learned_guess"]["ms_per_step"] /
            chorin["plain"]["ms_per_step"])
        chorin["mixed_over_plain"] = (chorin["mixed"]["ms_per_step"] /
                                      chorin["plain"]["ms_per_step"])
        chorin["mixed"].pop("Q")
        # the learned, mixed-precision march against the plain FP64 one
        # after the same steps (pressure to 1e-8 in both, FP32 fluxes)
        Qplain = chorin["plain"].pop("Q")
        chorin["learned_vs_plain_rel_l2"] = rel_l2_rows(
            chorin["learned"].pop("Q"), Qplain, rows=(1, 2, 3, 4))
        chorin["learned_guess_vs_plain_rel_l2"] = rel_l2_rows(
            chorin["learned_guess"].pop("Q"), Qplain, rows=(1, 2, 3, 4))
        if args.chorin_only:
            print(json.dumps({"chorin": chorin}))
            return

    mesh = Mesh.quad(NX, NY)
    # cells along the Morton curve: each 256-cell block is a 16 x 16 tile
    mesh = mesh.reorder(mesh.morton_permutation()).set_bc("wall")
    Q0 = build_state(mesh)
    N = mesh.n_cells
    if args.imex_only:
        print(json.dumps({"imex_global": imex_global_leg(mesh, Q0)}))
        return
    log("FP64 and FP32 marches on the 2048x2048 quad mesh")
    r64 = None if args.no_fp64 else gpu_leg(mesh, Q0, torch.float64,
                                             args.steps, args.warmup)
    r32 = None if args.no_fp32 else gpu_leg(mesh, Q0, torch.float32,
                                             args.steps, args.warmup)
    if r64 is None:
        cups32, roof32 = summarize_leg(r32, 4, N, args.steps)
        print(json.dumps({"fp32_only": True, "value": cups32,
                          "roofline": roof32, "kernel_ms": r32["ms"]}))
        return
    if not args.no_cpu:
        log("CPU baseline (OpenMP oracle)")
    cpu = None if args.no_cpu else cpu_leg()

    def summarize(r, w):
        return summarize_leg(r, w, N, args.steps)

    cups64, roof64 = summarize(r64, 8)
    out = {
        "metric": "cell-updates/s MUSCL-HLL",
        "value": cups64,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2-D SWE circular dam break over 32 Gaussian "
                               "bumps, 2048x2048 quads stored unstructured, "
                               "LSQ-MUSCL minmod + HLL + SSP-RK2, walls",
                   "cells": N, "faces": 8392704, "parallelism": "single"},
        "roofline": roof64,
        "kernel_ms": r64["ms"],
        "checks": {"finite": r64["finite"], "troubled": r64["troubled"],
                   "sim_time": r64["sim_t"], "mass": r64["mass"]},
    }
    if cpu is not None:
        out["checks"]["gpu_vs_oracle_rel_l2"] = cpu["gpu_vs_oracle"]
        out["checks"]["gpu_vs_oracle_sample"] = (
            f"256x256 sample of the same workload, {cpu['steps_checked']} "
            f"steps")
    if not args.no_diffusion:
        log("diffusion leg")
        out["diffusion"] = diffusion_leg(mesh, Q0)
    if not args.no_imex:
        log("global IMEX source-stage leg")
        out["imex_global"] = imex_global_leg(mesh, Q0)
    if not args.no_wetdry:
        log("wet/dry triangle leg (Delaunay of 2M points, RCM and Morton)")
        out["wetdry_tri"] = wetdry_tri_leg(args.steps, args.warmup)
    if not args.no_ensemble:
        log("batched ensemble leg (256 instances)")
        out["ensemble"] = ensemble_leg(args.ensemble_steps, args.warmup)
    if r32 is not None:
        cups32, roof32 = summarize(r32, 4)
        out["fp32"] = {"value": cups32,
                       "ms_per_step": r32["elapsed"] * 1e3 / args.steps,
                       "roofline": roof32, "kernel_ms": r32["ms"],
                       "checks": {"finite": r32["finite"],
                                  "troubled": r32["troubled"],
                                  "mass": r32["mass"],
                                  "rel_l2_vs_fp64": rel_l2_rows(r32["Q"],
                                                                r64["Q"])}}
    if chorin is not None:
        out["chorin"] = chorin
    if cpu is not None:
        ca, c1 = cpu["threads_all"], cpu["threads_1"]
        out["cpu_baseline"] = {
            "value": ca["value"], "unit": "cell-updates/s",
            "cores": cpu["cores"], "kind": "port",
            "cpu_model": cpu["cpu_model"],
            "value_1_thread": c1["value"],
            "sample": f"FP64 C oracle (OpenMP build, bit-identical to the "
                      f"serial checker), same workload on a 256x256 mesh: "
                      f"median of 5 runs of {ca['steps']} steps on "
                      f"{cpu['cores']} threads ({ca['median_s']:.2f} s) and "
                      f"of {c1['steps']} steps on 1 thread "
                      f"({c1['median_s']:.2f} s), each after a warm-up run",
            "runs_s": {"all": ca["runs_s"], "one": c1["runs_s"]}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
