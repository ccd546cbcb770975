"""GPU parity of the block contract (zoomy_jax/fvm/solver2/blocks.py) and
the per-face operator API: the reconstruction object, dt_pass / reduce_dt,
halo_bc, flux_pass, gather_update and the block-driven march, each through
the C ABI against the FP64 oracle (or against a composition of the oracle's
face numerics in numpy where the oracle has no block of that name)."""
import numpy as np
import pytest

from helpers import (circular_dam_state, ic_rows, rel_l2, stoker_ic,
                     wetdry_quads)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(kind):
    from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points
    if kind == "quad":
        mesh = Mesh.quad(24, 20).set_bc({"left": "wall", "right":
                                         "extrapolation", "bottom": "wall",
                                         "top": "wall"})
        Q = circular_dam_state(mesh, seed=3)
    elif kind == "periodic":
        mesh = Mesh.quad(16, 16).set_bc("periodic")
        Q = circular_dam_state(mesh, seed=4)
    elif kind == "tri":
        mesh = delaunay_mesh(jittered_points(18, 14, (0, 1, 0, 1), 0.3, 2),
                             (0, 1, 0, 1)).set_bc("wall")
        Q = circular_dam_state(mesh, seed=5)
    elif kind == "wetdry":
        mesh, Q = wetdry_quads(20)
    else:
        mesh = Mesh.line(60, 0.0, 10.0).set_bc("extrapolation")
        Q = ic_rows(mesh, stoker_ic)
    rng = np.random.default_rng(7)
    Q = Q.copy()
    Q[2] += 0.05 * rng.normal(size=mesh.n_cells) * Q[1]
    if mesh.dim == 2:
        Q[3] += 0.05 * rng.normal(size=mesh.n_cells) * Q[1]
    return mesh, Q


def _solver(mesh, recon, **kw):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    pos = kw.pop("positivity", "")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9,
                         reconstruction_variables=recon,
                         positivity_method=pos)
    s.setup_simulation(mesh, ShallowWater(dimension=mesh.dim, **kw))
    return s


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


@pytest.mark.parametrize("kind,recon,kw", [
    ("quad", "conservative", dict(order=2, limiter="minmod", riemann="hll")),
    ("quad", "conservative", dict(order=1, riemann="rusanov_nc")),
    ("periodic", "conservative", dict(order=2, limiter="vanleer",
                                      riemann="rusanov")),
    ("tri", "xz", dict(order=2, limiter="venkatakrishnan", riemann="hll")),
    ("wetdry", "eta", dict(order=2, limiter="vanleer", riemann="hll",
                           positivity="zhang_shu")),
    ("line", "eta_hu", dict(order=2, limiter="minmod", riemann="hll")),
])
def test_reconstruction_object_and_flux_pass(kind, recon, kw):
    """Face states Q_L, Q_R (reference face order), the limited gradient of
    reconstruct_with_grad and the per-face (F, Dp, Dm) against the oracle;
    the force_o1 mask of the eta reconstruction demotes exactly the masked
    cells."""
    from oracle.oracle import Oracle, face_flux
    from zoomy_hip.fvm import blocks as B
    mesh, Q = _case(kind)
    o = Oracle(mesh, recon=recon, **{**kw, "positivity":
                                     kw.get("positivity", "")})
    s = _solver(mesh, recon, **dict(kw))
    QLo, QRo, lgo = o.reconstruct(Q)
    rec = s.get_reconstruction()
    QL, QR, lg = rec.reconstruct_with_grad(torch.as_tensor(Q, device="cuda"))
    n = s.n_state
    rows = [r for r in range(n)]
    assert _close(QL.cpu().numpy(), QLo[rows])
    assert _close(QR.cpu().numpy(), QRo[rows])
    d = mesh.dim
    if kw["order"] >= 2:
        assert _close(lg.cpu().numpy(), lgo[rows][:, :d])
    # flux_pass against the oracle's face numerics on the oracle's faces
    meshrt = s._faces()
    ops = B.build_operators(s)
    Q4 = s._full(torch.as_tensor(Q, device="cuda"))
    QL4, QR4, _ = B.reconstruct(Q4, None, meshrt, ops, None)
    Fx, Dp, Dm = B.flux_pass(Q4, None, None, meshrt, ops, 0.0, (QL4, QR4))
    Fx, Dp, Dm = (a.cpu().numpy() for a in (Fx, Dp, Dm))
    fc = o.rm.face_cells
    nrm = o.rm.face_normals
    hinv = np.array([o.L.zjo_hinv(h, 1e-8) for h in Q[1]])
    Fw, Dpw, Dmw = np.zeros_like(Fx), np.zeros_like(Fx), np.zeros_like(Fx)
    for f in range(fc.shape[1]):
        A, Bc = fc[0, f], fc[1, f]
        # a boundary face's Q_R is BC(Q_L) (the partner cell at a seam)
        qr = QRo[:, f] if Bc >= 0 else _ghost(mesh, o, f, QLo[:, f], Q)
        hb = hinv[Bc] if Bc >= 0 else hinv[A]
        F_, Dp_, Dm_ = face_flux(QLo[:, f], qr, hinv[A], hb, nrm[:, f],
                                 riemann=kw["riemann"])
        Fw[:, f], Dpw[:, f], Dmw[:, f] = F_, Dp_, Dm_
    # one scale for the three arrays: the fluctuations vanish on flat beds,
    # where the device's fused multiply-adds leave ~1e-20 of the products
    scale = max(np.abs(Fw).max(), np.abs(Dpw).max(), np.abs(Dmw).max())
    for a, b in ((Fx, Fw), (Dp, Dpw), (Dm, Dmw)):
        assert np.abs(a - b).max() <= 1e-11 * scale
    if recon == "eta":
        # force_o1: masked cells reconstruct at first order
        mask = np.zeros(mesh.n_cells, np.uint8)
        mask[::3] = 1
        QLm, QRm = rec(torch.as_tensor(Q, device="cuda"), None,
                       force_o1=torch.as_tensor(mask))
        QLm, QRm = QLm.cpu().numpy(), QRm.cpu().numpy()
        for f in range(fc.shape[1]):
            A, Bc = fc[0, f], fc[1, f]
            if mask[A]:
                # (b, eta - b, u (eta - b)) of the cell: its state up to
                # the rounding of eta - b
                np.testing.assert_allclose(QLm[:n, f], Q[:n, A], rtol=1e-13,
                                           atol=1e-13)


def _ghost(mesh, o, f, q, Q):
    """Oracle ghost of face state q at boundary face f (BC(Q_L))."""
    from oracle.oracle import lib
    import ctypes
    b = int(np.nonzero(o.rm.bface_face == f)[0][0])
    kind = int(o.rm.bface_kind[b])
    if kind == 2:                                  # periodic: partner cell
        pc = o.rm.bface_cell[o.rm.bface_partner[b]]
        return Q[:, pc]
    n = np.ascontiguousarray(o.rm.face_normals[:, f])
    q = np.ascontiguousarray(q, np.float64)
    out = np.zeros(4)
    lib().zjo_bc_ghost(kind, q.ctypes.data_as(ctypes.c_void_p),
                       n.ctypes.data_as(ctypes.c_void_p),
                       out.ctypes.data_as(ctypes.c_void_p))
    return out


@pytest.mark.parametrize("kind", ["quad", "tri", "line", "wetdry"])
def test_dt_pass_reduce_dt_and_halo_bc(kind):
    """Per-face eigenvalue bounds (max |lambda| = |u.n| + c over both cell
    states, the owner on both sides of a boundary face), reduce_dt equal to
    the oracle's dt law, the wave-free and clamp caps, and halo_bc's ghost
    values against the oracle's boundary values."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import blocks as B
    mesh, Q = _case(kind)
    o = Oracle(mesh, order=2, riemann="hll", cfl=0.9)
    s = _solver(mesh, "conservative", order=2, riemann="hll")
    meshrt = s._faces()
    Qd = torch.as_tensor(Q, device="cuda")
    lo, hi = B.dt_pass(Qd, None, None, meshrt)
    amax = s.get_compute_max_abs_eigenvalue()(Qd, None).cpu().numpy()
    fc = o.rm.face_cells
    nrm = o.rm.face_normals
    hinv = np.array([o.L.zjo_hinv(h, 1e-8) for h in Q[1]])

    def lam(c, f):
        un = (Q[2, c] * nrm[0, f] + Q[3, c] * nrm[1, f]) * hinv[c]
        return un - np.sqrt(9.81 * max(Q[1, c], 0)), \
            un + np.sqrt(9.81 * max(Q[1, c], 0))
    want_lo = np.zeros(fc.shape[1])
    want_hi = np.zeros(fc.shape[1])
    for f in range(fc.shape[1]):
        A, Bc = fc[0, f], fc[1, f]
        Bc = A if Bc < 0 else Bc
        la, ha = lam(A, f)
        lb, hb = lam(Bc, f)
        want_lo[f] = min(0.0, la, lb)
        want_hi[f] = max(0.0, ha, hb)
    assert _close(lo.cpu().numpy(), want_lo, 1e-13)
    assert _close(hi.cpu().numpy(), want_hi, 1e-13)
    assert _close(amax, np.maximum(np.abs(want_lo), np.abs(want_hi)), 1e-13)
    dt = B.reduce_dt(lo, hi, meshrt.inradius_f, None, CFL=0.9,
                     dimension=mesh.dim, MeshRT=meshrt)
    assert abs(float(dt) - o.dt(Q)) <= 1e-13 * o.dt(Q)
    d2 = B.reduce_dt(lo, hi, meshrt.inradius_f, None, CFL=0.9,
                     dimension=mesh.dim, degree=1, MeshRT=meshrt)
    assert abs(float(d2) * 3.0 - o.dt(Q)) <= 1e-13 * o.dt(Q)
    d3 = B.reduce_dt(lo, hi, meshrt.inradius_f, 1e-9, CFL=0.9,
                     dimension=mesh.dim, t_remaining=5e-10, MeshRT=meshrt)
    assert float(d3) == 5e-10
    with pytest.raises(ValueError):
        B.reduce_dt(lo, hi, meshrt.inradius_f, None, CFL=0.9, dimension=2,
                    t_remaining=1.0, dt_window=1.0, MeshRT=meshrt)
    # wave-free: dt falls back to dt_max
    z = torch.zeros_like(lo)
    assert float(B.reduce_dt(z, z, meshrt.inradius_f, 0.25, CFL=0.9,
                             dimension=2, MeshRT=meshrt)) == 0.25
    with pytest.raises(FloatingPointError):
        B.assert_dt_admissible(torch.zeros((), device="cuda"), lo, hi,
                               meshrt.inradius_f, time=0.0, iteration=3)
    # halo_bc: the oracle's reconstruct at order 1 puts the ghosts of the
    # cell states into Q_R of the boundary faces
    _, _, bf = B.halo_bc(Qd, None, None, meshrt, 0.0)
    o1 = Oracle(mesh, order=1, riemann="hll")
    _, QR1, _ = o1.reconstruct(Q)
    want = QR1[:, o.rm.bface_face]
    assert _close(bf.cpu().numpy()[:, :mesh.n_bfaces], want, 1e-15)


def _oracle_stage_loop(mesh, o, Q0, Qaux_hinv, dt, g=9.81, manning=0.0,
                       o1=False):
    """The solver2 stage loop (march.py:126-157) composed in numpy from the
    oracle's reconstruction and face numerics: SSP-RK2, the step-start aux
    1/h in both stages, the interior term -g h grad(b)_lim, Manning source of
    the stage base."""
    from oracle.oracle import Oracle, face_flux
    fc = o.rm.face_cells
    nrm = o.rm.face_normals
    fv = o.rm.face_volumes
    cv = o.rm.cell_volumes
    Qk = Q0.copy()
    troubled = np.zeros(mesh.n_cells, bool)
    for alpha, beta in ((0.0, 1.0), (0.5, 1.0)):
        oo = o if not o1 else Oracle(mesh, order=1, riemann="hll")
        QL, QR, lg = oo.reconstruct(Qk)
        dQ = np.zeros_like(Qk)
        for f in range(fc.shape[1]):
            A, Bc = fc[0, f], fc[1, f]
            qr = QR[:, f] if Bc >= 0 else _ghost(mesh, o, f, QL[:, f], Qk)
            hb = Qaux_hinv[Bc] if Bc >= 0 else Qaux_hinv[A]
            F_, Dp_, Dm_ = face_flux(QL[:, f], qr, Qaux_hinv[A], hb,
                                     nrm[:, f], riemann="hll")
            dQ[:, A] -= (F_ + Dm_) * (fv[f] / cv[A])
            if Bc >= 0:
                dQ[:, Bc] -= (-F_ + Dp_) * (fv[f] / cv[Bc])
        if not o1:
            dQ[2] -= g * Qk[1] * lg[0, 0]
            dQ[3] -= g * Qk[1] * lg[0, 1]
        if manning > 0:
            u, v = Qk[2] * Qaux_hinv, Qk[3] * Qaux_hinv
            sp = np.sqrt(u * u + v * v)
            k = g * manning ** 2 * np.cbrt(Qaux_hinv)
            dQ[2] += -k * u * sp
            dQ[3] += -k * v * sp
        Qc = Qk + beta * dt * dQ
        troubled |= (Qc[1] < 0) | ~np.isfinite(Qc).all(0)
        Qk = Qc if alpha == 0.0 else alpha * Q0 + (1 - alpha) * Qc
    return Qk, troubled


@pytest.mark.parametrize("manning", [0.0, 0.03])
def test_block_march_against_oracle_composition(manning):
    """Three steps of the block-driven march (MarchSolver: dt_pass +
    reduce_dt + guard, stage loop of halo_bc / reconstruct / flux_pass /
    gather_update with the interior term and the Manning source, aux
    refresh) against the same composition of the oracle's blocks."""
    from oracle.oracle import Oracle
    from zoomy_hip.fvm import ShallowWater
    from zoomy_hip.fvm.blocks import MarchSolver
    mesh, Q = _case("quad")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll", manning_n=manning)
    ms = MarchSolver(time_end=np.inf, cfl=0.9)
    Qd, Qaux = ms.setup_simulation(mesh, model, Q)
    o = Oracle(mesh, order=2, limiter="minmod", riemann="hll", cfl=0.9)
    Qo = Q.copy()
    for _ in range(3):
        hinv = np.array([o.L.zjo_hinv(h, 1e-8) for h in Qo[1]])
        dt = o.dt(Qo)
        Qo, _ = _oracle_stage_loop(mesh, o, Qo, hinv, dt, manning=manning)
    Qd, Qaux, st = ms.march(Qd, Qaux, n_steps=3)
    assert st["steps"] == 3
    Qg = Qd.cpu().numpy()
    for r in (1, 2, 3):
        assert rel_l2(Qg[r], Qo[r]) < 1e-12, r
    assert rel_l2(Qaux.cpu().numpy()[-1], 1.0 / Qo[1]) < 1e-12


def test_block_march_mood_redo_and_flat_bed_matches_fused_step():
    """On a flat bed the block march (frozen step-start aux in both stages,
    as the solver2 stage loop hands its aux) equals the fused step with the
    frozen stage aux; on the wet/dry case with MOOD the troubled steps are
    redone at first order and the depth stays non-negative."""
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.fvm.blocks import MarchSolver
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(32, 32).set_bc("wall")
    Q = circular_dam_state(mesh, seed=9)
    Q[0] = 0.0
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll")
    ms = MarchSolver(time_end=np.inf, cfl=0.9)
    Qd, Qaux = ms.setup_simulation(mesh, model, Q)
    Qb, _, _ = ms.march(Qd, Qaux, n_steps=4)
    fused = HyperbolicSolver(time_end=np.inf, cfl=0.9, stage_aux="frozen")
    fused.setup_simulation(mesh, model)
    Qf = torch.as_tensor(Q, device="cuda")
    for _ in range(4):
        Qf = fused.step(fused.compute_timestep(Qf), 0.0, Qf)
    assert max(rel_l2(Qb.cpu().numpy()[r], Qf.cpu().numpy()[r])
               for r in (1, 2, 3)) < 1e-13
    mesh2, Q2 = wetdry_quads(24)
    ms2 = MarchSolver(time_end=np.inf, cfl=0.9, mood_redo=True,
                      reconstruction_variables="eta")
    Qd2, Qa2 = ms2.setup_simulation(mesh2, ShallowWater(
        dimension=2, order=2, limiter="vanleer", riemann="hll"), Q2)
    m0 = float((Qd2[1].cpu().numpy() * mesh2.cell_volumes()).sum())
    Qd2, Qa2, st = ms2.march(Qd2, Qa2, n_steps=6)
    assert st["troubled_steps"] >= 1
    assert float(Qd2[1].min()) >= -1e-20
    m1 = float((Qd2[1].cpu().numpy() * mesh2.cell_volumes()).sum())
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.parametrize("recon", ["eta", "conservative"])
def test_flux_operator_force_o1(recon):
    """flux_operator(..., force_o1=mask): the eta reconstruction demotes the
    masked cells (against the oracle with the same mask); the conservative
    one ignores the mask, as the reference's flux operator does for a
    reconstruction without supports_force_o1 (solver_jax.py:580-582)."""
    from oracle.oracle import Oracle
    mesh, Q = _case("wetdry")
    kw = dict(order=2, limiter="vanleer", riemann="hll")
    s = _solver(mesh, recon, **dict(kw))
    o = Oracle(mesh, recon=recon, **kw)
    mask = np.zeros(mesh.n_cells, np.uint8)
    mask[(np.arange(mesh.n_cells) % 5) == 2] = 1
    got = s.get_flux_operator()(0.0, 0.0, torch.as_tensor(Q, device="cuda"),
                                None, force_o1=torch.as_tensor(mask))
    got = got.cpu().numpy()
    want = o.flux_o1(Q, mask) if recon == "eta" else o.flux(Q)
    assert rel_l2(got[1], want[1]) < 1e-12
    qs = max(np.linalg.norm(want[2]), np.linalg.norm(want[3]))
    for r in (2, 3):
        assert np.linalg.norm(got[r] - want[r]) <= 1e-12 * qs
    if recon == "eta":
        assert rel_l2(want[1], o.flux(Q)[1]) > 1e-6   # the mask acted


@pytest.mark.parametrize("riemann", ["hll", "rusanov", "rusanov_nc"])
@pytest.mark.parametrize("dim", [1, 2])
def test_runtime_numerical_flux_slots(riemann, dim):
    """The runtime slots numerical_flux / numerical_fluctuations on random
    face states, cell aux and unit normals against the oracle's face
    numerics (with the aux 1/h, and with the face states' 1/h when no aux is
    given), FP64 to round-off."""
    from oracle.oracle import face_flux, lib
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = (Mesh.quad(8, 8) if dim == 2 else Mesh.line(16)).set_bc("wall")
    s = HyperbolicSolver(time_end=np.inf)
    s.setup_simulation(mesh, ShallowWater(dimension=dim, order=2,
                                          limiter="minmod", riemann=riemann))
    rng = np.random.default_rng(17)
    n = 300
    nv = 2 + dim
    qL = np.zeros((nv, n))
    qR = np.zeros((nv, n))
    for q in (qL, qR):
        q[0] = rng.uniform(0.0, 0.3, n)
        q[1] = rng.uniform(0.0, 2.0, n) * (rng.uniform(size=n) > 0.1)
        q[2:] = rng.normal(0.0, 0.5, (dim, n)) * q[1]
    ang = rng.uniform(0, 2 * np.pi, n)
    nrm = (np.array([np.cos(ang), np.sin(ang)]) if dim == 2
           else np.sign(rng.uniform(-1, 1, (1, n))))
    naux = 3 * dim + 1
    xL = rng.normal(0.0, 1.0, (naux, n))
    xR = rng.normal(0.0, 1.0, (naux, n))
    xL[-1] = [lib().zjo_hinv(h, 1e-8) for h in rng.uniform(0.1, 2.0, n)]
    xR[-1] = [lib().zjo_hinv(h, 1e-8) for h in rng.uniform(0.1, 2.0, n)]
    rt = s.runtime
    for aux in (True, False):
        F = rt.numerical_flux(qL, qR, xL if aux else None,
                              xR if aux else None, None, nrm)
        D = rt.numerical_fluctuations(qL, qR, xL if aux else None,
                                      xR if aux else None, None, nrm)
        F, D = F.cpu().numpy(), D.cpu().numpy()
        assert F.shape == (nv, n) and D.shape == (2, nv, n)
        for f in range(n):
            a = np.zeros(4)
            b = np.zeros(4)
            a[:nv], b[:nv] = qL[:, f], qR[:, f]
            hA = xL[-1, f] if aux else lib().zjo_hinv(a[1], 1e-8)
            hB = xR[-1, f] if aux else lib().zjo_hinv(b[1], 1e-8)
            nn = np.array([nrm[0, f], nrm[1, f] if dim == 2 else 0.0])
            Fo, Dpo, Dmo = face_flux(a, b, hA, hB, nn, riemann=riemann)
            sc = max(np.abs(Fo).max(), np.abs(Dpo).max(), np.abs(Dmo).max(),
                     1.0)
            assert np.abs(F[:, f] - Fo[:nv]).max() <= 1e-12 * sc
            assert np.abs(D[0, :, f] - Dpo[:nv]).max() <= 1e-12 * sc
            assert np.abs(D[1, :, f] - Dmo[:nv]).max() <= 1e-12 * sc
