"""GPU parity of the batched ensemble (many independent instances in one
mesh, per-instance dt, HIP-graph replay): each instance against the FP64
oracle run on that instance alone."""
import math

import numpy as np
import pytest

from helpers import max_rel_rows, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _ensemble(dtype, n=5, seed=9):
    from zoomy_hip.ensemble import (EnsembleSolver, ensemble_spec,
                                    instance_mesh, instance_state)
    spec = ensemble_spec(n, 2000, 20000, seed=seed)
    meshes = [instance_mesh(i) for i in spec]
    states = [instance_state(m, i) for m, i in zip(meshes, spec)]
    es = EnsembleSolver(dtype=dtype)
    es.setup(meshes, states)
    return es, meshes, states


def test_fp64_ensemble_march_equals_each_instance_alone():
    """Every instance of an FP64 ensemble march (own dt per instance) equals
    the oracle's march of that instance alone: state within relative L2
    1e-12, time to 1e-12."""
    from oracle.oracle import Oracle
    es, meshes, states = _ensemble(torch.float64)
    steps = 6
    es.march(steps)
    t = es.t.cpu().numpy()
    parts = [q.cpu().numpy() for q in es.split()]
    for k, (m, Q0) in enumerate(zip(meshes, states)):
        o = Oracle(m, order=2, limiter="minmod", riemann="hll", cfl=0.9)
        Qo, to, ko = o.march(Q0, n_steps=steps)
        assert abs(t[k] - to) <= 1e-12 * to, k
        assert max_rel_rows(parts[k], Qo) < 1e-12, k
    # instances really differ in dt
    assert np.ptp(es.dt.cpu().numpy()) > 0


def test_fp32_ensemble_step_and_dt_against_oracle():
    """FP32 ensemble: per-instance dt against the oracle's (FP32 wave
    speeds: 1e-6), and one step from FP32-representable states against the
    FP64 oracle's step of each instance within relative L2 1e-5 (h and the
    momentum rows, the latter on the momentum scale)."""
    from oracle.oracle import Oracle
    es, meshes, states = _ensemble(torch.float32)
    es.march(4)
    parts0 = [q.double().cpu().numpy() for q in es.split()]
    dts = es.compute_dt().cpu().numpy()
    es.march(1)
    parts1 = [q.double().cpu().numpy() for q in es.split()]
    for k, m in enumerate(meshes):
        o = Oracle(m, order=2, limiter="minmod", riemann="hll", cfl=0.9)
        dto = o.dt(parts0[k])
        assert abs(dts[k] - dto) <= 1e-6 * dto, k
        Qo, _ = o.step(dts[k], parts0[k])
        assert rel_l2(parts1[k][1], Qo[1]) < 1e-5, k
        qs = max(np.linalg.norm(Qo[2]), np.linalg.norm(Qo[3]))
        for r in (2, 3):
            assert np.linalg.norm(parts1[k][r] - Qo[r]) <= 1e-5 * qs, (k, r)


def test_graph_replays_equal_the_eager_march_and_conserve_mass():
    """Three replays of a captured 5-step graph, with eager work and
    synchronisations between them, equal 15 eager steps bit for bit, per
    instance t included; the per-instance compensated mass equals the
    exactly rounded sum and is conserved (walls)."""
    es, meshes, states = _ensemble(torch.float32, n=4, seed=3)
    eager, _, _ = _ensemble(torch.float32, n=4, seed=3)
    vols = [m.cell_volumes() for m in meshes]
    m0 = es.mass().cpu().numpy()
    for k, q in enumerate(es.split()):
        h = q[1].double().cpu().numpy()
        exact = math.fsum((h * vols[k]).tolist())
        # the device recomputes the cell areas from the vertices: last-bit
        # differences of the areas, not of the sum
        assert abs(m0[k] - exact) <= 1e-14 * exact
    Q_init = eager.Q.clone()
    es.capture(5)
    for r in range(3):
        es.replay()
        torch.cuda.synchronize()
        _ = (es.Q[1] * 2.0).sum().item()        # eager work in between
        # the eager march of the same 5 (r + 1) steps from the start
        eager.Q.copy_(Q_init)
        eager.t.zero_()
        eager.march(5 * (r + 1))
        torch.cuda.synchronize()
        assert torch.equal(es.Q, eager.Q), r
        assert torch.equal(es.t, eager.t), r
    m1 = es.mass().cpu().numpy()
    assert np.all(np.abs(m1 - m0) <= 1e-6 * m0)
    assert bool(torch.isfinite(es.Q).all()) and float(es.Q[1].min()) >= 0.0


def test_ensemble_rejects_unsupported_setups():
    """Instances must not share faces, hold more than 64 cells and the
    numerics must be order 2 without MOOD: the C ABI refuses otherwise."""
    import ctypes
    from zoomy_hip import _lib as L
    es, meshes, states = _ensemble(torch.float64, n=3)
    bad = es.offsets.copy()
    bad[1] += 7                    # an offset inside an instance
    h = ctypes.c_void_p()
    assert L.lib().zj_ensemble_create(es._rt.handle, 3, bad.ctypes.data,
                                      ctypes.byref(h)) == L.ZJ_EINVAL
    tiny = np.array([0, 10, es.offsets[-1]], np.int32)
    assert L.lib().zj_ensemble_create(es._rt.handle, 2, tiny.ctypes.data,
                                      ctypes.byref(h)) == L.ZJ_EINVAL
