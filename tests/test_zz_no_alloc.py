"""The allocation-free dt entry point (zj_compute_dt_noalloc) against
zj_compute_dt, through the C ABI: identical device results on two calls in a
row, the second on a smaller mesh, in FP64 and FP32; arguments checked."""
import ctypes

import numpy as np
import pytest

from helpers import circular_dam_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _setup(n, dtype):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(n, n).set_bc("wall")
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=dtype)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          riemann="hll"))
    Q = circular_dam_state(mesh, seed=n)
    rng = np.random.default_rng(n)
    Q[2:] = rng.normal(0.0, 0.3, (2, mesh.n_cells))
    return s, torch.as_tensor(Q, dtype=dtype, device="cuda").contiguous()


def _dt(fn, s, Q, t_rem):
    out = torch.full((), -1.0, dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(s.runtime.handle, ctypes.c_void_p(Q.data_ptr()), float(t_rem),
            ctypes.c_void_p(out.data_ptr()), st)
    assert rc == 0
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_dt_matches_compute_dt(dtype):
    from zoomy_hip import _lib
    L = _lib.lib()
    # two calls in a row, the second on a smaller input; each with and
    # without a time cap
    for n, t_rem in ((96, -1.0), (40, -1.0), (40, 1e-6)):
        s, Q = _setup(n, dtype)
        ref = _dt(L.zj_compute_dt, s, Q, t_rem)
        got = _dt(L.zj_compute_dt_noalloc, s, Q, t_rem)
        torch.cuda.synchronize()
        print(f"n={n} t_rem={t_rem}: dt {float(ref):.17g} vs "
              f"{float(got):.17g}", flush=True)
        assert float(got) == float(ref)
        if t_rem > 0:
            assert float(got) == t_rem
        else:
            assert 0.0 < float(got) < 1.0
    # the scratch keeps the last call's partials: a smaller mesh after a
    # larger one reads only its own blocks (checked above); a wave-free state
    # falls back to dt_max, as in compute_timestep
    s2, Q2 = _setup(24, dtype)
    Q2[1:].zero_()
    got = _dt(L.zj_compute_dt_noalloc, s2, Q2, -1.0)
    assert float(got) == float(_dt(L.zj_compute_dt, s2, Q2, -1.0))


def test_noalloc_dt_checks_arguments():
    from zoomy_hip import _lib
    L = _lib.lib()
    null = ctypes.c_void_p(None)
    assert L.zj_compute_dt_noalloc(null, null, -1.0, null,
                                   null) == _lib.ZJ_EINVAL
