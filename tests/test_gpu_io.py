"""Asynchronous HDF5 snapshots of device states (zj_io_write_async): each
snapshot holds the state at the moment of the call although the march goes
on overwriting Q in place right after it, with no host synchronisation in
between; read back with h5dump.  And run_simulation(write_output=True)
writes the reference's snapshot schedule (solver_jax.py:1264-1322)."""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from helpers import circular_dam_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

H5DUMP = shutil.which("h5dump") or (
    "/opt/conda/bin/h5dump" if pathlib.Path("/opt/conda/bin/h5dump").exists()
    else None)


def _dataset(path, name, dtype, shape, tmp):
    out = tmp / (name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(out), str(path)],
                   check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype).reshape(shape)


def _solver(mesh, dtype):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    s = HyperbolicSolver(time_end=np.inf, cfl=0.9, dtype=dtype)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          limiter="minmod", riemann="hll"))
    return s


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_async_snapshots_hold_the_state_of_the_call(tmp_path, dtype):
    from zoomy_hip.io import get_save_fields
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(96, 64).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=5)
    N = mesh.n_cells
    # reference run: the state after 0, 3, 6, ... steps, read back at once
    s = _solver(mesh, dtype)
    Q = torch.as_tensor(Q0, dtype=dtype, device="cuda").contiguous()
    want = []
    for k in range(6):
        want.append(Q.cpu().numpy().copy())
        s.march_device(Q, 3)
    # the same march with asynchronous snapshots and no synchronisation
    s2 = _solver(mesh, dtype)
    Q = torch.as_tensor(Q0, dtype=dtype, device="cuda").contiguous()
    Qaux = torch.zeros((7, N), dtype=dtype, device="cuda")
    f = tmp_path / "march.h5"
    save = get_save_fields(str(f), write_all=True)
    i = 0.0
    for k in range(6):
        i = save(0.1 * k, 0.0, i, Q, Qaux)
        s2.march_device(Q, 3)
    save.writer.close()
    assert i == 6.0
    et = "<f8" if dtype == torch.float64 else "<f4"
    for k in range(6):
        got = _dataset(f, f"/fields/iteration_{k}/Q", et, (4, N), tmp_path)
        assert np.array_equal(got, want[k]), k
        aux = _dataset(f, f"/fields/iteration_{k}/Qaux", et, (7, N), tmp_path)
        assert aux.shape == (7, N)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_run_simulation_writes_the_snapshot_schedule(tmp_path):
    from zoomy_hip.fvm import HyperbolicSolver, ShallowWater
    from zoomy_hip.mesh import Mesh
    mesh = Mesh.quad(32, 32).set_bc("wall")
    Q0 = circular_dam_state(mesh, seed=6)
    s = HyperbolicSolver(time_end=0.02, cfl=0.9, output_directory=str(tmp_path),
                         output_filename="dam", snapshots=3)
    s.setup_simulation(mesh, ShallowWater(dimension=2, order=2,
                                          limiter="minmod", riemann="hll"))
    Q = torch.as_tensor(Q0, device="cuda")
    Qaux = s.update_qaux(Q, None)
    Qf, _ = s.run_simulation(Q, Qaux, write_output=True)
    hdr = subprocess.run([H5DUMP, "-H", str(tmp_path / "dam.h5")], check=True,
                         capture_output=True, text=True).stdout
    groups = sorted(int(x) for x in
                    __import__("re").findall(r'GROUP "iteration_(\d+)"', hdr))
    # t = 0, then the first steps reaching 0.01 and 0.02
    assert groups == [0, 1, 2], groups
    t2 = float(_dataset(tmp_path / "dam.h5", "/fields/iteration_2/time", "<f8",
                        (), tmp_path))
    assert t2 == pytest.approx(0.02, rel=1e-12)
    last = _dataset(tmp_path / "dam.h5", "/fields/iteration_2/Q", "<f8",
                    (4, mesh.n_cells), tmp_path)
    assert np.array_equal(last, Qf.cpu().numpy())
    first = _dataset(tmp_path / "dam.h5", "/fields/iteration_0/Q", "<f8",
                     (4, mesh.n_cells), tmp_path)
    assert np.array_equal(first, Q0)
