"""HDF5 snapshots in the layout of ``zoomy_jax/misc/io.py:12-81``
(``fields/iteration_<k>/{time, Q, Qaux}``), written through the C ABI from
host arrays (no device needed) and read back with ``h5dump``, the reader of
the HDF5 distribution itself -- not with this package's code."""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or (
    "/opt/conda/bin/h5dump" if pathlib.Path("/opt/conda/bin/h5dump").exists()
    else None)

pytestmark = pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")


def _dataset(path, name, dtype, shape, tmp):
    """Raw little-endian values of one dataset, via h5dump's binary output."""
    out = tmp / (name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(out), str(path)],
                   check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype).reshape(shape)


def _header(path):
    return subprocess.run([H5DUMP, "-H", str(path)], check=True,
                          capture_output=True, text=True).stdout


def test_snapshots_read_back_by_h5dump(tmp_path):
    from zoomy_hip.io import SnapshotWriter
    rng = np.random.default_rng(4)
    f = tmp_path / "run.h5"
    snaps = []
    with SnapshotWriter(str(f)) as w:
        for k in range(3):
            Q = rng.normal(size=(4, 37))
            A = rng.normal(size=(7, 37))
            w.write(k, 0.125 * k, Q, A)
            snaps.append((Q, A))
    hdr = _header(f)
    for k in range(3):
        assert f'GROUP "iteration_{k}"' in hdr
    assert "H5T_IEEE_F64LE" in hdr and "( 4, 37 )" in hdr and "( 7, 37 )" in hdr
    assert 'DATASET "time"' in hdr and "SCALAR" in hdr
    for k, (Q, A) in enumerate(snaps):
        base = f"/fields/iteration_{k}"
        assert np.array_equal(_dataset(f, base + "/Q", "<f8", Q.shape, tmp_path), Q)
        assert np.array_equal(_dataset(f, base + "/Qaux", "<f8", A.shape, tmp_path), A)
        t = _dataset(f, base + "/time", "<f8", (), tmp_path)
        assert float(t) == 0.125 * k


def test_float32_state_and_no_aux(tmp_path):
    from zoomy_hip.io import SnapshotWriter
    f = tmp_path / "f32.h5"
    Q = np.arange(3 * 11, dtype=np.float32).reshape(3, 11) / 7
    with SnapshotWriter(str(f)) as w:
        w.write(0, 0.5, Q)
    hdr = _header(f)
    assert "H5T_IEEE_F32LE" in hdr and 'DATASET "Qaux"' not in hdr
    assert np.array_equal(_dataset(f, "/fields/iteration_0/Q", "<f4", Q.shape,
                                   tmp_path), Q)


def test_append_mode_and_overwrite(tmp_path):
    """h5py.File(path, "a"): a reopened file keeps its snapshots; an existing
    iteration group is replaced with overwrite and refused without it
    (misc/io.py:27-32)."""
    from zoomy_hip.io import SnapshotWriter
    f = tmp_path / "app.h5"
    Q0 = np.full((4, 5), 1.0)
    Q1 = np.full((4, 5), 2.0)
    with SnapshotWriter(str(f)) as w:
        w.write(0, 0.0, Q0)
    with SnapshotWriter(str(f)) as w:
        w.write(1, 0.1, Q1)
    assert np.array_equal(_dataset(f, "/fields/iteration_0/Q", "<f8", (4, 5),
                                   tmp_path), Q0)
    assert np.array_equal(_dataset(f, "/fields/iteration_1/Q", "<f8", (4, 5),
                                   tmp_path), Q1)
    with SnapshotWriter(str(f), overwrite=False) as w:
        with pytest.raises(ValueError):
            w.write(0, 0.0, Q1)
    with SnapshotWriter(str(f), overwrite=True) as w:
        w.write(0, 0.3, Q1)
    assert np.array_equal(_dataset(f, "/fields/iteration_0/Q", "<f8", (4, 5),
                                   tmp_path), Q1)
    assert float(_dataset(f, "/fields/iteration_0/time", "<f8", (), tmp_path)) == 0.3


def test_save_fields_condition_and_counter(tmp_path):
    """save_fields writes when write_all or time >= next_write_at and then
    returns i_snapshot + 1, else returns i_snapshot (misc/io.py:39-46)."""
    from zoomy_hip.io import get_save_fields
    f = tmp_path / "sf.h5"
    save = get_save_fields(str(f), write_all=False)
    Q = np.ones((4, 3))
    i = save(0.0, 0.0, 0.0, Q, None)
    assert i == 1.0
    i2 = save(0.05, 0.1, i, Q, None)       # before the next write time
    assert i2 == i
    i3 = save(0.1, 0.1, i2, 2 * Q, None)
    assert i3 == 2.0
    save.writer.close()
    hdr = _header(f)
    assert 'GROUP "iteration_0"' in hdr and 'GROUP "iteration_1"' in hdr
    assert 'GROUP "iteration_2"' not in hdr
    every = get_save_fields(str(tmp_path / "all.h5"), write_all=True)
    assert every(0.01, 5.0, 7.0, Q, None) == 8.0
    every.writer.close()
    assert 'GROUP "iteration_7"' in _header(tmp_path / "all.h5")


def test_bad_arguments():
    import ctypes
    from zoomy_hip import _lib
    L = _lib.lib()
    null = ctypes.c_void_p(None)
    assert L.zj_io_open(None, 1, ctypes.byref(ctypes.c_void_p())) == _lib.ZJ_EINVAL
    assert L.zj_io_write(null, 0, 0.0, null, 0, 4, 5, null, 0) == _lib.ZJ_EINVAL
    assert L.zj_io_flush(null) == _lib.ZJ_EINVAL
    assert L.zj_io_close(null) == _lib.ZJ_OK
