"""HDF5 snapshots of the state: the drop-in for ``zoomy_jax.misc.io``
(``misc/io.py:12-81``).

``get_save_fields(filepath, write_all=False, overwrite=True)`` returns
``save_fields(time, next_write_at, i_snapshot, Q, Qaux)``, which writes
``fields/iteration_<i_snapshot>/{time, Q, Qaux}`` when ``write_all`` or
``time >= next_write_at`` and returns ``i_snapshot + 1``, else returns
``i_snapshot`` unchanged -- the reference's condition and counter
(``misc/io.py:39-46``).  An existing iteration group is replaced when
``overwrite``, otherwise a ``ValueError`` is raised (``misc/io.py:27-32``;
with device arrays the error surfaces at the next ``flush``/``close``).

The writer is ``libzoomy_hip``'s ``zj_io_*`` (C++, HDF5 loaded at run
time).  Device tensors are snapshotted asynchronously: a device-to-device
copy on the current stream, then the host copy and the file write off it
(``zj_io_write_async``); host arrays are written synchronously.  The file
holds every completed snapshot after each write (the writer flushes it);
``flush()`` waits for the queued ones.  Relative paths are taken from the
current working directory (the reference resolves them against zoomy_core's
main directory).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import _lib as L

_DT = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}


class SnapshotWriter:
    """One HDF5 file in the reference's snapshot layout, open for appending
    (``h5py.File(path, "a")``)."""

    def __init__(self, path: str, overwrite: bool = True):
        self.path = os.path.abspath(path)
        self.overwrite = bool(overwrite)
        h = ctypes.c_void_p()
        rc = L.lib().zj_io_open(self.path.encode(), int(self.overwrite),
                                ctypes.byref(h))
        if rc == L.ZJ_EUNSUP:
            raise RuntimeError("no HDF5 library found (set ZJ_HDF5_LIB to a "
                               "libhdf5.so of the 1.10 series)")
        self._check(rc, "zj_io_open")
        self._h = h

    def _check(self, rc, what):
        if rc == L.ZJ_EEXIST:
            raise ValueError(f"a snapshot group already exists in {self.path}"
                             " (overwrite=False)")
        L.check(rc, what)

    def write(self, i_snapshot: int, time: float, Q, Qaux=None) -> None:
        """Writes fields/iteration_<i_snapshot>.  ``Q`` (rows, cells) and
        ``Qaux`` are torch device tensors (asynchronous) or host arrays
        (synchronous), float64 or float32."""
        import torch
        if self._h is None:
            raise ValueError("the snapshot writer is closed")
        on_dev = isinstance(Q, torch.Tensor) and Q.is_cuda
        if on_dev:
            Qd = Q.contiguous()
            Ad = None if Qaux is None else Qaux.to(dtype=Qd.dtype).contiguous()
            if Qd.dim() != 2 or (Ad is not None and (Ad.dim() != 2 or
                                                    Ad.shape[1] != Qd.shape[1])):
                raise ValueError("Q and Qaux must be (rows, cells) arrays")
            if Qd.dtype not in (torch.float64, torch.float32):
                raise ValueError("snapshots hold float64 or float32 states")
            dt = 0 if Qd.dtype == torch.float64 else 1
            stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            rc = L.lib().zj_io_write_async(
                self._h, int(i_snapshot), float(time),
                ctypes.c_void_p(Qd.data_ptr()), dt, Qd.shape[0], Qd.shape[1],
                None if Ad is None else ctypes.c_void_p(Ad.data_ptr()),
                0 if Ad is None else Ad.shape[0], stream)
            self._check(rc, "zj_io_write_async")
            return
        q = np.ascontiguousarray(Q.cpu().numpy() if hasattr(Q, "cpu") else Q)
        if q.dtype not in _DT:
            q = q.astype(np.float64)
        a = None
        if Qaux is not None:
            a = np.ascontiguousarray(
                (Qaux.cpu().numpy() if hasattr(Qaux, "cpu") else Qaux)
                .astype(q.dtype))
        if q.ndim != 2 or (a is not None and (a.ndim != 2 or
                                              a.shape[1] != q.shape[1])):
            raise ValueError("Q and Qaux must be (rows, cells) arrays")
        rc = L.lib().zj_io_write(
            self._h, int(i_snapshot), float(time), q.ctypes.data_as(ctypes.c_void_p),
            _DT[q.dtype], q.shape[0], q.shape[1],
            None if a is None else a.ctypes.data_as(ctypes.c_void_p),
            0 if a is None else a.shape[0])
        self._check(rc, "zj_io_write")

    def flush(self) -> None:
        """Waits until every queued snapshot is in the file."""
        if self._h is not None:
            self._check(L.lib().zj_io_flush(self._h), "zj_io_flush")

    def close(self) -> None:
        if self._h is not None:
            h, self._h = self._h, None
            self._check(L.lib().zj_io_close(h), "zj_io_close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                L.lib().zj_io_close(self._h)
                self._h = None
        except Exception:
            pass


def get_save_fields(filepath: str, write_all: bool = False,
                    overwrite: bool = True):
    """The reference's ``get_save_fields`` (``misc/io.py:12-81``).  The
    returned function carries its writer as ``save_fields.writer`` (for
    ``flush``/``close``)."""
    writer = SnapshotWriter(filepath, overwrite=overwrite)

    def save_fields(time, next_write_at, i_snapshot, Q, Qaux):
        if not write_all and float(time) < float(next_write_at):
            return i_snapshot
        writer.write(int(i_snapshot), float(time), Q, Qaux)
        return i_snapshot + 1.0

    save_fields.writer = writer
    return save_fields
