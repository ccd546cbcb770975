/* zoomy_io -- C ABI of the HDF5 snapshot writer (libzoomy_io.so).
 *
 * Replaces the h5py writer of zoomy_jax/misc/io.py:12-81
 * (get_save_fields / _save_hdf5): one HDF5 file holds a group "fields" with
 * one subgroup "iteration_<k>" per snapshot k, holding the datasets
 * "time" (scalar float64), "Q" (n_rows, n_cells) and, when given, "Qaux"
 * (n_aux_rows, n_cells), in the element type of the state (float64 or
 * float32).  The file is opened in append mode, as h5py.File(path, "a"):
 * an existing file keeps its snapshots; an existing iteration group is
 * replaced when overwrite != 0 and is an error (ZJ_EEXIST) otherwise.
 *
 * zj_io_write_async takes device arrays: it queues a device-to-device copy
 * into a staging buffer on `stream` (so work queued after it may overwrite
 * Q at once), then the copy into pinned host memory on a side stream of its
 * own, and returns; a writer thread waits for that copy and writes the
 * group.  So the solver's stream never waits for PCIe or the file system.
 * At most 4 snapshots are in flight; a further call waits for the writer.
 * A write error is kept and returned by the next zj_io_flush / zj_io_close.
 * The HDF5 C library (1.10 ABI) is loaded at zj_io_open: from $ZJ_HDF5_LIB,
 * else /opt/conda/lib/libhdf5.so.103, else the loader's search path; with
 * none, zj_io_open returns ZJ_EUNSUP.
 *
 * Error codes are those of zoomy_hip.h, plus ZJ_EIO (the HDF5 library
 * reported an error) and ZJ_EEXIST.
 */
#ifndef ZOOMY_IO_H
#define ZOOMY_IO_H

#include <stdint.h>

#include "zoomy_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ZJ_EIO 6        /* the HDF5 library reported an error          */
#define ZJ_EEXIST 7     /* the iteration group exists, overwrite == 0  */

typedef struct zj_io zj_io;

/* Opens (or creates) the file and its "fields" group (misc/io.py:21-25). */
int zj_io_open(const char* path, int32_t overwrite, zj_io** out);

/* Writes fields/iteration_<i_snapshot> from host arrays (row-major,
 * dtype ZJ_F64 or ZJ_F32; Qaux may be NULL), after every queued
 * asynchronous snapshot (misc/io.py:26-37). */
int zj_io_write(zj_io* io, int64_t i_snapshot, double time, const void* Q, int32_t dtype,
                int64_t n_rows, int64_t n_cells, const void* Qaux, int64_t n_aux_rows);

/* The same from device arrays, ordered after the work queued on `stream`;
 * returns before the data is copied or written (see above). */
int zj_io_write_async(zj_io* io, int64_t i_snapshot, double time, const void* Q,
                      int32_t dtype, int64_t n_rows, int64_t n_cells, const void* Qaux,
                      int64_t n_aux_rows, void* stream);

/* Waits until every queued snapshot is on disk; returns the first error of
 * a write since the last flush (or ZJ_OK). */
int zj_io_flush(zj_io* io);

/* Flushes, closes the file and frees the writer. */
int zj_io_close(zj_io* io);

#ifdef __cplusplus
}
#endif

#endif /* ZOOMY_IO_H */
