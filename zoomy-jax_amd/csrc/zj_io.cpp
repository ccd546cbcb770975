// HDF5 snapshot writer behind include/zoomy_io.h.
//
// Layout of zoomy_jax/misc/io.py:12-81: group "fields", one subgroup
// "iteration_<k>" per snapshot with the datasets "time" (scalar float64),
// "Q" and "Qaux" (row-major, the state's element type); append mode; an
// existing iteration group is replaced (overwrite) or refused.
//
// The HDF5 C library (1.10 ABI) is loaded at zj_io_open with dlopen -- from
// $ZJ_HDF5_LIB, else the image's /opt/conda/lib/libhdf5.so.103, else the
// loader's search path -- so the solver library has no link-time HDF5
// dependency; without one zj_io_open returns ZJ_EUNSUP.
//
// Asynchronous snapshots of device arrays: on the caller's stream, a
// device-to-device copy into a staging buffer (so later work on that stream
// may overwrite the state at once); then, ordered after it by an event, a
// copy into pinned host memory on a side stream of default priority; a
// writer thread waits for that copy and writes the group.  At most
// MAX_PENDING snapshots are in flight (a further call waits for the writer),
// which bounds the staging and pinned memory.  Both kinds of buffer are
// reused across snapshots and freed at close.  The HDF5 calls of one writer are serialised
// by its mutex (the synchronous path drains the queue first and holds it).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zoomy_io.h"

namespace {

// HDF5 1.10 C ABI: the handle and size types and the constants used here
// (H5Ipublic.h, H5public.h, H5Fpublic.h, H5Spublic.h).
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0, E_DEFAULT = 0;
constexpr unsigned ACC_RDWR = 0x0001u, ACC_TRUNC = 0x0002u;
constexpr int SCOPE_LOCAL = 0, S_SCALAR = 0;

struct H5Api {
    void* so = nullptr;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void*, void*);
    htri_t (*fis_hdf5)(const char*);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    herr_t (*fflush)(hid_t, int);
    htri_t (*lexists)(hid_t, const char*, hid_t);
    herr_t (*ldelete)(hid_t, const char*, hid_t);
    hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
    hid_t (*gopen)(hid_t, const char*, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dclose)(hid_t);
    const hid_t *f64le, *f32le, *native_double, *native_float;
};

std::mutex api_mu;
H5Api api;
bool api_ok = false;

// Loads the library and its entry points once; false when unavailable.
bool load_api() {
    std::lock_guard<std::mutex> lk(api_mu);
    if (api_ok) return true;
    const char* env = std::getenv("ZJ_HDF5_LIB");
    const char* names[] = {env, "/opt/conda/lib/libhdf5.so.103", "libhdf5.so.103", "libhdf5.so"};
    for (const char* n : names) {
        if (!n || !*n) continue;
        api.so = dlopen(n, RTLD_NOW | RTLD_LOCAL);
        if (api.so) break;
    }
    if (!api.so) return false;
    bool ok = true;
    auto sym = [&](auto& fp, const char* name) {
        void* v = dlsym(api.so, name);
        if (!v) ok = false;
        fp = reinterpret_cast<std::remove_reference_t<decltype(fp)>>(v);
    };
    sym(api.open, "H5open");
    sym(api.eset_auto, "H5Eset_auto2");
    sym(api.fis_hdf5, "H5Fis_hdf5");
    sym(api.fopen, "H5Fopen");
    sym(api.fcreate, "H5Fcreate");
    sym(api.fclose, "H5Fclose");
    sym(api.fflush, "H5Fflush");
    sym(api.lexists, "H5Lexists");
    sym(api.ldelete, "H5Ldelete");
    sym(api.gcreate, "H5Gcreate2");
    sym(api.gopen, "H5Gopen2");
    sym(api.gclose, "H5Gclose");
    sym(api.screate, "H5Screate");
    sym(api.screate_simple, "H5Screate_simple");
    sym(api.sclose, "H5Sclose");
    sym(api.dcreate, "H5Dcreate2");
    sym(api.dwrite, "H5Dwrite");
    sym(api.dclose, "H5Dclose");
    sym(api.f64le, "H5T_IEEE_F64LE_g");
    sym(api.f32le, "H5T_IEEE_F32LE_g");
    sym(api.native_double, "H5T_NATIVE_DOUBLE_g");
    sym(api.native_float, "H5T_NATIVE_FLOAT_g");
    // the type ids are set by the library's initialisation
    if (!ok || api.open() < 0) {
        dlclose(api.so);
        api.so = nullptr;
        return false;
    }
    // errors are returned as codes, not printed by the library
    api.eset_auto(E_DEFAULT, nullptr, nullptr);
    api_ok = true;
    return true;
}

}  // namespace

namespace {

constexpr int MAX_PENDING = 4;

struct Job {
    int64_t snap;
    double time;
    int dtype;
    int64_t nr, nc, na;
    void* hq;           // pinned copy of Q
    size_t cq;          // its capacity
    void* ha;           // pinned copy of Qaux (or nullptr)
    size_t ca;
    void* dq;           // device staging copies of Q and Qaux
    size_t cdq;
    void* da;
    size_t cda;
    hipEvent_t copied;  // recorded on the side stream after the copies
};

size_t esize(int dtype) { return dtype == ZJ_F64 ? 8 : 4; }

}  // namespace

struct zj_io {
    hid_t file = -1;
    hid_t fields = -1;
    int overwrite = 1;
    hipStream_t side = nullptr;
    std::thread writer;
    std::mutex mu;
    std::condition_variable cv_job, cv_idle;
    std::deque<Job> queue;
    int busy = 0;          // jobs popped and not finished
    bool stop = false;
    int err = ZJ_OK;       // first error since the last flush
    std::vector<std::pair<void*, size_t>> pool;   // free pinned buffers
    std::vector<std::pair<void*, size_t>> dpool;  // free device staging buffers
};

namespace {

// Writes fields/iteration_<snap> (misc/io.py:26-37).
int write_group(zj_io* io, int64_t snap, double time, int dtype, const void* q, int64_t nr,
                int64_t nc, const void* a, int64_t na) {
    const std::string name = "iteration_" + std::to_string(snap);
    const htri_t ex = api.lexists(io->fields, name.c_str(), P_DEFAULT);
    if (ex < 0) return ZJ_EIO;
    if (ex > 0) {
        if (!io->overwrite) return ZJ_EEXIST;
        if (api.ldelete(io->fields, name.c_str(), P_DEFAULT) < 0) return ZJ_EIO;
    }
    const hid_t g = api.gcreate(io->fields, name.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT);
    if (g < 0) return ZJ_EIO;
    int rc = ZJ_OK;
    auto put = [&](const char* dname, int rank, const hsize_t* dims, hid_t ftype, hid_t mtype,
                   const void* data) {
        if (rc) return;
        const hid_t sp = rank ? api.screate_simple(rank, dims, nullptr) : api.screate(S_SCALAR);
        if (sp < 0) {
            rc = ZJ_EIO;
            return;
        }
        const hid_t ds = api.dcreate(g, dname, ftype, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
        if (ds < 0 || api.dwrite(ds, mtype, S_ALL, S_ALL, P_DEFAULT, data) < 0) rc = ZJ_EIO;
        if (ds >= 0) api.dclose(ds);
        api.sclose(sp);
    };
    const hid_t ft = dtype == ZJ_F64 ? *api.f64le : *api.f32le;
    const hid_t mt = dtype == ZJ_F64 ? *api.native_double : *api.native_float;
    put("time", 0, nullptr, *api.f64le, *api.native_double, &time);
    const hsize_t dq[2] = {(hsize_t)nr, (hsize_t)nc};
    put("Q", 2, dq, ft, mt, q);
    if (a) {
        const hsize_t da[2] = {(hsize_t)na, (hsize_t)nc};
        put("Qaux", 2, da, ft, mt, a);
    }
    if (api.gclose(g) < 0 && !rc) rc = ZJ_EIO;
    if (!rc && api.fflush(io->file, SCOPE_LOCAL) < 0) rc = ZJ_EIO;
    return rc;
}

// A buffer of at least `bytes` from the pool (the smallest that fits), or a
// new one (pinned host memory, or device memory when dev).
void* take_buffer(zj_io* io, size_t bytes, size_t* cap, bool dev) {
    auto& pool = dev ? io->dpool : io->pool;
    {
        std::lock_guard<std::mutex> lk(io->mu);
        size_t best = (size_t)-1;
        for (size_t i = 0; i < pool.size(); ++i)
            if (pool[i].second >= bytes &&
                (best == (size_t)-1 || pool[i].second < pool[best].second))
                best = i;
        if (best != (size_t)-1) {
            void* p = pool[best].first;
            *cap = pool[best].second;
            pool.erase(pool.begin() + (long)best);
            return p;
        }
    }
    void* p = nullptr;
    const hipError_t e = dev ? hipMalloc(&p, bytes) : hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return nullptr;
    *cap = bytes;
    return p;
}

// Returns a job's buffers to the pools (io->mu held).
void recycle(zj_io* io, const Job& j) {
    if (j.hq) io->pool.emplace_back(j.hq, j.cq);
    if (j.ha) io->pool.emplace_back(j.ha, j.ca);
    if (j.dq) io->dpool.emplace_back(j.dq, j.cdq);
    if (j.da) io->dpool.emplace_back(j.da, j.cda);
}

void writer_loop(zj_io* io) {
    std::unique_lock<std::mutex> lk(io->mu);
    while (true) {
        io->cv_job.wait(lk, [&] { return io->stop || !io->queue.empty(); });
        if (io->queue.empty()) break;  // stop requested and nothing left
        Job j = io->queue.front();
        io->queue.pop_front();
        io->busy += 1;
        lk.unlock();
        int rc = hipEventSynchronize(j.copied) == hipSuccess ? ZJ_OK : ZJ_EHIP;
        if (!rc) rc = write_group(io, j.snap, j.time, j.dtype, j.hq, j.nr, j.nc, j.ha, j.na);
        (void)hipEventDestroy(j.copied);
        lk.lock();
        recycle(io, j);
        if (rc && !io->err) io->err = rc;
        io->busy -= 1;
        io->cv_idle.notify_all();
    }
}

// Waits until the queue is empty and the writer idle; returns and clears
// the first error.
int drain(zj_io* io) {
    std::unique_lock<std::mutex> lk(io->mu);
    io->cv_idle.wait(lk, [&] { return io->queue.empty() && io->busy == 0; });
    const int rc = io->err;
    io->err = ZJ_OK;
    return rc;
}

bool bad_shape(int32_t dtype, int64_t nr, int64_t nc, const void* a, int64_t na) {
    return (dtype != ZJ_F64 && dtype != ZJ_F32) || nr < 1 || nc < 1 || (a && na < 1);
}

}  // namespace

extern "C" {

int zj_io_open(const char* path, int32_t overwrite, zj_io** out) {
    if (!path || !out) return ZJ_EINVAL;
    *out = nullptr;
    if (!load_api()) return ZJ_EUNSUP;
    zj_io* io = new (std::nothrow) zj_io();
    if (!io) return ZJ_ENOMEM;
    io->overwrite = overwrite ? 1 : 0;
    const htri_t is = api.fis_hdf5(path);
    io->file = is > 0 ? api.fopen(path, ACC_RDWR, P_DEFAULT)
                      : api.fcreate(path, ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    if (io->file < 0) {
        delete io;
        return ZJ_EIO;
    }
    const htri_t has = api.lexists(io->file, "fields", P_DEFAULT);
    io->fields = has > 0 ? api.gopen(io->file, "fields", P_DEFAULT)
                         : api.gcreate(io->file, "fields", P_DEFAULT, P_DEFAULT, P_DEFAULT);
    if (has < 0 || io->fields < 0) {
        api.fclose(io->file);
        delete io;
        return ZJ_EIO;
    }
    io->writer = std::thread(writer_loop, io);
    *out = io;
    return ZJ_OK;
}

int zj_io_write(zj_io* io, int64_t i_snapshot, double time, const void* Q, int32_t dtype,
                int64_t n_rows, int64_t n_cells, const void* Qaux, int64_t n_aux_rows) {
    if (!io || !Q || i_snapshot < 0 || bad_shape(dtype, n_rows, n_cells, Qaux, n_aux_rows))
        return ZJ_EINVAL;
    if (int rc = drain(io)) return rc;
    return write_group(io, i_snapshot, time, dtype, Q, n_rows, n_cells, Qaux, n_aux_rows);
}

int zj_io_write_async(zj_io* io, int64_t i_snapshot, double time, const void* Q,
                      int32_t dtype, int64_t n_rows, int64_t n_cells, const void* Qaux,
                      int64_t n_aux_rows, void* stream) {
    if (!io || !Q || i_snapshot < 0 || bad_shape(dtype, n_rows, n_cells, Qaux, n_aux_rows))
        return ZJ_EINVAL;
    {
        // bound the snapshots in flight (and so the staging and pinned memory)
        std::unique_lock<std::mutex> lk(io->mu);
        io->cv_idle.wait(lk, [&] { return (int)io->queue.size() + io->busy < MAX_PENDING; });
    }
    if (!io->side && hipStreamCreateWithFlags(&io->side, hipStreamNonBlocking) != hipSuccess) {
        io->side = nullptr;
        return ZJ_EHIP;
    }
    Job j{};
    j.snap = i_snapshot;
    j.time = time;
    j.dtype = dtype;
    j.nr = n_rows;
    j.nc = n_cells;
    j.na = Qaux ? n_aux_rows : 0;
    const size_t bq = esize(dtype) * (size_t)n_rows * (size_t)n_cells;
    const size_t ba = esize(dtype) * (size_t)j.na * (size_t)n_cells;
    j.hq = take_buffer(io, bq, &j.cq, false);
    j.dq = take_buffer(io, bq, &j.cdq, true);
    if (Qaux) {
        j.ha = take_buffer(io, ba, &j.ca, false);
        j.da = take_buffer(io, ba, &j.cda, true);
    }
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t ready = nullptr;
    int rc = ZJ_OK;
    if (!j.hq || !j.dq || (Qaux && (!j.ha || !j.da))) rc = ZJ_ENOMEM;
    if (!rc && (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&j.copied, hipEventDisableTiming) != hipSuccess))
        rc = ZJ_EHIP;
    // staging copies on the caller's stream, host copies on the side stream
    if (!rc && (hipMemcpyAsync(j.dq, Q, bq, hipMemcpyDeviceToDevice, st) != hipSuccess ||
                (Qaux && hipMemcpyAsync(j.da, Qaux, ba, hipMemcpyDeviceToDevice, st) !=
                             hipSuccess) ||
                hipEventRecord(ready, st) != hipSuccess ||
                hipStreamWaitEvent(io->side, ready, 0) != hipSuccess ||
                hipMemcpyAsync(j.hq, j.dq, bq, hipMemcpyDeviceToHost, io->side) != hipSuccess ||
                (Qaux && hipMemcpyAsync(j.ha, j.da, ba, hipMemcpyDeviceToHost, io->side) !=
                             hipSuccess) ||
                hipEventRecord(j.copied, io->side) != hipSuccess))
        rc = ZJ_EHIP;
    if (ready) (void)hipEventDestroy(ready);
    if (rc) {
        // no copy in flight into the buffers when they go back to the pools
        (void)hipStreamSynchronize(st);
        (void)hipStreamSynchronize(io->side);
        std::lock_guard<std::mutex> lk(io->mu);
        recycle(io, j);
        if (j.copied) (void)hipEventDestroy(j.copied);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(io->mu);
        io->queue.push_back(j);
    }
    io->cv_job.notify_one();
    return ZJ_OK;
}

int zj_io_flush(zj_io* io) {
    if (!io) return ZJ_EINVAL;
    return drain(io);
}

int zj_io_close(zj_io* io) {
    if (!io) return ZJ_OK;
    int rc = drain(io);
    {
        std::lock_guard<std::mutex> lk(io->mu);
        io->stop = true;
    }
    io->cv_job.notify_all();
    if (io->writer.joinable()) io->writer.join();
    if (io->fields >= 0 && api.gclose(io->fields) < 0 && !rc) rc = ZJ_EIO;
    if (io->file >= 0 && api.fclose(io->file) < 0 && !rc) rc = ZJ_EIO;
    for (auto& b : io->pool) (void)hipHostFree(b.first);
    for (auto& b : io->dpool) (void)hipFree(b.first);
    if (io->side) (void)hipStreamDestroy(io->side);
    delete io;
    return rc;
}

}  // extern "C"
