// Implicit (Crank-Nicolson) scalar diffusion on the unstructured FV mesh,
// behind the zj_diffusion_* C ABI (include/zoomy_hip.h).
//
// The operator is the reference's two-point Laplacian (DiffusionOperatorJAX,
// zoomy_jax/fvm/reconstruction_jax.py:1423-1520): per interior face f=(a,b)
// g_f = |f| / |x_b - x_a| / max(|n.e|, 0.1), (L u)_a = nu/|A_a| sum_f g_f
// (u_b - u_a); boundary faces carry no flux.  One CN step solves
//     (V - dt/2 nu G) u = (V + dt/2 nu G) u*       (G u)_a = sum_f g_f (u_b - u_a)
// which is the reference's (I - dt/2 L) u = (I + dt/2 L) u* multiplied by the
// cell volumes V: symmetric positive definite, so it runs as Jacobi-
// preconditioned CG (the reference solves it densely).
//
// Layout: the mesh's slot tables (mf, N) -- neighbour ids and the face
// coefficients g (0 on boundary slots) -- and per-cell volumes, all FP64.
// A CG iteration is two grid-stride passes, each followed by a single-block
// fold of its dot products: the direction update fused into M p (p = z +
// beta p_old recomputed for the neighbours, double-buffered), then the x, r,
// z update.  The scalars (alpha, beta, residuals) stay on the device, and
// the host reads a convergence flag once per chunk of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/zoomy_hip.h"

namespace {

constexpr int DB = 256;        // threads per block (4 wave64)
constexpr int DFIN = 1024;     // threads of the single-block folds
constexpr int DCHUNK = 8;      // CG iterations per host check
constexpr int DMAXB = 2048;    // grid cap of the CG passes (grid-stride)

#define DCHK(x)                                    \
    do {                                           \
        if ((x) != hipSuccess) return ZJ_EHIP;     \
    } while (0)

// device scalars of one solve
struct DifScal {
    double rz, bb, pq, rr, alpha, beta, tol2;
    int done, iters, maxit;
};

struct DifView {
    int N, mf;
    const int* nbr;     // (mf, N)
    const double* g;    // (mf, N), 0 on boundary slots
    const double* vol;  // (N)
};

__device__ __forceinline__ double gsum(const DifView& d, const double* __restrict__ u, int c,
                                       double& gs) {
    const double uc = u[c];
    double acc = 0.0;
    gs = 0.0;
    for (int k = 0; k < d.mf; ++k) {
        const int nb = d.nbr[(size_t)k * d.N + c];
        const double gk = d.g[(size_t)k * d.N + c];
        if (nb >= 0) {
            acc += gk * (u[nb] - uc);
            gs += gk;
        }
    }
    return acc;
}

// block sums of (a, b) into part[2 * block]
__device__ __forceinline__ void block_sum2(double a, double b, double* __restrict__ part) {
    __shared__ double la[DB / 64], lb[DB / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { la[w] = a; lb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < DB / 64; ++i) { sa += la[i]; sb += lb[i]; }
        part[2 * (size_t)blockIdx.x] = sa;
        part[2 * (size_t)blockIdx.x + 1] = sb;
    }
}

// Single-block fold of nb block pairs (fixed order: deterministic), then
// `what`: 0 start (rz, bb), 1 alpha = rz / p.q, 2 beta and convergence.
// (A fold in the last block of the producing kernel needs a device-scope
// fence per block, which costs an L2 write-back on every XCD: a separate
// launch is far cheaper.)
__global__ __launch_bounds__(DFIN) void k_dif_fin(const double* __restrict__ part, int nb,
                                                  DifScal* __restrict__ s, int what) {
    if (s->done && what != 0) return;
    __shared__ double la[DFIN / 64], lb[DFIN / 64];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += DFIN) {
        a += part[2 * (size_t)i];
        b += part[2 * (size_t)i + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { la[w] = a; lb[w] = b; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a = 0.0;
    b = 0.0;
    for (int i = 0; i < DFIN / 64; ++i) { a += la[i]; b += lb[i]; }
    if (what == 0) {
        s->rz = a;
        s->bb = b;
        s->rr = a;
        s->beta = 0.0;
        s->done = (a == 0.0);  // u* already solves the system (G u* = 0)
        s->iters = 0;
    } else if (what == 1) {
        s->pq = a;
        s->alpha = a != 0.0 ? s->rz / a : 0.0;
    } else {
        s->rr = b;
        s->beta = s->rz != 0.0 ? a / s->rz : 0.0;
        s->rz = a;
        s->iters += 1;
        if (b <= s->tol2 * s->bb || s->iters >= s->maxit) s->done = 1;
    }
}

// x = u*, r = dt nu G u*, z = r / diag, p_old = z (beta = 0 for the first
// direction); partials (r.z, b.b).  Grid-stride over the cells.
__global__ __launch_bounds__(DB) void k_dif_init(DifView d, double c_half, const double* __restrict__ u,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ z, double* __restrict__ po,
                                                 double* __restrict__ diag, double* __restrict__ part) {
    double rz = 0.0, bb = 0.0;
    for (int c = blockIdx.x * DB + threadIdx.x; c < d.N; c += gridDim.x * DB) {
        double gs;
        const double gu = gsum(d, u, c, gs);
        const double vc = d.vol[c];
        const double dg = vc + c_half * gs;
        const double rc = 2.0 * c_half * gu;
        const double bc = vc * u[c] + c_half * gu;
        const double zc = rc / dg;
        x[c] = u[c];
        r[c] = rc;
        z[c] = zc;
        po[c] = zc;
        diag[c] = dg;
        rz += rc * zc;
        bb += bc * bc;
    }
    block_sum2(rz, bb, part);
}

// the new direction p = z + beta p_old (own cell stored to pn, neighbours
// recomputed), q = M p = V p - c_half G p; partials p.q
__global__ __launch_bounds__(DB) void k_dif_mp(DifView d, double c_half, const double* __restrict__ z,
                                               const double* __restrict__ po, double* __restrict__ pn,
                                               double* __restrict__ q, double* __restrict__ part,
                                               const DifScal* __restrict__ s) {
    if (s->done) return;
    const double beta = s->beta;
    double pq = 0.0;
    for (int c = blockIdx.x * DB + threadIdx.x; c < d.N; c += gridDim.x * DB) {
        const double pc = z[c] + beta * po[c];
        double acc = 0.0;
        for (int k = 0; k < d.mf; ++k) {
            const int nb = d.nbr[(size_t)k * d.N + c];
            if (nb >= 0) acc += d.g[(size_t)k * d.N + c] * ((z[nb] + beta * po[nb]) - pc);
        }
        const double qc = d.vol[c] * pc - c_half * acc;
        pn[c] = pc;
        q[c] = qc;
        pq += pc * qc;
    }
    block_sum2(pq, 0.0, part);
}

// x += alpha p, r -= alpha q, z = r / diag; partials (r.z, r.r)
__global__ __launch_bounds__(DB) void k_dif_upd(int N, const double* __restrict__ p,
                                                const double* __restrict__ q,
                                                const double* __restrict__ diag, double* __restrict__ x,
                                                double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ part, const DifScal* __restrict__ s) {
    if (s->done) return;
    const double al = s->alpha;
    double rz = 0.0, rr = 0.0;
    for (int c = blockIdx.x * DB + threadIdx.x; c < N; c += gridDim.x * DB) {
        x[c] += al * p[c];
        const double rc = r[c] - al * q[c];
        const double zc = rc / diag[c];
        r[c] = rc;
        z[c] = zc;
        rz += rc * zc;
        rr += rc * rc;
    }
    block_sum2(rz, rr, part);
}

// (L u)_c = nu / |A_c| (G u)_c
__global__ __launch_bounds__(DB) void k_dif_explicit(DifView d, double nu, const double* __restrict__ u,
                                                     double* __restrict__ out) {
    for (int c = blockIdx.x * DB + threadIdx.x; c < d.N; c += gridDim.x * DB) {
        double gs;
        out[c] = nu * gsum(d, u, c, gs) / d.vol[c];
    }
}

}  // namespace

struct zj_diffusion {
    int N = 0, mf = 0, nblk = 0;
    int* nbr = nullptr;
    double *g = nullptr, *vol = nullptr;
    double *r = nullptr, *z = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *diag = nullptr;
    double* part = nullptr;
    DifScal* s = nullptr;
    DifScal* hs = nullptr;  // 2 pinned host slots of the convergence scalars
    hipEvent_t ev[2] = {nullptr, nullptr};
    int last_iters = 0;
    double last_relres = 0.0;
    DifView view() const { return DifView{N, mf, nbr, g, vol}; }
};

extern "C" {

int zj_diffusion_destroy(zj_diffusion* d) {
    if (!d) return ZJ_OK;
    void* ptrs[] = {d->nbr, d->g, d->vol, d->r,    d->z,    d->p,
                    d->p2,  d->q, d->diag, d->part, d->s};
    for (void* x : ptrs)
        if (x) (void)hipFree(x);
    if (d->hs) (void)hipHostFree(d->hs);
    for (hipEvent_t e : d->ev)
        if (e) (void)hipEventDestroy(e);
    delete d;
    return ZJ_OK;
}

int zj_diffusion_create(const zj_mesh_desc* m, zj_diffusion** out) {
    if (!m || !out || m->n_cells <= 0 || m->max_faces < 2 || m->max_faces > 4 ||
        m->dim < 1 || m->dim > 2 || (m->dim == 1 && m->max_faces != 2) || !m->cell_vertices ||
        !m->cell_nbr || !m->verts || !m->cell_center)
        return ZJ_EINVAL;
    const int N = m->n_cells, mf = m->max_faces;
    std::vector<double> g((size_t)mf * N, 0.0), vol(N);
    std::vector<int> nbr((size_t)mf * N);
    auto vx = [&](int c, int k) { return m->verts[2 * (size_t)m->cell_vertices[(size_t)k * N + c]]; };
    auto vy = [&](int c, int k) { return m->verts[2 * (size_t)m->cell_vertices[(size_t)k * N + c] + 1]; };
    for (size_t i = 0; i < (size_t)mf * N; ++i) {
        const int v = m->cell_vertices[i];
        if (v < 0 || v >= m->n_verts || m->cell_nbr[i] >= N) return ZJ_EINVAL;
    }
    for (int c = 0; c < N; ++c) {
        if (m->dim == 1) {
            vol[c] = vx(c, 1) - vx(c, 0);
        } else {
            double a = 0.0;
            for (int k = 0; k < mf; ++k) {
                const int k1 = (k + 1) % mf;
                a += vx(c, k) * vy(c, k1) - vx(c, k1) * vy(c, k);
            }
            vol[c] = 0.5 * a;
        }
        if (!(vol[c] > 0.0)) return ZJ_EINVAL;
        for (int k = 0; k < mf; ++k) {
            const int nb = m->cell_nbr[(size_t)k * N + c];
            nbr[(size_t)k * N + c] = nb;
            if (nb < 0) continue;
            double nx, ny, len;
            if (m->dim == 1) {
                nx = k == 0 ? -1.0 : 1.0;
                ny = 0.0;
                len = 1.0;
            } else {
                const int k1 = (k + 1) % mf;
                const double ex = vx(c, k1) - vx(c, k), ey = vy(c, k1) - vy(c, k);
                len = std::sqrt(ex * ex + ey * ey);
                nx = ey / len;
                ny = -ex / len;
            }
            const double dx = m->cell_center[nb] - m->cell_center[c];
            const double dy = m->dim == 1 ? 0.0 : m->cell_center[N + nb] - m->cell_center[N + c];
            const double dist = std::sqrt(dx * dx + dy * dy);
            if (dist < 1e-30) continue;
            double ne = (nx * dx + ny * dy) / dist;
            ne = std::fmax(std::fabs(ne), 0.1);
            g[(size_t)k * N + c] = len / dist / ne;
        }
    }
    zj_diffusion* d = new (std::nothrow) zj_diffusion();
    if (!d) return ZJ_ENOMEM;
    d->N = N;
    d->mf = mf;
    d->nblk = std::min((N + DB - 1) / DB, DMAXB);
    const size_t vb = sizeof(double) * N;
    bool ok = hipMalloc((void**)&d->nbr, sizeof(int) * nbr.size()) == hipSuccess &&
              hipMalloc((void**)&d->g, sizeof(double) * g.size()) == hipSuccess &&
              hipMalloc((void**)&d->vol, vb) == hipSuccess && hipMalloc((void**)&d->r, vb) == hipSuccess &&
              hipMalloc((void**)&d->z, vb) == hipSuccess && hipMalloc((void**)&d->p, vb) == hipSuccess &&
              hipMalloc((void**)&d->q, vb) == hipSuccess && hipMalloc((void**)&d->diag, vb) == hipSuccess &&
              hipMalloc((void**)&d->p2, vb) == hipSuccess &&
              hipMalloc((void**)&d->part, sizeof(double) * 2 * d->nblk) == hipSuccess &&
              hipMalloc((void**)&d->s, sizeof(DifScal)) == hipSuccess &&
              hipHostMalloc((void**)&d->hs, 2 * sizeof(DifScal), hipHostMallocDefault) == hipSuccess &&
              hipEventCreateWithFlags(&d->ev[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&d->ev[1], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        zj_diffusion_destroy(d);
        return ZJ_ENOMEM;
    }
    if (hipMemcpy(d->nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->g, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->vol, vol.data(), vb, hipMemcpyHostToDevice) != hipSuccess) {
        zj_diffusion_destroy(d);
        return ZJ_EHIP;
    }
    *out = d;
    return ZJ_OK;
}

int zj_diffusion_explicit(zj_diffusion* d, double nu, const double* u, double* Lu, void* stream) {
    if (!d || !u || !Lu || !(nu >= 0.0)) return ZJ_EINVAL;
    hipLaunchKernelGGL(k_dif_explicit, dim3(d->nblk), dim3(DB), 0, (hipStream_t)stream, d->view(), nu,
                       u, Lu);
    DCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_diffusion_cn(zj_diffusion* d, double nu, double dt, const double* u_star, double* u_out,
                    double tol, int32_t maxiter, void* stream) {
    if (!d || !u_star || !u_out || u_star == u_out || !(nu >= 0.0) || !(dt >= 0.0) ||
        !(tol > 0.0) || maxiter < 1)
        return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const double c_half = 0.5 * dt * nu;
    DifScal s0{};
    s0.tol2 = tol * tol;
    s0.maxit = maxiter;
    DCHK(hipMemcpyAsync(d->s, &s0, sizeof(DifScal), hipMemcpyHostToDevice, st));
    const DifView v = d->view();
    hipLaunchKernelGGL(k_dif_init, dim3(d->nblk), dim3(DB), 0, st, v, c_half, u_star, u_out, d->r, d->z,
                       d->p, d->diag, d->part);
    hipLaunchKernelGGL(k_dif_fin, dim3(1), dim3(DFIN), 0, st, (const double*)d->part, d->nblk, d->s, 0);
    // p_old / p_new alternate between the two direction buffers
    double* pb[2] = {d->p, d->p2};
    int cur = 0;
    auto chunk = [&]() {
        for (int i = 0; i < DCHUNK; ++i) {
            hipLaunchKernelGGL(k_dif_mp, dim3(d->nblk), dim3(DB), 0, st, v, c_half, (const double*)d->z,
                               (const double*)pb[cur], pb[cur ^ 1], d->q, d->part,
                               (const DifScal*)d->s);
            hipLaunchKernelGGL(k_dif_fin, dim3(1), dim3(DFIN), 0, st, (const double*)d->part, d->nblk,
                               d->s, 1);
            hipLaunchKernelGGL(k_dif_upd, dim3(d->nblk), dim3(DB), 0, st, d->N,
                               (const double*)pb[cur ^ 1], (const double*)d->q, (const double*)d->diag,
                               u_out, d->r, d->z, d->part, (const DifScal*)d->s);
            hipLaunchKernelGGL(k_dif_fin, dim3(1), dim3(DFIN), 0, st, (const double*)d->part, d->nblk,
                               d->s, 2);
            cur ^= 1;
        }
    };
    // the host reads chunk k's flag while chunk k + 1 runs (the passes after
    // convergence return at once), so the device never waits for the host
    chunk();
    DCHK(hipGetLastError());
    DCHK(hipMemcpyAsync(&d->hs[0], d->s, sizeof(DifScal), hipMemcpyDeviceToHost, st));
    DCHK(hipEventRecord(d->ev[0], st));
    DifScal h{};
    for (int k = 0;; ++k) {
        const int a = k & 1, b = a ^ 1;
        chunk();
        DCHK(hipGetLastError());
        DCHK(hipMemcpyAsync(&d->hs[b], d->s, sizeof(DifScal), hipMemcpyDeviceToHost, st));
        DCHK(hipEventRecord(d->ev[b], st));
        DCHK(hipEventSynchronize(d->ev[a]));
        h = d->hs[a];
        if (h.done) break;
    }
    DCHK(hipStreamSynchronize(st));
    d->last_iters = h.iters;
    d->last_relres = h.bb > 0.0 ? std::sqrt(h.rr / h.bb) : 0.0;
    if (h.bb > 0.0 && h.rr > s0.tol2 * h.bb) return ZJ_ESOLVE;
    return ZJ_OK;
}

int zj_diffusion_stats(zj_diffusion* d, int32_t* iters, double* relres) {
    if (!d || !iters || !relres) return ZJ_EINVAL;
    *iters = d->last_iters;
    *relres = d->last_relres;
    return ZJ_OK;
}

}  // extern "C"
