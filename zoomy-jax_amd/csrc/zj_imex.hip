// Global implicit source stage of the IMEX solver: Newton-GMRES on
//   R(Q) = Q - Qexp - dt S(Q, aux(Q)) = 0
// for sources that read the derivative aux rows (the reference's
// _build_implicit_global, zoomy_jax/fvm/solver_imex_jax.py:344-396, chosen by
// source_mode "auto" when the model has derivative specs, :182-188).  The aux
// rows are the LSQ gradients of the post-step refresh (zj_update_aux), so the
// stage couples every cell to its stencil.
//
// Source (rows q_d, d < dim): the Manning friction of the model and the
// depth-gradient term -alpha g h dh/dx_d, which reads aux row "dh/dx_d".
// The Jacobian-vector product is the exact linearisation (what the
// reference's "ad" backend differentiates): J v = v - dt (dS/dQ v +
// dS/daux . aux(v)), where aux(v) -- the LSQ gradients of v -- is linear in
// v, and the 1/h of the Manning term is differentiated through its
// desingularised formula.  The stage changes the momentum rows only (S has
// no b or h rows, and Q starts at Qexp), so the Newton unknowns and every
// Krylov vector live in the dim momentum rows: v_h = 0, hence the aux(v)
// rows the source reads (dh/dx_d) vanish and aux(Q) is the same for every
// Newton step -- it is refreshed once per solve.  GMRES is restarted
// GMRES(m) with classical Gram-Schmidt and one reorthogonalisation, Givens
// rotations on the host.  Every reduction stores one partial per block and
// a fold kernel sums them in a fixed order, so the stage is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <new>
#include <vector>

#include "../../include/zoomy_hip.h"
#include "zj_chorin.hpp"

namespace {

constexpr int IB = 256;        // threads per block
constexpr int IGRID = 512;     // blocks of the grid-stride vector kernels
constexpr int MAXK = 32;       // largest Krylov space (restart length)

#define IHIPCHK(x)                                                    \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "zoomy_hip: %s failed: %s\n", #x,         \
                    hipGetErrorString(e_));                           \
            return ZJ_EHIP;                                           \
        }                                                             \
    } while (0)

struct SrcPrm {
    double g, n2, eps, alpha, dt;
    int N, dim;
};

__device__ __forceinline__ double hinv_d(double h, double eps) {
    const double m = fmax(h, eps);
    return 1.4142135623730951 * h / sqrt(h * h * h * h + m * m * m * m);
}

// aux row of dh/dx_d: rows per direction d are [dq_d/dx_d, dh/dx_d]
__device__ __forceinline__ int row_dh(int d) { return 2 * d + 1; }

// S(q, aux) rows q_d at cell c; kf = g n^2 hinv(h)^(7/3), the Manning
// coefficient of the cell (h does not change during the stage)
__device__ __forceinline__ void source_at(const SrcPrm& p, const double* q, double kf,
                                          const double* hx, double* S) {
    const double qq = sqrt(q[2] * q[2] + q[3] * q[3]);
    const double m = kf * qq;
    S[0] = -p.alpha * p.g * q[1] * hx[0] - m * q[2];
    S[1] = p.dim == 2 ? -p.alpha * p.g * q[1] * hx[1] - m * q[3] : -m * q[3];
}

// kf[c] = g n^2 hinv(h_c)^(7/3), once per stage
__global__ __launch_bounds__(IB) void k_friction(SrcPrm p, const double* __restrict__ Q,
                                                 double* __restrict__ kf) {
    for (int c = blockIdx.x * IB + threadIdx.x; c < p.N; c += gridDim.x * IB)
        kf[c] = p.g * p.n2 * pow(hinv_d(Q[(size_t)p.N + c], p.eps), 7.0 / 3.0);
}

__device__ __forceinline__ void block_sum(double v, double* lds, double* out) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < IB / 64; ++i) s += lds[i];
        *out = s;
    }
    __syncthreads();
}

// R = Q - Qexp - dt S(Q, aux) in the momentum rows (R[d N + c], d < dim;
// the b and h rows of Q equal Qexp's); part[block] = sum of R^2 over the
// block
__global__ __launch_bounds__(IB) void k_src_residual(SrcPrm p, const double* __restrict__ Q,
                                                     const double* __restrict__ Qe,
                                                     const double* __restrict__ aux,
                                                     const double* __restrict__ kf,
                                                     double* __restrict__ R,
                                                     double* __restrict__ part) {
    __shared__ double lds[IB / 64];
    double acc = 0.0;
    const int N = p.N;
    for (int c = blockIdx.x * IB + threadIdx.x; c < N; c += gridDim.x * IB) {
        double q[4] = {0.0, 0.0, 0.0, 0.0}, hx[2] = {0.0, 0.0};
        for (int v = 1; v < 4; ++v) q[v] = Q[(size_t)v * N + c];
        for (int d = 0; d < p.dim; ++d) hx[d] = aux[(size_t)row_dh(d) * N + c];
        double S[2];
        source_at(p, q, kf[c], hx, S);
        for (int d = 0; d < p.dim; ++d) {
            const double r = q[2 + d] - Qe[(size_t)(2 + d) * N + c] - p.dt * S[d];
            R[(size_t)d * N + c] = r;
            acc += r * r;
        }
    }
    block_sum(acc, lds, part + blockIdx.x);
}

// out = J v = v - dt dS/dq v for momentum vectors v (v[d N + c], d < dim):
// with v_h = 0 the terms of dh and of aux(v) (whose dh/dx_d rows are the
// gradients of v_h) vanish
__global__ __launch_bounds__(IB) void k_src_jvp(SrcPrm p, const double* __restrict__ Q,
                                                const double* __restrict__ kf,
                                                const double* __restrict__ V,
                                                double* __restrict__ out) {
    const int N = p.N;
    for (int c = blockIdx.x * IB + threadIdx.x; c < N; c += gridDim.x * IB) {
        double q[4] = {0.0, 0.0, 0.0, 0.0}, v[2] = {0.0, 0.0};
        for (int k = 2; k < 4; ++k) q[k] = Q[(size_t)k * N + c];
        for (int d = 0; d < p.dim; ++d) v[d] = V[(size_t)d * N + c];
        const double qq = sqrt(q[2] * q[2] + q[3] * q[3]);
        const double k = kf[c];
        // d(|q| k) = k (q.dq / |q|) at fixed h
        const double dqq = qq > 0.0 ? (q[2] * v[0] + (p.dim == 2 ? q[3] * v[1] : 0.0)) / qq : 0.0;
        const double dm = k * dqq;
        const double m = k * qq;
        for (int d = 0; d < p.dim; ++d) {
            const double dS = -(dm * q[2 + d] + m * v[d]);
            out[(size_t)d * N + c] = v[d] - p.dt * dS;
        }
    }
}

struct Coefs {
    double a[MAXK + 1];
};

// part[i * gridDim + block] = <w, v_i> over the block's entries, i <= k
__global__ __launch_bounds__(IB) void k_mdot(int n, int k, const double* __restrict__ w,
                                             const double* __restrict__ Vb,
                                             double* __restrict__ part) {
    __shared__ double lds[IB / 64];
    for (int i = 0; i <= k; ++i) {
        const double* vi = Vb + (size_t)i * n;
        double acc = 0.0;
        for (int j = blockIdx.x * IB + threadIdx.x; j < n; j += gridDim.x * IB)
            acc += w[j] * vi[j];
        block_sum(acc, lds, part + (size_t)i * gridDim.x + blockIdx.x);
    }
}

// out[i] = sum of the nb partials of dot i, in a fixed order
__global__ __launch_bounds__(IB) void k_fold(int nb, const double* __restrict__ part,
                                             double* __restrict__ out) {
    __shared__ double lds[IB / 64];
    const double* p = part + (size_t)blockIdx.x * nb;
    double acc = 0.0;
    for (int j = threadIdx.x; j < nb; j += IB) acc += p[j];
    block_sum(acc, lds, out + blockIdx.x);
}

// w += sum_i a_i v_i (i < m) (and w *= s first when scale)
__global__ __launch_bounds__(IB) void k_maxpy(int n, int m, Coefs c, double s,
                                              const double* __restrict__ Vb,
                                              double* __restrict__ w) {
    for (int j = blockIdx.x * IB + threadIdx.x; j < n; j += gridDim.x * IB) {
        double x = w[j] * s;
        for (int i = 0; i < m; ++i) x += c.a[i] * Vb[(size_t)i * n + j];
        w[j] = x;
    }
}

// y = a x
__global__ __launch_bounds__(IB) void k_scale(int n, double a, const double* __restrict__ x,
                                              double* __restrict__ y) {
    for (int j = blockIdx.x * IB + threadIdx.x; j < n; j += gridDim.x * IB) y[j] = a * x[j];
}

}  // namespace

struct zj_imex_global {
    zj_solver* s = nullptr;
    int N = 0, n = 0, dim = 2, m = 20;
    double alpha = 0.0;
    double *aux = nullptr, *kf = nullptr, *R = nullptr, *W = nullptr, *Vb = nullptr,
           *dX = nullptr, *part = nullptr, *dots = nullptr, *host = nullptr;
    int newton_its = 0, gmres_its = 0, fallbacks = 0;
    double rnorm = 0.0;
};

namespace {

// dot products <w, v_i>, i <= k, on the host (synchronising)
int mdot(zj_imex_global* g, const double* w, const double* Vb, int k, double* out,
         hipStream_t st) {
    k_mdot<<<IGRID, IB, 0, st>>>(g->n, k, w, Vb, g->part);
    k_fold<<<k + 1, IB, 0, st>>>(IGRID, g->part, g->dots);
    IHIPCHK(hipMemcpyAsync(g->host, g->dots, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, st));
    IHIPCHK(hipStreamSynchronize(st));
    std::copy(g->host, g->host + k + 1, out);
    return ZJ_OK;
}

int norm2(zj_imex_global* g, const double* w, double* out, hipStream_t st) {
    double d = 0.0;
    if (int rc = mdot(g, w, w, 0, &d, st)) return rc;
    *out = std::sqrt(std::max(d, 0.0));
    return ZJ_OK;
}

// R(Q) into g->R and its norm (g->aux holds aux(Q), whose dh/dx_d rows do
// not change during the stage)
int residual(zj_imex_global* g, const SrcPrm& p, const double* Q, const double* Qe,
             double* rnorm, hipStream_t st) {
    k_src_residual<<<IGRID, IB, 0, st>>>(p, Q, Qe, g->aux, g->kf, g->R, g->part);
    k_fold<<<1, IB, 0, st>>>(IGRID, g->part, g->dots);
    IHIPCHK(hipMemcpyAsync(g->host, g->dots, sizeof(double), hipMemcpyDeviceToHost, st));
    IHIPCHK(hipStreamSynchronize(st));
    *rnorm = std::sqrt(std::max(g->host[0], 0.0));
    return ZJ_OK;
}

// J v at the current Q
int jvp(zj_imex_global* g, const SrcPrm& p, const double* Q, const double* v, double* out,
        hipStream_t st) {
    k_src_jvp<<<IGRID, IB, 0, st>>>(p, Q, g->kf, v, out);
    IHIPCHK(hipGetLastError());
    return ZJ_OK;
}

// GMRES(m) for J x = b (b = -R in g->R's negation), x accumulated into dx
// (zero start), relative tolerance tol, at most maxres restart cycles.
// Returns 0 when the tolerance was met.
int gmres(zj_imex_global* g, const SrcPrm& p, const double* Q, double* dx, double tol,
          int maxres, int* its, int* converged, hipStream_t st) {
    const int n = g->n, m = g->m;
    double* V = g->Vb;                     // (m + 1) basis vectors
    double* w = g->W;
    // b = -R
    double bnorm = 0.0;
    if (int rc = norm2(g, g->R, &bnorm, st)) return rc;
    IHIPCHK(hipMemsetAsync(dx, 0, sizeof(double) * n, st));
    *converged = bnorm == 0.0;
    *its = 0;
    if (*converged) return ZJ_OK;
    const double target = tol * bnorm;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), e(m + 1), h(m + 1), h2(m + 1);
    for (int cyc = 0; cyc < maxres; ++cyc) {
        // r = b - J dx
        double beta;
        if (cyc == 0) {
            k_scale<<<IGRID, IB, 0, st>>>(n, -1.0, g->R, V);
            beta = bnorm;
        } else {
            // V0 = -R - J dx
            if (int rc = jvp(g, p, Q, dx, w, st)) return rc;
            k_scale<<<IGRID, IB, 0, st>>>(n, -1.0, g->R, V);
            Coefs c1{};
            c1.a[0] = -1.0;
            k_maxpy<<<IGRID, IB, 0, st>>>(n, 1, c1, 1.0, w, V);
            if (int rc = norm2(g, V, &beta, st)) return rc;
            if (beta <= target) {
                *converged = 1;
                return ZJ_OK;
            }
        }
        k_scale<<<IGRID, IB, 0, st>>>(n, 1.0 / beta, V, V);
        std::fill(e.begin(), e.end(), 0.0);
        e[0] = beta;
        int k = 0;
        for (; k < m; ++k) {
            double* vk1 = V + (size_t)(k + 1) * n;
            if (int rc = jvp(g, p, Q, V + (size_t)k * n, vk1, st)) return rc;
            // classical Gram-Schmidt, twice
            if (int rc = mdot(g, vk1, V, k, h.data(), st)) return rc;
            Coefs c{};
            for (int i = 0; i <= k; ++i) c.a[i] = -h[i];
            k_maxpy<<<IGRID, IB, 0, st>>>(n, k + 1, c, 1.0, V, vk1);
            if (int rc = mdot(g, vk1, V, k, h2.data(), st)) return rc;
            for (int i = 0; i <= k; ++i) {
                c.a[i] = -h2[i];
                h[i] += h2[i];
            }
            k_maxpy<<<IGRID, IB, 0, st>>>(n, k + 1, c, 1.0, V, vk1);
            double hn = 0.0;
            if (int rc = norm2(g, vk1, &hn, st)) return rc;
            h[k + 1] = hn;
            if (hn > 0.0) k_scale<<<IGRID, IB, 0, st>>>(n, 1.0 / hn, vk1, vk1);
            // apply the previous rotations, then a new one
            for (int i = 0; i < k; ++i) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double r = std::hypot(h[k], h[k + 1]);
            cs[k] = r > 0.0 ? h[k] / r : 1.0;
            sn[k] = r > 0.0 ? h[k + 1] / r : 0.0;
            h[k] = r;
            h[k + 1] = 0.0;
            e[k + 1] = -sn[k] * e[k];
            e[k] = cs[k] * e[k];
            for (int i = 0; i <= k; ++i) H[(size_t)i * m + k] = h[i];
            *its += 1;
            if (std::fabs(e[k + 1]) <= target || hn == 0.0) {
                ++k;
                break;
            }
        }
        // y = H^-1 e (upper triangular k x k), dx += V y
        std::vector<double> y(k);
        for (int i = k - 1; i >= 0; --i) {
            double s = e[i];
            for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
            y[i] = s / H[(size_t)i * m + i];
        }
        Coefs c{};
        for (int i = 0; i < k; ++i) c.a[i] = y[i];
        k_maxpy<<<IGRID, IB, 0, st>>>(n, k, c, 1.0, V, dx);
        IHIPCHK(hipGetLastError());
        if (std::fabs(e[k]) <= target) {
            *converged = 1;
            return ZJ_OK;
        }
    }
    return ZJ_OK;
}

}  // namespace

extern "C" {

int zj_imex_global_create(zj_solver* s, double alpha, int32_t restart, zj_imex_global** out) {
    if (!s || !out || !(alpha >= 0.0) || restart < 1 || restart > MAXK) return ZJ_EINVAL;
    if (zj::solver_dtype(s) != ZJ_F64) return ZJ_EUNSUP;
    zj_imex_global* g = new (std::nothrow) zj_imex_global();
    if (!g) return ZJ_ENOMEM;
    g->s = s;
    g->N = zj::solver_cells(s);
    g->dim = zj::solver_dim(s);
    g->n = g->dim * g->N;  // the momentum rows
    g->m = restart;
    g->alpha = alpha;
    const size_t naux = (size_t)(3 * g->dim + 1) * g->N;
    if (hipMalloc((void**)&g->aux, sizeof(double) * naux) != hipSuccess ||
        hipMalloc((void**)&g->kf, sizeof(double) * g->N) != hipSuccess ||
        hipMalloc((void**)&g->R, sizeof(double) * g->n) != hipSuccess ||
        hipMalloc((void**)&g->W, sizeof(double) * g->n) != hipSuccess ||
        hipMalloc((void**)&g->dX, sizeof(double) * g->n) != hipSuccess ||
        hipMalloc((void**)&g->Vb, sizeof(double) * g->n * (size_t)(g->m + 1)) != hipSuccess ||
        hipMalloc((void**)&g->part, sizeof(double) * IGRID * (MAXK + 1)) != hipSuccess ||
        hipMalloc((void**)&g->dots, sizeof(double) * (MAXK + 1)) != hipSuccess ||
        hipHostMalloc((void**)&g->host, sizeof(double) * (MAXK + 1)) != hipSuccess) {
        zj_imex_global_destroy(g);
        return ZJ_ENOMEM;
    }
    *out = g;
    return ZJ_OK;
}

int zj_imex_global_destroy(zj_imex_global* g) {
    if (!g) return ZJ_OK;
    for (double* p : {g->aux, g->kf, g->R, g->W, g->Vb, g->dX, g->part, g->dots})
        if (p) (void)hipFree(p);
    if (g->host) (void)hipHostFree(g->host);
    delete g;
    return ZJ_OK;
}

int zj_imex_global_solve(zj_imex_global* g, double dt, const double* Qexp, double* Qout,
                         int32_t maxiter, double tol, double gmres_tol, int32_t gmres_maxiter,
                         void* stream) {
    if (!g || !Qexp || !Qout || Qexp == Qout || maxiter < 0 || !(tol >= 0.0) ||
        !(gmres_tol > 0.0) || gmres_maxiter < 1)
        return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const zj_params prm = zj::solver_params(g->s);
    SrcPrm p{prm.g, prm.manning_n * prm.manning_n, prm.eps_wet, g->alpha, dt, g->N, g->dim};
    double* corr = g->dX;  // the Newton correction (momentum rows)
    double* Qmom = Qout + (size_t)2 * g->N;
    IHIPCHK(hipMemcpyAsync(Qout, Qexp, sizeof(double) * 4 * g->N, hipMemcpyDeviceToDevice, st));
    if (int rc = zj_update_aux(g->s, Qout, g->aux, st)) return rc;
    k_friction<<<IGRID, IB, 0, st>>>(p, Qout, g->kf);
    double rn = 0.0;
    if (int rc = residual(g, p, Qout, Qexp, &rn, st)) return rc;
    g->newton_its = 0;
    g->gmres_its = 0;
    g->fallbacks = 0;
    int rc = ZJ_OK;
    while (g->newton_its < maxiter && rn > tol) {
        int its = 0, ok = 0;
        rc = gmres(g, p, Qout, corr, gmres_tol, gmres_maxiter, &its, &ok, st);
        if (rc) break;
        g->gmres_its += its;
        if (!ok) {
            // GMRES missed its tolerance: the step falls back to -R, as the
            // reference's jnp.where(info == 0, delta, b)
            k_scale<<<IGRID, IB, 0, st>>>(g->n, -1.0, g->R, corr);
            g->fallbacks += 1;
        }
        Coefs c{};
        c.a[0] = 1.0;
        k_maxpy<<<IGRID, IB, 0, st>>>(g->n, 1, c, 1.0, corr, Qmom);
        if ((rc = residual(g, p, Qout, Qexp, &rn, st))) break;
        g->newton_its += 1;
    }
    g->rnorm = rn;
    return rc;
}

int zj_imex_global_stats(zj_imex_global* g, int32_t* out3, double* rnorm) {
    if (!g || !out3 || !rnorm) return ZJ_EINVAL;
    out3[0] = g->newton_its;
    out3[1] = g->gmres_its;
    out3[2] = g->fallbacks;
    *rnorm = g->rnorm;
    return ZJ_OK;
}

}  // extern "C"
