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
// GMRES(m) with classical Gram-Schmidt and one reorthogonalisation.  The
// dot products, the Hessenberg columns, the Givens rotations, the
// back-substitution and the convergence test stay on the device (one-thread
// kernels on a GmresDev block); the host enqueues Arnoldi iterations in
// chunks (the first as long as the previous solve took) and reads the
// block's flags once per chunk, and the Arnoldi kernels of a chunk return at
// once after convergence.  Every reduction stores one partial per block and
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
                                                double* __restrict__ out,
                                                const int* __restrict__ stop) {
    if (stop && *stop) return;
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
                                             double* __restrict__ part,
                                             const int* __restrict__ stop) {
    if (stop && *stop) return;
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
                                             double* __restrict__ out,
                                             const int* __restrict__ stop) {
    if (stop && *stop) return;
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

// GMRES state on the device (one restart cycle): the Hessenberg columns
// H[i * MAXK + k], the rotations, the rotated right-hand side e, the
// solution coefficients y, the two Gram-Schmidt passes' dot products d1, d2
// and the norm dot dn; stop ends the Arnoldi kernels of the cycle, k counts
// its iterations, its every iteration of the solve, conv marks convergence.
struct GmresDev {
    double H[(MAXK + 1) * MAXK];
    double cs[MAXK], sn[MAXK], e[MAXK + 1], y[MAXK];
    double d1[MAXK + 1], d2[MAXK + 1];
    double dn, hn, beta, target;
    int stop, k, its, conv;
};

// w += sign sum_i coef_i v_i over m columns (m from *mdev when given)
__global__ __launch_bounds__(IB) void k_maxpy_dev(int n, int m, const int* __restrict__ mdev,
                                                  const double* __restrict__ coef, double sign,
                                                  const double* __restrict__ Vb,
                                                  double* __restrict__ w,
                                                  const int* __restrict__ stop) {
    if (stop && *stop) return;
    if (mdev) m = *mdev;
    for (int j = blockIdx.x * IB + threadIdx.x; j < n; j += gridDim.x * IB) {
        double x = w[j];
        for (int i = 0; i < m; ++i) x += (sign * coef[i]) * Vb[(size_t)i * n + j];
        w[j] = x;
    }
}

// x *= 1 / s with s = beta (which == 0) or hn (which == 1), unless stopped
// or s == 0
__global__ __launch_bounds__(IB) void k_gm_unit(int n, const GmresDev* __restrict__ G, int which,
                                                double* __restrict__ x) {
    if (G->stop) return;
    const double sv = which ? G->hn : G->beta;
    if (!(sv > 0.0)) return;
    const double a = 1.0 / sv;
    for (int j = blockIdx.x * IB + threadIdx.x; j < n; j += gridDim.x * IB) x[j] = a * x[j];
}

// Start of a cycle: beta = |r| from dn, the target of the solve on the first
// cycle (tol |b|), e = (beta, 0, ...); converged (stop) when |r| <= target
// (first cycle: when b = 0).
__global__ void k_gm_cycle(GmresDev* G, double tol, int first) {
    const double beta = sqrt(fmax(G->dn, 0.0));
    G->beta = beta;
    if (first) {
        G->target = tol * beta;
        G->its = 0;
    }
    for (int i = 0; i <= MAXK; ++i) G->e[i] = 0.0;
    G->e[0] = beta;
    G->k = 0;
    G->conv = first ? beta == 0.0 : beta <= G->target;
    G->stop = G->conv;
}

// Arnoldi iteration k, after both Gram-Schmidt passes and the norm dot: the
// Hessenberg column (h = d1 + d2, h_{k+1} = |v_{k+1}|), the previous
// rotations applied, a new one, e rotated; converged when |e_{k+1}| <=
// target, stop then or on breakdown (|v_{k+1}| = 0).
__global__ void k_gm_givens(GmresDev* G, int k) {
    if (G->stop) return;
    double h[MAXK + 1];
    for (int i = 0; i <= k; ++i) h[i] = G->d1[i] + G->d2[i];
    const double hn = sqrt(fmax(G->dn, 0.0));
    G->hn = hn;
    h[k + 1] = hn;
    for (int i = 0; i < k; ++i) {
        const double t = G->cs[i] * h[i] + G->sn[i] * h[i + 1];
        h[i + 1] = -G->sn[i] * h[i] + G->cs[i] * h[i + 1];
        h[i] = t;
    }
    const double r = hypot(h[k], h[k + 1]);
    G->cs[k] = r > 0.0 ? h[k] / r : 1.0;
    G->sn[k] = r > 0.0 ? h[k + 1] / r : 0.0;
    h[k] = r;
    G->e[k + 1] = -G->sn[k] * G->e[k];
    G->e[k] = G->cs[k] * G->e[k];
    for (int i = 0; i <= k; ++i) G->H[(size_t)i * MAXK + k] = h[i];
    G->k = k + 1;
    G->its += 1;
    G->conv = fabs(G->e[k + 1]) <= G->target;
    if (G->conv || hn == 0.0) G->stop = 1;
}

// End of a cycle: y = H^-1 e over its k iterations (upper triangular).
__global__ void k_gm_finish(GmresDev* G) {
    const int k = G->k;
    for (int i = k - 1; i >= 0; --i) {
        double sv = G->e[i];
        for (int j = i + 1; j < k; ++j) sv -= G->H[(size_t)i * MAXK + j] * G->y[j];
        G->y[i] = sv / G->H[(size_t)i * MAXK + i];
    }
}

}  // namespace

struct zj_imex_global {
    zj_solver* s = nullptr;
    int N = 0, n = 0, dim = 2, m = 20;
    double alpha = 0.0;
    double *aux = nullptr, *kf = nullptr, *R = nullptr, *W = nullptr, *Vb = nullptr,
           *dX = nullptr, *part = nullptr, *dots = nullptr, *host = nullptr;
    GmresDev* gd = nullptr;       // the GMRES block on the device
    GmresDev* gd_host = nullptr;  // its pinned copy (flags read per chunk)
    int last_its = 4;             // the first chunk of the next solve
    int newton_its = 0, gmres_its = 0, fallbacks = 0;
    double rnorm = 0.0;
};

namespace {

// R(Q) into g->R and its norm (g->aux holds aux(Q), whose dh/dx_d rows do
// not change during the stage)
int residual(zj_imex_global* g, const SrcPrm& p, const double* Q, const double* Qe,
             double* rnorm, hipStream_t st) {
    k_src_residual<<<IGRID, IB, 0, st>>>(p, Q, Qe, g->aux, g->kf, g->R, g->part);
    k_fold<<<1, IB, 0, st>>>(IGRID, g->part, g->dots, nullptr);
    IHIPCHK(hipMemcpyAsync(g->host, g->dots, sizeof(double), hipMemcpyDeviceToHost, st));
    IHIPCHK(hipStreamSynchronize(st));
    *rnorm = std::sqrt(std::max(g->host[0], 0.0));
    return ZJ_OK;
}

// GMRES(m) for J x = b (b = -R in g->R's negation), x accumulated into dx
// (zero start), relative tolerance tol, at most maxres restart cycles; the
// iteration state lives in g->gd (see GmresDev).  Sets *converged when the
// tolerance was met.
int gmres(zj_imex_global* g, const SrcPrm& p, const double* Q, double* dx, double tol,
          int maxres, int* its, int* converged, hipStream_t st) {
    const int n = g->n, m = g->m;
    double* V = g->Vb;                     // (m + 1) basis vectors
    double* w = g->W;
    GmresDev* G = g->gd;
    const int* stop = &G->stop;
    IHIPCHK(hipMemsetAsync(dx, 0, sizeof(double) * n, st));
    *converged = 0;
    *its = 0;
    auto read = [&]() -> int {
        IHIPCHK(hipMemcpyAsync(g->gd_host, G, sizeof(GmresDev), hipMemcpyDeviceToHost, st));
        IHIPCHK(hipStreamSynchronize(st));
        return ZJ_OK;
    };
    // one Arnoldi iteration k (the kernels return at once after a stop)
    auto arnoldi = [&](int k) {
        double* vk1 = V + (size_t)(k + 1) * n;
        k_src_jvp<<<IGRID, IB, 0, st>>>(p, Q, g->kf, V + (size_t)k * n, vk1, stop);
        // classical Gram-Schmidt, twice
        k_mdot<<<IGRID, IB, 0, st>>>(n, k, vk1, V, g->part, stop);
        k_fold<<<k + 1, IB, 0, st>>>(IGRID, g->part, G->d1, stop);
        k_maxpy_dev<<<IGRID, IB, 0, st>>>(n, k + 1, nullptr, G->d1, -1.0, V, vk1, stop);
        k_mdot<<<IGRID, IB, 0, st>>>(n, k, vk1, V, g->part, stop);
        k_fold<<<k + 1, IB, 0, st>>>(IGRID, g->part, G->d2, stop);
        k_maxpy_dev<<<IGRID, IB, 0, st>>>(n, k + 1, nullptr, G->d2, -1.0, V, vk1, stop);
        k_mdot<<<IGRID, IB, 0, st>>>(n, 0, vk1, vk1, g->part, stop);
        k_fold<<<1, IB, 0, st>>>(IGRID, g->part, &G->dn, stop);
        k_gm_givens<<<1, 1, 0, st>>>(G, k);
        k_gm_unit<<<IGRID, IB, 0, st>>>(n, G, 1, vk1);
    };
    for (int cyc = 0; cyc < maxres; ++cyc) {
        // V0 = r = b - J dx
        k_scale<<<IGRID, IB, 0, st>>>(n, -1.0, g->R, V);
        if (cyc > 0) {
            k_src_jvp<<<IGRID, IB, 0, st>>>(p, Q, g->kf, dx, w, nullptr);
            Coefs c1{};
            c1.a[0] = -1.0;
            k_maxpy<<<IGRID, IB, 0, st>>>(n, 1, c1, 1.0, w, V);
        }
        k_mdot<<<IGRID, IB, 0, st>>>(n, 0, V, V, g->part, nullptr);
        k_fold<<<1, IB, 0, st>>>(IGRID, g->part, &G->dn, nullptr);
        k_gm_cycle<<<1, 1, 0, st>>>(G, tol, cyc == 0);
        k_gm_unit<<<IGRID, IB, 0, st>>>(n, G, 0, V);
        int launched = 0;
        int chunk = std::max(1, std::min(m, g->last_its));
        while (true) {
            for (int k = launched; k < launched + chunk; ++k) arnoldi(k);
            IHIPCHK(hipGetLastError());
            launched += chunk;
            if (int rc = read()) return rc;
            if (g->gd_host->stop || launched >= m) break;
            chunk = std::min(2, m - launched);
        }
        // the cycle converged on its start (stop with no iteration) or after
        // its iterations: y and dx += V y
        if (g->gd_host->k > 0) {
            k_gm_finish<<<1, 1, 0, st>>>(G);
            k_maxpy_dev<<<IGRID, IB, 0, st>>>(n, 0, &G->k, G->y, 1.0, V, dx, nullptr);
            IHIPCHK(hipGetLastError());
        }
        *its = g->gd_host->its;
        // otherwise (the cycle's m iterations, or a breakdown short of the
        // target) the next cycle restarts from the true residual
        if (g->gd_host->conv) {
            *converged = 1;
            break;
        }
    }
    g->last_its = std::max(1, *its);
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
        hipHostMalloc((void**)&g->host, sizeof(double) * (MAXK + 1)) != hipSuccess ||
        hipMalloc((void**)&g->gd, sizeof(GmresDev)) != hipSuccess ||
        hipHostMalloc((void**)&g->gd_host, sizeof(GmresDev)) != hipSuccess) {
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
    if (g->gd) (void)hipFree(g->gd);
    if (g->gd_host) (void)hipHostFree(g->gd_host);
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
