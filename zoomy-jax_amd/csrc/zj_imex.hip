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
// GMRES(m) with classical Gram-Schmidt and one reorthogonalisation.  An
// Arnoldi iteration is three one-cell-per-thread kernels, each emitting its
// dot products as block partials: the Jacobian product of v_k (scaled by
// the norm of the previous iteration on the way, so the normalisation costs
// no pass of its own) with <w, v_i>; the first Gram-Schmidt update with the
// second pass's dots; the second update with <w, w>.  Fixed-order fold
// kernels sum the partials; the Hessenberg column, the Givens rotations and
// the convergence test run in the last fold's block, and the
// back-substitution in every block of the solution update.  The host
// enqueues Arnoldi iterations in chunks (the first as long as the previous
// solve took) and reads four flags once per chunk; the kernels of a chunk
// return at once after convergence.  The stage is deterministic.
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

// Start of the stage, one cell per thread: Q = Qexp (the 4 rows) and
// kf[c] = g n^2 hinv(h_c)^(7/3) (h does not change during the stage)
__global__ __launch_bounds__(IB) void k_start(SrcPrm p, const double* __restrict__ Qe,
                                              double* __restrict__ Q, double* __restrict__ kf) {
    const int c = blockIdx.x * IB + threadIdx.x;
    if (c >= p.N) return;
    for (int v = 0; v < 4; ++v) Q[(size_t)v * p.N + c] = Qe[(size_t)v * p.N + c];
    kf[c] = p.g * p.n2 * pow(hinv_d(Qe[(size_t)p.N + c], p.eps), 7.0 / 3.0);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// *out = the sum of v over a block of B threads (wave sums, then the waves
// in order)
template <int B>
__device__ __forceinline__ void block_sum(double v, double* lds, double* out) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < B / 64; ++i) s += lds[i];
        *out = s;
    }
    __syncthreads();
}

// Newton update and residual, one cell per thread: Q_d += corr_d (corr =
// -R_d, the residual of the last call, when neg_r; nothing when corr is
// null and neg_r is 0), then R = Q - Qexp - dt S(Q, aux) in the momentum
// rows (R[d N + c], d < dim; the b and h rows of Q equal Qexp's);
// part[block] = sum of R^2 over the block
__global__ __launch_bounds__(IB) void k_src_residual(SrcPrm p, double* __restrict__ Q,
                                                     const double* __restrict__ Qe,
                                                     const double* __restrict__ aux,
                                                     const double* __restrict__ kf,
                                                     const double* __restrict__ corr, int neg_r,
                                                     double* __restrict__ R,
                                                     double* __restrict__ part) {
    __shared__ double lds[IB / 64];
    double acc = 0.0;
    const int N = p.N;
    const int c = blockIdx.x * IB + threadIdx.x;
    if (c < N) {
        double q[4] = {0.0, 0.0, 0.0, 0.0}, hx[2] = {0.0, 0.0};
        for (int v = 1; v < 4; ++v) q[v] = Q[(size_t)v * N + c];
        if (corr || neg_r) {
            for (int d = 0; d < p.dim; ++d) {
                const size_t j = (size_t)d * N + c;
                q[2 + d] += neg_r ? -R[j] : corr[j];
                Q[(size_t)(2 + d) * N + c] = q[2 + d];
            }
        }
        for (int d = 0; d < p.dim; ++d) hx[d] = aux[(size_t)row_dh(d) * N + c];
        double S[2];
        source_at(p, q, kf[c], hx, S);
        for (int d = 0; d < p.dim; ++d) {
            const double r = q[2 + d] - Qe[(size_t)(2 + d) * N + c] - p.dt * S[d];
            R[(size_t)d * N + c] = r;
            acc += r * r;
        }
    }
    block_sum<IB>(acc, lds, part + blockIdx.x);
}

// (J v)_d at cell c = v_d - dt (dS/dq v)_d for momentum vectors v: with
// v_h = 0 the terms of dh and of aux(v) (whose dh/dx_d rows are the
// gradients of v_h) vanish; q2, q3 the momenta, k the cell's kf
__device__ __forceinline__ void jvp_at(const SrcPrm& p, double q2, double q3, double k,
                                       const double* v, double* out) {
    const double qq = sqrt(q2 * q2 + q3 * q3);
    // d(|q| k) = k (q.dq / |q|) at fixed h
    const double dqq = qq > 0.0 ? (q2 * v[0] + (p.dim == 2 ? q3 * v[1] : 0.0)) / qq : 0.0;
    const double dm = k * dqq;
    const double m = k * qq;
    const double qd[2] = {q2, q3};
    for (int d = 0; d < p.dim; ++d) out[d] = v[d] - p.dt * (-(dm * qd[d] + m * v[d]));
}

// GMRES state on the device (one restart cycle): the Hessenberg columns
// H[i * MAXK + k], the rotations, the rotated right-hand side e, the
// Gram-Schmidt passes' dot products d1, d2 and the norm dot dn of the
// current iteration, hn = |w| of the last one (v_{k+1} is stored unscaled
// and scaled by 1 / hn when the next iteration reads it; v_0 by 1 / beta);
// stop ends the Arnoldi kernels of the cycle, k counts its iterations, its
// every iteration of the solve, conv marks convergence.
struct GmresDev {
    double H[(MAXK + 1) * MAXK];
    double cs[MAXK], sn[MAXK], e[MAXK + 1];
    double d1[MAXK + 1], d2[MAXK + 1];
    double dn, hn, beta, target;
    int stop, k, its, conv;
};

// Block partials of the dots of a one-cell-per-thread kernel: every thread
// of the block calls emit(v, i) for the same i in turn (0 for cells past
// the end), then finish(nd) stores part[i * gridDim + block], i < nd.
struct BlockDots {
    double (*lds)[IB / 64];
    __device__ __forceinline__ void emit(double v, int i) const {
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) lds[i][threadIdx.x >> 6] = v;
    }
    __device__ __forceinline__ void finish(int nd, double* __restrict__ part) const {
        __syncthreads();
        if ((int)threadIdx.x < nd) {
            double s = 0.0;
            for (int j = 0; j < IB / 64; ++j) s += lds[threadIdx.x][j];
            part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
        }
    }
};

// The cell's momentum entries of basis vector i (zero past the end)
__device__ __forceinline__ void load_v(const double* Vb, int i, int n, int N, int c, int dim,
                                       double* v) {
    v[0] = 0.0;
    v[1] = 0.0;
    if (c < N)
        for (int d = 0; d < dim; ++d) v[d] = Vb[(size_t)i * n + (size_t)d * N + c];
}

// Start of a cycle, one cell per thread: v_0 = -R (first) or -R - J dx;
// the partials of <v_0, v_0>
__global__ __launch_bounds__(IB) void k_gm_start(SrcPrm p, const double* __restrict__ Q,
                                                 const double* __restrict__ kf,
                                                 const double* __restrict__ R,
                                                 const double* __restrict__ dx, int first,
                                                 double* __restrict__ V0,
                                                 double* __restrict__ part) {
    __shared__ double lds[1][IB / 64];
    const BlockDots bd{lds};
    const int N = p.N, c = blockIdx.x * IB + threadIdx.x;
    double acc = 0.0;
    if (c < N) {
        double v[2] = {0.0, 0.0}, jv[2] = {0.0, 0.0};
        for (int d = 0; d < p.dim; ++d) v[d] = -R[(size_t)d * N + c];
        if (!first) {
            double x[2] = {0.0, 0.0};
            for (int d = 0; d < p.dim; ++d) x[d] = dx[(size_t)d * N + c];
            jvp_at(p, Q[(size_t)2 * N + c], Q[(size_t)3 * N + c], kf[c], x, jv);
        }
        for (int d = 0; d < p.dim; ++d) {
            const double r = v[d] - jv[d];
            V0[(size_t)d * N + c] = r;
            acc += r * r;
        }
    }
    bd.emit(acc, 0);
    bd.finish(1, part);
}

// Arnoldi iteration k, first kernel, one cell per thread: v_k *= 1 / beta
// (k = 0) or 1 / hn (the deferred normalisation, written back), w = J v_k
// into v_{k+1}; the partials of <w, v_i>, i <= k
__global__ __launch_bounds__(IB) void k_arn_jvp(SrcPrm p, const double* __restrict__ Q,
                                                const double* __restrict__ kf, double* Vb,
                                                int n, int k, const GmresDev* __restrict__ G,
                                                double* __restrict__ part) {
    if (G->stop) return;
    __shared__ double lds[MAXK + 1][IB / 64];
    const BlockDots bd{lds};
    const int N = p.N, c = blockIdx.x * IB + threadIdx.x;
    const double sv = k == 0 ? G->beta : G->hn;
    const double s = sv > 0.0 ? 1.0 / sv : 1.0;
    double v[2], w[2] = {0.0, 0.0};
    load_v(Vb, k, n, N, c, p.dim, v);
    if (c < N) {
        for (int d = 0; d < p.dim; ++d) {
            v[d] *= s;
            Vb[(size_t)k * n + (size_t)d * N + c] = v[d];
        }
        jvp_at(p, Q[(size_t)2 * N + c], Q[(size_t)3 * N + c], kf[c], v, w);
        for (int d = 0; d < p.dim; ++d) Vb[(size_t)(k + 1) * n + (size_t)d * N + c] = w[d];
    }
    for (int i = 0; i < k; ++i) {
        double u[2];
        load_v(Vb, i, n, N, c, p.dim, u);
        bd.emit(w[0] * u[0] + w[1] * u[1], i);
    }
    bd.emit(w[0] * v[0] + w[1] * v[1], k);
    bd.finish(k + 1, part);
}

// Arnoldi iteration k, a Gram-Schmidt pass, one cell per thread: w -=
// sum_i coef_i v_i (i <= k, coef = d1 or d2); then the partials of
// <w, v_i>, i <= k (pass 1; the v_i read again, from cache) or of <w, w>
// (pass 2)
template <int PASS>
__global__ __launch_bounds__(IB) void k_arn_gs(double* Vb, int n, int N, int dim, int k,
                                               const GmresDev* __restrict__ G,
                                               double* __restrict__ part) {
    if (G->stop) return;
    __shared__ double lds[PASS == 1 ? MAXK + 1 : 1][IB / 64];
    const BlockDots bd{lds};
    const int c = blockIdx.x * IB + threadIdx.x;
    const double* coef = PASS == 1 ? G->d1 : G->d2;
    double w[2];
    load_v(Vb, k + 1, n, N, c, dim, w);
    for (int i = 0; i <= k; ++i) {
        double u[2];
        load_v(Vb, i, n, N, c, dim, u);
        const double a = coef[i];
        w[0] -= a * u[0];
        w[1] -= a * u[1];
    }
    if (c < N)
        for (int d = 0; d < dim; ++d) Vb[(size_t)(k + 1) * n + (size_t)d * N + c] = w[d];
    if (PASS == 1) {
        for (int i = 0; i <= k; ++i) {
            double u[2];
            load_v(Vb, i, n, N, c, dim, u);
            bd.emit(w[0] * u[0] + w[1] * u[1], i);
        }
        bd.finish(k + 1, part);
    } else {
        bd.emit(w[0] * w[0] + w[1] * w[1], 0);
        bd.finish(1, part);
    }
}

constexpr int FB = 1024;  // threads of a fold block

// out[i] = sum of the nb partials of dot i (block i), in a fixed order
__global__ __launch_bounds__(FB) void k_fold(int nb, const double* __restrict__ part,
                                             double* __restrict__ out,
                                             const int* __restrict__ stop) {
    if (stop && *stop) return;
    __shared__ double lds[FB / 64];
    const double* pp = part + (size_t)blockIdx.x * nb;
    double acc = 0.0;
    for (int j = threadIdx.x; j < nb; j += FB) acc += pp[j];
    block_sum<FB>(acc, lds, out + blockIdx.x);
}

// Start of a cycle (one block): dn = |v_0|^2 from the partials, beta, the
// target of the solve on the first cycle (tol |b|), e = (beta, 0, ...);
// converged (stop) when |r| <= target (first cycle: when b = 0).
__global__ __launch_bounds__(FB) void k_gm_cycle(int nb, const double* __restrict__ part,
                                                 GmresDev* G, double tol, int first) {
    __shared__ double lds[FB / 64];
    double acc = 0.0;
    for (int j = threadIdx.x; j < nb; j += FB) acc += part[j];
    block_sum<FB>(acc, lds, &G->dn);
    if (threadIdx.x != 0) return;
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

// End of Arnoldi iteration k (one block): dn = |w|^2 from the partials; the
// Hessenberg column (h = d1 + d2, h_{k+1} = |w|), the previous rotations
// applied, a new one, e rotated; converged when |e_{k+1}| <= target, stop
// then or on breakdown (|w| = 0).
__global__ __launch_bounds__(FB) void k_gm_givens(int nb, const double* __restrict__ part,
                                                  GmresDev* G, int k) {
    if (G->stop) return;
    __shared__ double lds[FB / 64];
    double acc = 0.0;
    for (int j = threadIdx.x; j < nb; j += FB) acc += part[j];
    block_sum<FB>(acc, lds, &G->dn);
    if (threadIdx.x != 0) return;
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

// End of a cycle, one cell per thread: y = H^-1 e over the cycle's k
// iterations (upper triangular; every block solves it alike in LDS), then
// dx = V y (first cycle) or dx += V y
__global__ __launch_bounds__(IB) void k_gm_update(const GmresDev* __restrict__ G,
                                                  const double* __restrict__ Vb, int n, int N,
                                                  int dim, int first, double* __restrict__ dx) {
    __shared__ double y[MAXK];
    const int k = G->k;
    if (threadIdx.x == 0) {
        for (int i = k - 1; i >= 0; --i) {
            double sv = G->e[i];
            for (int j = i + 1; j < k; ++j) sv -= G->H[(size_t)i * MAXK + j] * y[j];
            y[i] = sv / G->H[(size_t)i * MAXK + i];
        }
    }
    __syncthreads();
    const int c = blockIdx.x * IB + threadIdx.x;
    if (c >= N) return;
    for (int d = 0; d < dim; ++d) {
        const size_t j = (size_t)d * N + c;
        double x = first ? 0.0 : dx[j];
        for (int i = 0; i < k; ++i) x += y[i] * Vb[(size_t)i * n + j];
        dx[j] = x;
    }
}

}  // namespace

struct zj_imex_global {
    zj_solver* s = nullptr;
    int N = 0, n = 0, dim = 2, m = 20, nb = 0;
    double alpha = 0.0;
    double *aux = nullptr, *kf = nullptr, *R = nullptr, *Vb = nullptr, *dX = nullptr,
           *part = nullptr, *dots = nullptr, *host = nullptr;
    GmresDev* gd = nullptr;  // the GMRES block on the device
    int* flags = nullptr;    // pinned copy of its stop, k, its, conv
    int last_its = 4;        // the first chunk of the next solve
    int newton_its = 0, gmres_its = 0, fallbacks = 0;
    double rnorm = 0.0;
};

namespace {

// Q_mom += corr (or -R when neg_r; neither when both are off), then R(Q)
// into g->R and its norm (g->aux holds aux(Q), whose dh/dx_d rows do not
// change during the stage)
int residual(zj_imex_global* g, const SrcPrm& p, double* Q, const double* Qe,
             const double* corr, int neg_r, double* rnorm, hipStream_t st) {
    k_src_residual<<<g->nb, IB, 0, st>>>(p, Q, Qe, g->aux, g->kf, corr, neg_r, g->R, g->part);
    k_fold<<<1, FB, 0, st>>>(g->nb, g->part, g->dots, nullptr);
    IHIPCHK(hipMemcpyAsync(g->host, g->dots, sizeof(double), hipMemcpyDeviceToHost, st));
    IHIPCHK(hipStreamSynchronize(st));
    *rnorm = std::sqrt(std::max(g->host[0], 0.0));
    return ZJ_OK;
}

// GMRES(m) for J x = b (b = -R), x into dx (zero start), relative
// tolerance tol, at most maxres restart cycles; the iteration state lives
// in g->gd (see GmresDev).  Sets *converged when the tolerance was met.
int gmres(zj_imex_global* g, const SrcPrm& p, const double* Q, double* dx, double tol,
          int maxres, int* its, int* converged, hipStream_t st) {
    const int n = g->n, m = g->m, nb = g->nb, N = g->N, dim = g->dim;
    double* V = g->Vb;  // (m + 1) basis vectors
    GmresDev* G = g->gd;
    const int* stop = &G->stop;
    int* fl = g->flags;  // stop, k, its, conv
    *converged = 0;
    *its = 0;
    bool have_dx = false;
    for (int cyc = 0; cyc < maxres; ++cyc) {
        k_gm_start<<<nb, IB, 0, st>>>(p, Q, g->kf, g->R, dx, !have_dx, V, g->part);
        k_gm_cycle<<<1, FB, 0, st>>>(nb, g->part, G, tol, cyc == 0);
        int launched = 0;
        int chunk = std::max(1, std::min(m, g->last_its));
        while (true) {
            for (int k = launched; k < launched + chunk; ++k) {
                k_arn_jvp<<<nb, IB, 0, st>>>(p, Q, g->kf, V, n, k, G, g->part);
                k_fold<<<k + 1, FB, 0, st>>>(nb, g->part, G->d1, stop);
                k_arn_gs<1><<<nb, IB, 0, st>>>(V, n, N, dim, k, G, g->part);
                k_fold<<<k + 1, FB, 0, st>>>(nb, g->part, G->d2, stop);
                k_arn_gs<2><<<nb, IB, 0, st>>>(V, n, N, dim, k, G, g->part);
                k_gm_givens<<<1, FB, 0, st>>>(nb, g->part, G, k);
            }
            IHIPCHK(hipGetLastError());
            launched += chunk;
            IHIPCHK(hipMemcpyAsync(fl, &G->stop, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
            IHIPCHK(hipStreamSynchronize(st));
            if (fl[0] || launched >= m) break;
            chunk = std::min(2, m - launched);
        }
        // y and dx (+)= V y over the cycle's iterations (none when the
        // cycle converged on its start)
        if (fl[1] > 0) {
            k_gm_update<<<nb, IB, 0, st>>>(G, V, n, N, dim, !have_dx, dx);
            IHIPCHK(hipGetLastError());
            have_dx = true;
        }
        *its = fl[2];
        // otherwise (the cycle's m iterations, or a breakdown short of the
        // target) the next cycle restarts from the true residual
        if (fl[3]) {
            *converged = 1;
            break;
        }
    }
    if (!have_dx) IHIPCHK(hipMemsetAsync(dx, 0, sizeof(double) * n, st));
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
    g->nb = (g->N + IB - 1) / IB;
    g->alpha = alpha;
    const size_t naux = (size_t)(3 * g->dim + 1) * g->N;
    if (hipMalloc((void**)&g->aux, sizeof(double) * naux) != hipSuccess ||
        hipMalloc((void**)&g->kf, sizeof(double) * g->N) != hipSuccess ||
        hipMalloc((void**)&g->R, sizeof(double) * g->n) != hipSuccess ||
        hipMalloc((void**)&g->dX, sizeof(double) * g->n) != hipSuccess ||
        hipMalloc((void**)&g->Vb, sizeof(double) * g->n * (size_t)(g->m + 1)) != hipSuccess ||
        hipMalloc((void**)&g->part, sizeof(double) * (size_t)g->nb * (MAXK + 1)) != hipSuccess ||
        hipMalloc((void**)&g->dots, sizeof(double) * (MAXK + 1)) != hipSuccess ||
        hipHostMalloc((void**)&g->host, sizeof(double) * (MAXK + 1)) != hipSuccess ||
        hipMalloc((void**)&g->gd, sizeof(GmresDev)) != hipSuccess ||
        hipHostMalloc((void**)&g->flags, 4 * sizeof(int)) != hipSuccess) {
        zj_imex_global_destroy(g);
        return ZJ_ENOMEM;
    }
    *out = g;
    return ZJ_OK;
}

int zj_imex_global_destroy(zj_imex_global* g) {
    if (!g) return ZJ_OK;
    for (double* p : {g->aux, g->kf, g->R, g->Vb, g->dX, g->part, g->dots})
        if (p) (void)hipFree(p);
    if (g->host) (void)hipHostFree(g->host);
    if (g->gd) (void)hipFree(g->gd);
    if (g->flags) (void)hipHostFree(g->flags);
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
    k_start<<<g->nb, IB, 0, st>>>(p, Qexp, Qout, g->kf);
    IHIPCHK(hipGetLastError());
    if (int rc = zj_update_aux(g->s, Qout, g->aux, st)) return rc;
    double rn = 0.0;
    if (int rc = residual(g, p, Qout, Qexp, nullptr, 0, &rn, st)) return rc;
    g->newton_its = 0;
    g->gmres_its = 0;
    g->fallbacks = 0;
    int rc = ZJ_OK;
    while (g->newton_its < maxiter && rn > tol) {
        int its = 0, ok = 0;
        rc = gmres(g, p, Qout, corr, gmres_tol, gmres_maxiter, &its, &ok, st);
        if (rc) break;
        g->gmres_its += its;
        // GMRES missed its tolerance: the step falls back to -R, as the
        // reference's jnp.where(info == 0, delta, b)
        if (!ok) g->fallbacks += 1;
        if ((rc = residual(g, p, Qout, Qexp, ok ? corr : nullptr, !ok, &rn, st))) break;
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
