// Chorin projection on a structured nx x ny quad grid (cell id i*ny + j,
// walls): predictor, pressure operator, geometric multigrid, preconditioned
// CG and corrector, behind the zj_chorin_* C ABI (include/zoomy_hip.h).
//
// Pressure operator (5-point, SPD; a_c = h_c + b_c - b_e, a_e = h_e + b_e -
// b_c, c_f = |f|/d_f hinv(h_f) on the face between cells c and e):
//     S_cc += c_f a_c^2,  S_ce = -c_f a_c a_e,  S_cc += 4 |A_c| hinv(h_c)
// stored per cell as (diag, cE = S[c, c+ny], cN = S[c, c+1]).
//
// Multigrid: coarse levels are re-discretised from 4-cell averages of h and
// b; red-black Gauss-Seidel smoothing (pre: red, black; post: black, red);
// cell-centred bilinear prolongation and its transpose as restriction; the
// coarsest level is smoothed to convergence by one workgroup.  The V-cycle is
// symmetric, so plain CG uses it as preconditioner.
//
// Every CG kernel takes the device flag `done`; once the residual test has
// passed, later kernels of an already launched chunk return at once, so the
// host enqueues iterations in chunks and synchronises once per chunk.
// Reductions are two-pass (per-block partials in a fixed order, then one
// workgroup), hence bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/zoomy_hip.h"
#include "zj_chorin.hpp"
#include "zj_device.hpp"

using namespace zj;

namespace {

constexpr int CB = 256;      // threads per block of the grid kernels
constexpr int RED_MAX = 1024;

__device__ __forceinline__ bool is_done(const int* done) { return done && *done; }

// ───────────────────────── hierarchy: averages and stencils ─────────────────
// One level of the hierarchy.  On level 0 (avg == 0) h and b are the
// level's own depth and bed; on coarser levels (avg != 0) they are the finer
// level's, and each cell forms the 4-cell averages 0.25 (((f00 + f10) + f01)
// + f11) of itself and of its four neighbours (the same sums, so the same
// values, as a separate averaging pass would store), storing its own in
// (hc, bc) for the next level.  The stencil coefficients (diag, cE, cN) go
// to the FP64 arrays and, when given, rounded to the FP32 arrays of the
// low-precision hierarchy.
__global__ __launch_bounds__(CB) void k_level(int nx, int ny, double dx, double dy, double eps,
                                              const double* __restrict__ h,
                                              const double* __restrict__ b, int avg,
                                              double* __restrict__ hc_out,
                                              double* __restrict__ bc_out,
                                              double* __restrict__ diag, double* __restrict__ cE,
                                              double* __restrict__ cN, float* __restrict__ diag32,
                                              float* __restrict__ cE32, float* __restrict__ cN32) {
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= nx * ny) return;
    int i = c / ny, j = c - i * ny;
    const int nyf = 2 * ny;
    auto val = [&](const double* __restrict__ a, int q) {
        if (!avg) return a[q];
        const int I = q / ny, J = q - I * ny;
        const int f00 = (2 * I) * nyf + 2 * J, f10 = f00 + nyf;
        return 0.25 * (((a[f00] + a[f10]) + a[f00 + 1]) + a[f10 + 1]);
    };
    double hc = val(h, c), bc = val(b, c);
    if (avg) {
        hc_out[c] = hc;
        bc_out[c] = bc;
    }
    double d = 4.0 * dx * dy * hinv_of(hc, eps);
    double e = 0.0, n = 0.0;
    const double rx = dy / dx, ry = dx / dy;
    // faces to the four neighbours: (offset, ratio |f|/d_f, is the E/N face)
    if (i + 1 < nx) {
        int o = c + ny;
        const double ho = val(h, o), bo = val(b, o);
        double cf = rx * hinv_of(0.5 * (hc + ho), eps);
        double ac = hc + bc - bo, ae = ho + bo - bc;
        d += cf * ac * ac;
        e = -cf * ac * ae;
    }
    if (i > 0) {
        int o = c - ny;
        const double ho = val(h, o), bo = val(b, o);
        double cf = rx * hinv_of(0.5 * (ho + hc), eps);
        double ac = hc + bc - bo;
        d += cf * ac * ac;
    }
    if (j + 1 < ny) {
        int o = c + 1;
        const double ho = val(h, o), bo = val(b, o);
        double cf = ry * hinv_of(0.5 * (hc + ho), eps);
        double ac = hc + bc - bo, ae = ho + bo - bc;
        d += cf * ac * ac;
        n = -cf * ac * ae;
    }
    if (j > 0) {
        int o = c - 1;
        const double ho = val(h, o), bo = val(b, o);
        double cf = ry * hinv_of(0.5 * (ho + hc), eps);
        double ac = hc + bc - bo;
        d += cf * ac * ac;
    }
    diag[c] = d;
    cE[c] = e;
    cN[c] = n;
    if (diag32) {
        diag32[c] = (float)d;
        cE32[c] = (float)e;
        cN32[c] = (float)n;
    }
}

// sum_{e != c} S_ce x_e
template <typename T>
__device__ __forceinline__ T offdiag(int c, int i, int j, int nx, int ny, const T* __restrict__ cE,
                                     const T* __restrict__ cN, const T* __restrict__ x) {
    T s = T(0);
    if (i + 1 < nx) s += cE[c] * x[c + ny];
    if (i > 0) s += cE[c - ny] * x[c - ny];
    if (j + 1 < ny) s += cN[c] * x[c + 1];
    if (j > 0) s += cN[c - 1] * x[c - 1];
    return s;
}

// ───────────────────────── multigrid kernels ───────────────────────────────
// One thread per cell of the given colour ((i + j) & 1 == colour).  With
// first != 0 the iterate is zero on entry (colour 0 only).
template <typename T>
__global__ __launch_bounds__(CB) void k_rbgs(int nx, int ny, const T* __restrict__ diag,
                                             const T* __restrict__ cE, const T* __restrict__ cN,
                                             T* __restrict__ x, const T* __restrict__ f, int colour,
                                             int first, const int* __restrict__ done) {
    if (is_done(done)) return;
    const int half = (ny + 1) >> 1;
    int t = blockIdx.x * CB + threadIdx.x;
    if (t >= nx * half) return;
    int i = t / half, k = t - i * half;
    int j = 2 * k + ((i + colour) & 1);
    if (j >= ny) return;
    int c = i * ny + j;
    if (first) {
        x[c] = f[c] / diag[c];
        return;
    }
    x[c] = (f[c] - offdiag(c, i, j, nx, ny, cE, cN, x)) / diag[c];
}

template <typename T>
__global__ __launch_bounds__(CB) void k_resid(int nx, int ny, const T* __restrict__ diag,
                                              const T* __restrict__ cE, const T* __restrict__ cN,
                                              const T* __restrict__ x, const T* __restrict__ f,
                                              T* __restrict__ r, const int* __restrict__ done) {
    if (is_done(done)) return;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= nx * ny) return;
    int i = c / ny, j = c - i * ny;
    r[c] = f[c] - (diag[c] * x[c] + offdiag(c, i, j, nx, ny, cE, cN, x));
}

// 1-D weight of fine cell 2I + d (d in -1..2) in coarse cell I of nc.
__device__ __forceinline__ double rw(int I, int d, int nc) {
    if (d == -1) return I >= 1 ? 0.25 : 0.0;
    if (d == 0) return I == 0 ? 1.0 : 0.75;
    if (d == 1) return I == nc - 1 ? 1.0 : 0.75;
    return I + 1 < nc ? 0.25 : 0.0;
}

template <typename T>
__global__ __launch_bounds__(CB) void k_restrict(int nxc, int nyc, const T* __restrict__ r,
                                                 T* __restrict__ fc, const int* __restrict__ done) {
    if (is_done(done)) return;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= nxc * nyc) return;
    int I = c / nyc, J = c - I * nyc;
    const int nyf = 2 * nyc;
    double s = 0.0;
#pragma unroll
    for (int di = -1; di <= 2; ++di) {
        double wx = rw(I, di, nxc);
        if (wx == 0.0) continue;
        const T* row = r + (size_t)(2 * I + di) * nyf;
        double sr = 0.0;
#pragma unroll
        for (int dj = -1; dj <= 2; ++dj) {
            double wy = rw(J, dj, nyc);
            if (wy != 0.0) sr += wy * (double)row[2 * J + dj];
        }
        s += wx * sr;
    }
    fc[c] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(CB) void k_prolong(int nx, int ny, const T* __restrict__ xc,
                                                T* __restrict__ x, const int* __restrict__ done) {
    if (is_done(done)) return;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= nx * ny) return;
    int i = c / ny, j = c - i * ny;
    const int nxc = nx >> 1, nyc = ny >> 1;
    int I = i >> 1, Io = I + ((i & 1) ? 1 : -1);
    int J = j >> 1, Jo = J + ((j & 1) ? 1 : -1);
    double wI = 1.0, wIo = 0.0, wJ = 1.0, wJo = 0.0;
    if (Io >= 0 && Io < nxc) { wI = 0.75; wIo = 0.25; } else Io = I;
    if (Jo >= 0 && Jo < nyc) { wJ = 0.75; wJo = 0.25; } else Jo = J;
    double v = wI * (wJ * (double)xc[I * nyc + J] + wJo * (double)xc[I * nyc + Jo]) +
               wIo * (wJ * (double)xc[Io * nyc + J] + wJo * (double)xc[Io * nyc + Jo]);
    x[c] = (T)((double)x[c] + v);
}

// Coarsest level: sweeps x (red, black, black, red) from x = 0, one block.
template <typename T>
__global__ __launch_bounds__(CB) void k_coarse(int nx, int ny, const T* __restrict__ diag,
                                               const T* __restrict__ cE, const T* __restrict__ cN,
                                               T* __restrict__ x, const T* __restrict__ f,
                                               int sweeps, const int* __restrict__ done) {
    if (is_done(done)) return;
    const int n = nx * ny;
    for (int c = threadIdx.x; c < n; c += CB) x[c] = T(0);
    __syncthreads();
    for (int s = 0; s < sweeps; ++s) {
        for (int q = 0; q < 4; ++q) {
            const int colour = (q == 0 || q == 3) ? 0 : 1;
            for (int c = threadIdx.x; c < n; c += CB) {
                int i = c / ny, j = c - i * ny;
                if (((i + j) & 1) != colour) continue;
                x[c] = (f[c] - offdiag(c, i, j, nx, ny, cE, cN, x)) / diag[c];
            }
            __syncthreads();
        }
    }
}

// ───────────────────────── fused multigrid kernels (nu = 1) ────────────────
// Same arithmetic, in the same order, as the separate kernels above, so the
// fused V-cycle is bit-identical to the unfused one.
constexpr int TC = 8;                // coarse cells per tile side (8: 4x the
                                     // blocks of 16 at 1/3 the LDS; fastest measured)
constexpr int WD = 2 * TC + 6;       // down-kernel fine window side (halo 3)
constexpr int WU = 2 * TC + 4;       // up-kernel fine window side (halo 2)
constexpr int WC = TC + 4;           // up-kernel coarse window side
constexpr int TAIL_CELLS = 1365;     // cells of all tail levels together
// the tail's LDS: five arrays per level and the residual of one level
static_assert(6 * TAIL_CELLS * sizeof(double) <= 64 * 1024, "tail LDS");
constexpr int TAIL_MAXL = 10;
constexpr int TB = 512;              // tail kernel threads

// offdiag on an LDS window of side W: cell (wi, wj) at global (gi, gj)
template <typename T, int W>
__device__ __forceinline__ T offdiag_lds(int wi, int wj, int gi, int gj, int nx, int ny,
                                         const T* se, const T* sn, const T* sx) {
    const int w = wi * W + wj;
    T s = T(0);
    if (gi + 1 < nx) s += se[w] * sx[w + W];
    if (gi > 0) s += se[w - W] * sx[w - W];
    if (gj + 1 < ny) s += sn[w] * sx[w + 1];
    if (gj > 0) s += sn[w - 1] * sx[w - 1];
    return s;
}

// Pre-smoothing from zero (red, black), residual and restriction of one
// tile of TC x TC coarse cells; writes the smoothed fine iterate of the
// tile's 2TC x 2TC fine cells and the coarse right-hand side (f is only
// read, x and fc only written, so tiles do not interact).
template <typename T>
__global__ __launch_bounds__(CB) void k_mg_down(int nx, int ny, const T* __restrict__ diag,
                                                const T* __restrict__ cE,
                                                const T* __restrict__ cN,
                                                const T* __restrict__ f, T* __restrict__ x,
                                                T* __restrict__ fc,
                                                const int* __restrict__ done) {
    if (is_done(done)) return;
    __shared__ T sf[WD * WD], sd[WD * WD], se[WD * WD], sn[WD * WD], sx[WD * WD];
    const int nxc = nx >> 1, nyc = ny >> 1;
    const int I0 = blockIdx.x * TC, J0 = blockIdx.y * TC;
    const int gi0 = 2 * I0 - 3, gj0 = 2 * J0 - 3;
    const int tid = threadIdx.x;
    {
        // all loads of the window first, then the LDS stores
        constexpr int NIT = (WD * WD + CB - 1) / CB;
        T vf[NIT], vd[NIT], ve[NIT], vn[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int w = tid + k * CB;
            const int wi = w / WD, wj = w - wi * WD;
            const int gi = gi0 + wi, gj = gj0 + wj;
            vf[k] = vd[k] = ve[k] = vn[k] = T(0);
            if (w < WD * WD && gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
                const int c = gi * ny + gj;
                vf[k] = f[c];
                vd[k] = diag[c];
                ve[k] = cE[c];
                vn[k] = cN[c];
            }
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int w = tid + k * CB;
            if (w >= WD * WD) continue;
            const int wi = w / WD, wj = w - wi * WD;
            const int gi = gi0 + wi, gj = gj0 + wj;
            const bool in = gi >= 0 && gi < nx && gj >= 0 && gj < ny;
            sf[w] = vf[k];
            sd[w] = vd[k];
            se[w] = ve[k];
            sn[w] = vn[k];
            sx[w] = (!in || ((gi + gj) & 1)) ? T(0) : vf[k] / vd[k];
        }
    }
    __syncthreads();
    // black sweep on [1, WD-2]^2 (reads only red values)
    for (int w = tid; w < WD * WD; w += CB) {
        const int wi = w / WD, wj = w - wi * WD;
        const int gi = gi0 + wi, gj = gj0 + wj;
        if (wi < 1 || wi > WD - 2 || wj < 1 || wj > WD - 2) continue;
        if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || !((gi + gj) & 1)) continue;
        sx[w] = (sf[w] - offdiag_lds<T, WD>(wi, wj, gi, gj, nx, ny, se, sn, sx)) / sd[w];
    }
    __syncthreads();
    // residual on [2, WD-3]^2, in place of f
    for (int w = tid; w < WD * WD; w += CB) {
        const int wi = w / WD, wj = w - wi * WD;
        const int gi = gi0 + wi, gj = gj0 + wj;
        if (wi < 2 || wi > WD - 3 || wj < 2 || wj > WD - 3) continue;
        if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
        const T r = sf[w] - (sd[w] * sx[w] + offdiag_lds<T, WD>(wi, wj, gi, gj, nx, ny, se, sn, sx));
        sf[w] = r;
    }
    __syncthreads();
    {
        const int I = I0 + tid / TC, J = J0 + (tid % TC);
        if (tid < TC * TC && I < nxc && J < nyc) {
            const int bi = 2 * (I - I0) + 3, bj = 2 * (J - J0) + 3;
            double s = 0.0;
#pragma unroll
            for (int di = -1; di <= 2; ++di) {
                double wx = rw(I, di, nxc);
                if (wx == 0.0) continue;
                const T* row = sf + (bi + di) * WD + bj;
                double sr = 0.0;
#pragma unroll
                for (int dj = -1; dj <= 2; ++dj) {
                    double wy = rw(J, dj, nyc);
                    if (wy != 0.0) sr += wy * (double)row[dj];
                }
                s += wx * sr;
            }
            fc[I * nyc + J] = (T)s;
        }
    }
    for (int u = tid; u < 4 * TC * TC; u += CB) {
        const int li = u / (2 * TC), lj = u - li * (2 * TC);
        const int gi = 2 * I0 + li, gj = 2 * J0 + lj;
        if (gi < nx && gj < ny) x[gi * ny + gj] = sx[(li + 3) * WD + lj + 3];
    }
}

// Prolongation of the coarse correction and post-smoothing (black, red) of
// one tile of 2TC x 2TC fine cells: reads the pre-smoothed iterate xin
// (tile and halo), writes the tile of x (xin and x must not alias).
template <typename T>
__global__ __launch_bounds__(CB) void k_mg_up(int nx, int ny, const T* __restrict__ diag,
                                              const T* __restrict__ cE,
                                              const T* __restrict__ cN,
                                              const T* __restrict__ f,
                                              const T* __restrict__ xin, T* __restrict__ x,
                                              const T* __restrict__ xc,
                                              const int* __restrict__ done) {
    if (is_done(done)) return;
    __shared__ T sf[WU * WU], sd[WU * WU], se[WU * WU], sn[WU * WU], sx[WU * WU];
    __shared__ T sc[WC * WC];
    const int nxc = nx >> 1, nyc = ny >> 1;
    const int I0 = blockIdx.x * TC, J0 = blockIdx.y * TC;
    const int gi0 = 2 * I0 - 2, gj0 = 2 * J0 - 2;
    const int ci0 = I0 - 2, cj0 = J0 - 2;
    const int tid = threadIdx.x;
    {
        constexpr int NC = (WC * WC + CB - 1) / CB;
        constexpr int NIT = (WU * WU + CB - 1) / CB;
        T vc[NC], vf[NIT], vd[NIT], ve[NIT], vn[NIT], vx[NIT];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int w = tid + k * CB;
            const int a = w / WC, b = w - a * WC;
            const int I = ci0 + a, J = cj0 + b;
            vc[k] = (w < WC * WC && I >= 0 && I < nxc && J >= 0 && J < nyc) ? xc[I * nyc + J]
                                                                              : T(0);
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int w = tid + k * CB;
            const int wi = w / WU, wj = w - wi * WU;
            const int gi = gi0 + wi, gj = gj0 + wj;
            vf[k] = vd[k] = ve[k] = vn[k] = vx[k] = T(0);
            if (w < WU * WU && gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
                const int c = gi * ny + gj;
                vf[k] = f[c];
                vd[k] = diag[c];
                ve[k] = cE[c];
                vn[k] = cN[c];
                vx[k] = xin[c];
            }
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int w = tid + k * CB;
            if (w < WC * WC) sc[w] = vc[k];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int w = tid + k * CB;
            if (w >= WU * WU) continue;
            sf[w] = vf[k];
            sd[w] = vd[k];
            se[w] = ve[k];
            sn[w] = vn[k];
            sx[w] = vx[k];
        }
    }
    __syncthreads();
    for (int w = tid; w < WU * WU; w += CB) {
        const int wi = w / WU, wj = w - wi * WU;
        const int i = gi0 + wi, j = gj0 + wj;
        if (i < 0 || i >= nx || j < 0 || j >= ny) continue;
        int I = i >> 1, Io = I + ((i & 1) ? 1 : -1);
        int J = j >> 1, Jo = J + ((j & 1) ? 1 : -1);
        double wI = 1.0, wIo = 0.0, wJ = 1.0, wJo = 0.0;
        if (Io >= 0 && Io < nxc) { wI = 0.75; wIo = 0.25; } else Io = I;
        if (Jo >= 0 && Jo < nyc) { wJ = 0.75; wJo = 0.25; } else Jo = J;
        const int a = I - ci0, ao = Io - ci0, b = J - cj0, bo = Jo - cj0;
        double v = wI * (wJ * (double)sc[a * WC + b] + wJo * (double)sc[a * WC + bo]) +
                   wIo * (wJ * (double)sc[ao * WC + b] + wJo * (double)sc[ao * WC + bo]);
        sx[w] = (T)((double)sx[w] + v);
    }
    __syncthreads();
    for (int colour = 1; colour >= 0; --colour) {
        const int lo = colour ? 1 : 2, hi = colour ? WU - 2 : WU - 3;
        for (int w = tid; w < WU * WU; w += CB) {
            const int wi = w / WU, wj = w - wi * WU;
            const int gi = gi0 + wi, gj = gj0 + wj;
            if (wi < lo || wi > hi || wj < lo || wj > hi) continue;
            if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || ((gi + gj) & 1) != colour) continue;
            sx[w] = (sf[w] - offdiag_lds<T, WU>(wi, wj, gi, gj, nx, ny, se, sn, sx)) / sd[w];
        }
        __syncthreads();
    }
    for (int u = tid; u < 4 * TC * TC; u += CB) {
        const int li = u / (2 * TC), lj = u - li * (2 * TC);
        const int gi = 2 * I0 + li, gj = 2 * J0 + lj;
        if (gi < nx && gj < ny) x[gi * ny + gj] = sx[(li + 2) * WU + lj + 2];
    }
}

template <typename T>
struct TailDesc {
    int nl;
    int nx[TAIL_MAXL], ny[TAIL_MAXL];
    const T *diag[TAIL_MAXL], *cE[TAIL_MAXL], *cN[TAIL_MAXL];
    const T* f;  // right-hand side of the first tail level
    T* x;        // its solution
};

// The V-cycle of the last levels (together <= TAIL_CELLS cells) in one
// workgroup with every array in LDS.  Level sizes and LDS offsets live in
// LDS too (a runtime-indexed kernel-argument array would go to scratch).
template <typename T>
__global__ __launch_bounds__(TB) void k_mg_tail(TailDesc<T> d, int coarse_sweeps,
                                                const int* __restrict__ done) {
    if (is_done(done)) return;
    extern __shared__ __align__(16) unsigned char tail_lds[];
    __shared__ int snx[TAIL_MAXL], sny[TAIL_MAXL], soff[TAIL_MAXL];
    T* base = (T*)tail_lds;
    const int tid = threadIdx.x;
    const int nl = d.nl;
    {
        int off = 0;
#pragma unroll
        for (int l = 0; l < TAIL_MAXL; ++l) {
            if (l >= nl) break;
            const int n = d.nx[l] * d.ny[l];
            if (tid == 0) {
                snx[l] = d.nx[l];
                sny[l] = d.ny[l];
                soff[l] = off;
            }
            T* SD = base + off;
            for (int c = tid; c < n; c += TB) {
                SD[c] = d.diag[l][c];
                SD[n + c] = d.cE[l][c];
                SD[2 * n + c] = d.cN[l][c];
                if (l == 0) SD[4 * n + c] = d.f[c];
            }
            off += 5 * n;
        }
    }
    __syncthreads();
    // arrays of level l: diag, cE, cN, x, f at base + soff[l] + k n; the
    // residual of the level being restricted after the last level
    auto arr = [&](int l, int k) { return base + soff[l] + k * snx[l] * sny[l]; };
    T* SR = base + soff[nl - 1] + 5 * snx[nl - 1] * sny[nl - 1];
    auto od = [&](int nx, int ny, const T* SE, const T* SN, int c, int i, int j, const T* xv) {
        T s = T(0);
        if (i + 1 < nx) s += SE[c] * xv[c + ny];
        if (i > 0) s += SE[c - ny] * xv[c - ny];
        if (j + 1 < ny) s += SN[c] * xv[c + 1];
        if (j > 0) s += SN[c - 1] * xv[c - 1];
        return s;
    };
    const int L = nl - 1;
    for (int l = 0; l < L; ++l) {
        const int nx = snx[l], ny = sny[l], n = nx * ny;
        const T *SD = arr(l, 0), *SE = arr(l, 1), *SN = arr(l, 2);
        T *SX = arr(l, 3), *SF = arr(l, 4);
        for (int c = tid; c < n; c += TB) {
            const int i = c / ny, j = c - i * ny;
            if (!((i + j) & 1)) SX[c] = SF[c] / SD[c];
        }
        __syncthreads();
        for (int c = tid; c < n; c += TB) {
            const int i = c / ny, j = c - i * ny;
            if ((i + j) & 1) SX[c] = (SF[c] - od(nx, ny, SE, SN, c, i, j, SX)) / SD[c];
        }
        __syncthreads();
        // the residual once per cell, then its restriction
        for (int c = tid; c < n; c += TB) {
            const int i = c / ny, j = c - i * ny;
            SR[c] = SF[c] - (SD[c] * SX[c] + od(nx, ny, SE, SN, c, i, j, SX));
        }
        __syncthreads();
        const int nxc = snx[l + 1], nyc = sny[l + 1];
        T* SFc = arr(l + 1, 4);
        for (int cc = tid; cc < nxc * nyc; cc += TB) {
            const int I = cc / nyc, J = cc - I * nyc;
            double s = 0.0;
            for (int di = -1; di <= 2; ++di) {
                double wx = rw(I, di, nxc);
                if (wx == 0.0) continue;
                const int i = 2 * I + di;
                double sr = 0.0;
                for (int dj = -1; dj <= 2; ++dj) {
                    double wy = rw(J, dj, nyc);
                    if (wy == 0.0) continue;
                    sr += wy * (double)SR[i * ny + 2 * J + dj];
                }
                s += wx * sr;
            }
            SFc[cc] = (T)s;
        }
        __syncthreads();
    }
    {
        const int nx = snx[L], ny = sny[L], n = nx * ny;
        const T *SD = arr(L, 0), *SE = arr(L, 1), *SN = arr(L, 2), *SF = arr(L, 4);
        T* SX = arr(L, 3);
        for (int c = tid; c < n; c += TB) SX[c] = T(0);
        __syncthreads();
        for (int sw = 0; sw < coarse_sweeps; ++sw)
            for (int q = 0; q < 4; ++q) {
                const int colour = (q == 0 || q == 3) ? 0 : 1;
                for (int c = tid; c < n; c += TB) {
                    const int i = c / ny, j = c - i * ny;
                    if (((i + j) & 1) != colour) continue;
                    SX[c] = (SF[c] - od(nx, ny, SE, SN, c, i, j, SX)) / SD[c];
                }
                __syncthreads();
            }
    }
    for (int l = L - 1; l >= 0; --l) {
        const int nx = snx[l], ny = sny[l], n = nx * ny;
        const int nxc = snx[l + 1], nyc = sny[l + 1];
        const T *SD = arr(l, 0), *SE = arr(l, 1), *SN = arr(l, 2), *SF = arr(l, 4);
        T* SX = arr(l, 3);
        const T* xc = arr(l + 1, 3);
        for (int c = tid; c < n; c += TB) {
            const int i = c / ny, j = c - i * ny;
            int I = i >> 1, Io = I + ((i & 1) ? 1 : -1);
            int J = j >> 1, Jo = J + ((j & 1) ? 1 : -1);
            double wI = 1.0, wIo = 0.0, wJ = 1.0, wJo = 0.0;
            if (Io >= 0 && Io < nxc) { wI = 0.75; wIo = 0.25; } else Io = I;
            if (Jo >= 0 && Jo < nyc) { wJ = 0.75; wJo = 0.25; } else Jo = J;
            double v = wI * (wJ * (double)xc[I * nyc + J] + wJo * (double)xc[I * nyc + Jo]) +
                       wIo * (wJ * (double)xc[Io * nyc + J] + wJo * (double)xc[Io * nyc + Jo]);
            SX[c] = (T)((double)SX[c] + v);
        }
        __syncthreads();
        for (int colour = 1; colour >= 0; --colour) {
            for (int c = tid; c < n; c += TB) {
                const int i = c / ny, j = c - i * ny;
                if (((i + j) & 1) != colour) continue;
                SX[c] = (SF[c] - od(nx, ny, SE, SN, c, i, j, SX)) / SD[c];
            }
            __syncthreads();
        }
    }
    const int n0 = snx[0] * sny[0];
    const T* SX0 = arr(0, 3);
    for (int c = tid; c < n0; c += TB) d.x[c] = SX0[c];
}

// ───────────────────────── CG kernels ──────────────────────────────────────
// Every reduction kernel stores one partial per block; k_fin folds them in
// a fixed order and performs the scalar step of CG.  (Folding them in the
// last block of the reduction kernel needs a device-scope release fence per
// block, which writes back the L2 of every XCD and costs more than the
// extra launch.)
struct Red {
    double* part;       // partials (2 x RED_MAX)
    double* sc;         // CG_* slots
    int* fl;            // {done, iterations}
    double tol2;
    int maxit;
    int op;             // FIN_* step
};

__device__ void scalar_step(const Red& R, double v) {
    int* fl = R.fl;
    double* sc = R.sc;
    if (R.op != FIN_B2 && fl[0]) return;
    switch (R.op) {
        case FIN_B2:
            sc[CG_B2] = v;
            sc[CG_RR] = v;
            fl[1] = 0;
            fl[0] = !(v > 0.0);
            break;
        case FIN_RR0:  // residual of a non-zero initial guess
            sc[CG_RR] = v;
            if (v <= R.tol2 * sc[CG_B2]) fl[0] = 1;
            break;
        case FIN_RR:
            sc[CG_RR] = v;
            fl[1] += 1;
            if (v <= R.tol2 * sc[CG_B2] || fl[1] >= R.maxit) fl[0] = 1;
            break;
    }
}

// per-block partial sums of v (and v2 when two), folded by k_fin
__device__ void reduce_finish(const Red& R, double v, double v2, bool two) {
    __shared__ double lds[CB / 64];
    auto add = [](double a, double b) { return a + b; };
    v = block_reduce<double, decltype(add), CB>(v, add, lds);
    __syncthreads();
    if (two) v2 = block_reduce<double, decltype(add), CB>(v2, add, lds);
    if (threadIdx.x == 0) {
        R.part[blockIdx.x] = v;
        if (two) R.part[gridDim.x + blockIdx.x] = v2;
    }
}

// Sums the np partials of the preceding reduction in a fixed order and
// performs the scalar step R.op.
__global__ __launch_bounds__(RED_MAX) void k_fin(Red R, int np) {
    if (R.op != FIN_B2 && R.fl[0]) return;
    __shared__ double lds[RED_MAX / 64];
    auto add = [](double a, double b) { return a + b; };
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += RED_MAX) v += R.part[i];
    v = block_reduce<double, decltype(add), RED_MAX>(v, add, lds);
    if (threadIdx.x == 0) scalar_step(R, v);
}

template <typename T>
__global__ __launch_bounds__(CB) void k_spmv(int nx, int ny, const T* __restrict__ diag,
                                             const T* __restrict__ cE, const T* __restrict__ cN,
                                             const T* __restrict__ x, T* __restrict__ y, Red R,
                                             const int* __restrict__ done) {
    if (is_done(done)) return;
    const int n = nx * ny;
    double acc = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB) {
        int i = c / ny, j = c - i * ny;
        T v = diag[c] * x[c] + offdiag(c, i, j, nx, ny, cE, cN, x);
        y[c] = v;
        acc += (double)x[c] * (double)v;
    }
    if (R.part) reduce_finish(R, acc, 0.0, false);
}

template <typename T>
__global__ __launch_bounds__(CB) void k_dot(int n, const T* __restrict__ a, const T* __restrict__ b,
                                            Red R, const int* __restrict__ done) {
    if (is_done(done)) return;
    double acc = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB)
        acc += (double)a[c] * (double)b[c];
    reduce_finish(R, acc, 0.0, false);
}

// r = rhs - S x (zero_x: zero initial guess, x is zeroed); sum of rhs.rhs
template <typename T>
__global__ __launch_bounds__(CB) void k_cg_init(int nx, int ny, const T* __restrict__ diag,
                                                const T* __restrict__ cE,
                                                const T* __restrict__ cN, T* __restrict__ x,
                                                int zero_x, const T* __restrict__ rhs,
                                                T* __restrict__ r, Red R) {
    const int n = nx * ny;
    double acc = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB) {
        T f = rhs[c];
        if (zero_x) {
            x[c] = T(0);
            r[c] = f;
        } else {
            int i = c / ny, j = c - i * ny;
            r[c] = f - (diag[c] * x[c] + offdiag(c, i, j, nx, ny, cE, cN, x));
        }
        acc += (double)f * (double)f;
    }
    reduce_finish(R, acc, 0.0, false);
}

// Sum of the np partials of a preceding reduction, folded in a fixed order
// that every block of a launch repeats: all blocks, and all their threads,
// hold the same value.
__device__ double fold_all(const double* __restrict__ part, int np) {
    __shared__ double lds[CB / 64 + 1];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += CB) v += part[i];
    auto add = [](double a, double b) { return a + b; };
    v = block_reduce<double, decltype(add), CB>(v, add, lds);
    if (threadIdx.x == 0) lds[CB / 64] = v;
    __syncthreads();
    v = lds[CB / 64];
    __syncthreads();
    return v;
}

// CG iteration k, first half: rz_k = r.z folded from the partials prz
// (flexible CG: the Polak-Ribiere numerator z.(r - r_old), the second half
// of prz), beta = rz_k / rz_{k-1}; the direction d = z + beta d_prev (k = 0:
// d = z) of the cell and, redundantly, of its four neighbours; q = S d and
// the p.q partial.  rz_k goes to slot CG_RZ (k even) or CG_RZ2 (k odd), and
// rz_{k-1} is read from the other one, so no block reads a slot that another
// block of the launch writes.
template <typename T>
__global__ __launch_bounds__(CB) void k_spmv_dir(int nx, int ny, const T* __restrict__ diag,
                                                 const T* __restrict__ cE,
                                                 const T* __restrict__ cN,
                                                 const T* __restrict__ z,
                                                 const T* __restrict__ dprev, T* __restrict__ d,
                                                 T* __restrict__ q,
                                                 const double* __restrict__ prz, int np,
                                                 int flex, int k, Red R,
                                                 const int* __restrict__ done) {
    if (is_done(done)) return;
    const double rz = fold_all(prz, np);
    double num = rz;
    if (flex && k > 0) num -= fold_all(prz + np, np);
    const T beta = k == 0 ? T(0) : (T)(num / R.sc[(k & 1) ? CG_RZ : CG_RZ2]);
    if (blockIdx.x == 0 && threadIdx.x == 0) R.sc[(k & 1) ? CG_RZ2 : CG_RZ] = rz;
    const int n = nx * ny;
    auto dir = [&](int e) { return k == 0 ? z[e] : z[e] + beta * dprev[e]; };
    double acc = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB) {
        const int i = c / ny, j = c - i * ny;
        const T dc = dir(c);
        T s = T(0);
        if (i + 1 < nx) s += cE[c] * dir(c + ny);
        if (i > 0) s += cE[c - ny] * dir(c - ny);
        if (j + 1 < ny) s += cN[c] * dir(c + 1);
        if (j > 0) s += cN[c - 1] * dir(c - 1);
        const T v = diag[c] * dc + s;
        d[c] = dc;
        q[c] = v;
        acc += (double)dc * (double)v;
    }
    reduce_finish(R, acc, 0.0, false);
}

// CG iteration k, second half: alpha = rz_k / p.q (p.q folded from the
// partials ppq; p.q <= 0 is a breakdown: the flag becomes 2 and nothing is
// updated); x += alpha d, r -= alpha q (r_old keeps r before the update when
// given, r32 receives r rounded to FP32 when given) and the r.r partial.
template <typename T>
__global__ __launch_bounds__(CB) void k_axpy2_fold(int n, T* __restrict__ x, T* __restrict__ r,
                                                   T* __restrict__ r_old, float* __restrict__ r32,
                                                   const T* __restrict__ d,
                                                   const T* __restrict__ q,
                                                   const double* __restrict__ ppq, int np, int k,
                                                   Red R, const int* __restrict__ done) {
    if (is_done(done)) return;
    const double pq = fold_all(ppq, np);
    if (!(pq > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) R.fl[0] = 2;  // not positive definite
        return;
    }
    const T alpha = (T)(R.sc[(k & 1) ? CG_RZ2 : CG_RZ] / pq);
    double acc = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB) {
        x[c] += alpha * d[c];
        T rv = r[c];
        if (r_old) r_old[c] = rv;
        rv -= alpha * q[c];
        r[c] = rv;
        if (r32) r32[c] = (float)rv;
        acc += (double)rv * (double)rv;
    }
    reduce_finish(R, acc, 0.0, false);
}

// z = the FP32 preconditioner output widened to FP64, with the partials of
// z.r and (when r_old is given) z.r_old.
__global__ __launch_bounds__(CB) void k_widen_dot(int n, const float* __restrict__ z32,
                                                  double* __restrict__ z,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ r_old, Red R,
                                                  const int* __restrict__ done) {
    if (is_done(done)) return;
    double s1 = 0.0, s2 = 0.0;
    for (int c = blockIdx.x * CB + threadIdx.x; c < n; c += gridDim.x * CB) {
        const double zv = (double)z32[c];
        z[c] = zv;
        s1 += zv * r[c];
        if (r_old) s2 += zv * r_old[c];
    }
    reduce_finish(R, s1, s2, r_old != nullptr);
}

// ───────────────────────── Chorin stage kernels ────────────────────────────
// Qs = Q + dt dQ (rows b, h, hu, hv) and the upwind advection of hw with the
// central face mass flux of Q.
// (dQ in the FV solver's cell order: cell c at inv[c], or c when inv is NULL;
// FP64 or, for the mixed-precision predictor, FP32)
template <typename TQ>
__global__ __launch_bounds__(CB) void k_predict(int nx, int ny, double dx, double dy, double dt,
                                                double eps, const double* __restrict__ Q,
                                                const TQ* __restrict__ dQ,
                                                const int* __restrict__ inv,
                                                double* __restrict__ Qs) {
    const int n = nx * ny;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= n) return;
    int i = c / ny, j = c - i * ny;
    const double *h = Q + n, *hu = Q + 2 * n, *hv = Q + 3 * n, *hw = Q + 4 * n;
    const int cf = inv ? inv[c] : c;
    Qs[c] = Q[c] + dt * (double)dQ[cf];
    Qs[n + c] = h[c] + dt * (double)dQ[n + cf];
    Qs[2 * n + c] = hu[c] + dt * (double)dQ[2 * n + cf];
    Qs[3 * n + c] = hv[c] + dt * (double)dQ[3 * n + cf];
    const double wc = hw[c] * hinv_of(h[c], eps);
    double acc = 0.0;
    if (i + 1 < nx) {
        int o = c + ny;
        double m = 0.5 * (hu[c] + hu[o]) * dy;
        acc += m * (m > 0.0 ? wc : hw[o] * hinv_of(h[o], eps));
    }
    if (i > 0) {
        int o = c - ny;
        double m = 0.5 * (hu[o] + hu[c]) * dy;
        acc -= m * (m > 0.0 ? hw[o] * hinv_of(h[o], eps) : wc);
    }
    if (j + 1 < ny) {
        int o = c + 1;
        double m = 0.5 * (hv[c] + hv[o]) * dx;
        acc += m * (m > 0.0 ? wc : hw[o] * hinv_of(h[o], eps));
    }
    if (j > 0) {
        int o = c - 1;
        double m = 0.5 * (hv[o] + hv[c]) * dx;
        acc -= m * (m > 0.0 ? hw[o] * hinv_of(h[o], eps) : wc);
    }
    Qs[4 * n + c] = hw[c] - dt / (dx * dy) * acc;
}

// rhs = -D(Qs)/dt, D_c = sum_f |f| v_f a_c + 2 |A_c| w_c, v_f the average of
// the two cells' outward normal velocities
template <typename T>
__global__ __launch_bounds__(CB) void k_rhs(int nx, int ny, double dx, double dy, double dt,
                                            double eps, const double* __restrict__ Qs,
                                            T* __restrict__ rhs) {
    const int n = nx * ny;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= n) return;
    int i = c / ny, j = c - i * ny;
    const double *b = Qs, *h = Qs + n, *hu = Qs + 2 * n, *hv = Qs + 3 * n, *hw = Qs + 4 * n;
    const double hc = h[c], bc = b[c], hic = hinv_of(hc, eps);
    const double uc = hu[c] * hic, vc = hv[c] * hic;
    double D = 2.0 * dx * dy * hw[c] * hic;
    if (i + 1 < nx) {
        int o = c + ny;
        D += dy * 0.5 * (uc + hu[o] * hinv_of(h[o], eps)) * (hc + bc - b[o]);
    }
    if (i > 0) {
        int o = c - ny;
        D -= dy * 0.5 * (hu[o] * hinv_of(h[o], eps) + uc) * (hc + bc - b[o]);
    }
    if (j + 1 < ny) {
        int o = c + 1;
        D += dx * 0.5 * (vc + hv[o] * hinv_of(h[o], eps)) * (hc + bc - b[o]);
    }
    if (j > 0) {
        int o = c - 1;
        D -= dx * 0.5 * (hv[o] * hinv_of(h[o], eps) + vc) * (hc + bc - b[o]);
    }
    rhs[c] = (T)(-D / dt);
}

// hu -= dt * (average of the two x-face B p), hv likewise, hw += 2 dt p;
// with Q0 the result is averaged with it: Qout = (Q0 + corrected) / 2.
__global__ __launch_bounds__(CB) void k_correct(int nx, int ny, double dx, double dy, double dt,
                                                const double* __restrict__ Qs,
                                                const double* __restrict__ p,
                                                const double* __restrict__ Q0,
                                                double* __restrict__ Qout) {
    const int n = nx * ny;
    int c = blockIdx.x * CB + threadIdx.x;
    if (c >= n) return;
    int i = c / ny, j = c - i * ny;
    const double *b = Qs, *h = Qs + n;
    const double hc = h[c], bc = b[c], pc = p[c];
    double fx = 0.0, fy = 0.0;
    if (i + 1 < nx) {  // face (c, c+ny) oriented +x
        int o = c + ny;
        fx += ((h[o] + b[o] - bc) * p[o] - (hc + bc - b[o]) * pc) / dx;
    }
    if (i > 0) {
        int o = c - ny;
        fx += ((hc + bc - b[o]) * pc - (h[o] + b[o] - bc) * p[o]) / dx;
    }
    if (j + 1 < ny) {
        int o = c + 1;
        fy += ((h[o] + b[o] - bc) * p[o] - (hc + bc - b[o]) * pc) / dy;
    }
    if (j > 0) {
        int o = c - 1;
        fy += ((hc + bc - b[o]) * pc - (h[o] + b[o] - bc) * p[o]) / dy;
    }
    double q0 = bc, q1 = hc;
    double q2 = Qs[2 * n + c] - dt * 0.5 * fx;
    double q3 = Qs[3 * n + c] - dt * 0.5 * fy;
    double q4 = Qs[4 * n + c] + 2.0 * dt * pc;
    if (Q0) {
        q0 = 0.5 * (Q0[c] + q0);
        q1 = 0.5 * (Q0[n + c] + q1);
        q2 = 0.5 * (Q0[2 * n + c] + q2);
        q3 = 0.5 * (Q0[3 * n + c] + q3);
        q4 = 0.5 * (Q0[4 * n + c] + q4);
    }
    Qout[c] = q0;
    Qout[n + c] = q1;
    Qout[2 * n + c] = q2;
    Qout[3 * n + c] = q3;
    Qout[4 * n + c] = q4;
}

// Qm[r][k] = Q[r][ord[k]] (ord NULL: identity) for the four FV rows, in
// the FV solver's precision
template <typename TM>
__global__ __launch_bounds__(CB) void k_gather4(int n, const int* __restrict__ ord,
                                                const double* __restrict__ Q,
                                                TM* __restrict__ Qm) {
    int k = blockIdx.x * CB + threadIdx.x;
    if (k >= n) return;
    const int c = ord ? ord[k] : k;
#pragma unroll
    for (int r = 0; r < 4; ++r) Qm[(size_t)r * n + k] = (TM)Q[(size_t)r * n + c];
}

template <typename T>
__global__ __launch_bounds__(CB) void k_fill(int n, T v, T* __restrict__ o) {
    int c = blockIdx.x * CB + threadIdx.x;
    if (c < n) o[c] = v;
}

template <typename T>
__global__ __launch_bounds__(CB) void k_cast(int n, const double* __restrict__ a, T* __restrict__ o) {
    int c = blockIdx.x * CB + threadIdx.x;
    if (c < n) o[c] = (T)a[c];
}

// FP32 preconditioner output back to the FP64 Krylov vectors
__global__ __launch_bounds__(CB) void k_widen(int n, const float* __restrict__ a,
                                              double* __restrict__ o) {
    int c = blockIdx.x * CB + threadIdx.x;
    if (c < n) o[c] = (double)a[c];
}

inline int nblk(long n) { return (int)((n + CB - 1) / CB); }
inline int nred(long n) { return (int)std::min<long>(RED_MAX, (n + CB - 1) / CB); }

}  // namespace

// ═════════════════════════════ host side ═══════════════════════════════════
#define HIPCHK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "zoomy_hip: %s failed: %s\n", #x,         \
                    hipGetErrorString(e_));                           \
            return ZJ_EHIP;                                           \
        }                                                             \
    } while (0)

namespace zj {

int Hier::alloc(int nx0, int ny0, double dx0, double dy0, int min_coarse, size_t esz_) {
    esz = esz_;
    levels.clear();
    int nx = nx0, ny = ny0;
    double dx = dx0, dy = dy0;
    while (true) {
        HLevel L;
        L.nx = nx; L.ny = ny; L.dx = dx; L.dy = dy;
        levels.push_back(L);
        if ((nx & 1) || (ny & 1) || std::min(nx, ny) / 2 < min_coarse) break;
        nx /= 2; ny /= 2; dx *= 2; dy *= 2;
    }
    size_t tot = 0;
    for (size_t l = 0; l < levels.size(); ++l) {
        size_t n = (size_t)levels[l].nx * levels[l].ny;
        tot += 6 * n * esz + (l ? 2 * n * sizeof(double) : 0) + 3 * 256;
    }
    if (hipMalloc(&mem, tot) != hipSuccess) return ZJ_ENOMEM;
    char* p = (char*)mem;
    auto take = [&](size_t bytes) {
        void* r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    for (size_t l = 0; l < levels.size(); ++l) {
        HLevel& L = levels[l];
        size_t n = (size_t)L.nx * L.ny;
        L.diag = take(n * esz); L.cE = take(n * esz); L.cN = take(n * esz);
        L.x = take(n * esz); L.f = take(n * esz); L.r = take(n * esz);
        if (l) { L.h = (double*)take(n * 8); L.b = (double*)take(n * 8); }
    }
    return ZJ_OK;
}

void Hier::release() {
    if (mem) (void)hipFree(mem);
    mem = nullptr;
}

// The FP64 hierarchy H from (h0, b0), and with H32 the FP32 coefficients of
// the same levels: one k_level launch per level.
static void build_levels(Hier& H, Hier* H32, const double* h0, const double* b0, double eps,
                         hipStream_t st) {
    for (size_t l = 0; l < H.levels.size(); ++l) {
        HLevel& L = H.levels[l];
        const int n = L.nx * L.ny;
        const double* hf = l == 0 ? h0 : (l == 1 ? h0 : H.levels[l - 1].h);
        const double* bf = l == 0 ? b0 : (l == 1 ? b0 : H.levels[l - 1].b);
        HLevel* S = H32 ? &H32->levels[l] : nullptr;
        k_level<<<nblk(n), CB, 0, st>>>(L.nx, L.ny, L.dx, L.dy, eps, hf, bf, l ? 1 : 0, L.h, L.b,
                                         (double*)L.diag, (double*)L.cE, (double*)L.cN,
                                         S ? (float*)S->diag : nullptr,
                                         S ? (float*)S->cE : nullptr,
                                         S ? (float*)S->cN : nullptr);
    }
}

// z = V(r) from level l0 down; r and z are the vectors of level l0 (on
// level 0 they are caller vectors, on deeper levels L.f and L.x).
static int cells_of(const Hier& H, int l0) {
    int c = 0;
    for (size_t l = l0; l < H.levels.size(); ++l) c += H.levels[l].nx * H.levels[l].ny;
    return c;
}

// Fused V-cycle (nu = 1): one down kernel per level above the tail, the
// tail levels in one workgroup, one up kernel per level.  Returns false
// when the tail does not fit (then the unfused V-cycle runs).
template <typename T>
static bool vcycle_fused(Hier& H, const T* r0, T* z0, int coarse_sweeps, const int* done,
                         hipStream_t st, int l0) {
    const int nl = (int)H.levels.size();
    int lt = nl, cells = 0;
    for (int l = nl - 1; l >= l0; --l) {
        cells += H.levels[l].nx * H.levels[l].ny;
        if (cells > TAIL_CELLS || nl - l > TAIL_MAXL) break;
        lt = l;
    }
    if (lt == nl) return false;
    auto X = [&](int l) { return l == l0 ? z0 : (T*)H.levels[l].x; };
    auto F = [&](int l) { return l == l0 ? r0 : (const T*)H.levels[l].f; };
    // the pre-smoothed iterate goes to the level's r array: the up kernel
    // reads it in halos that other blocks of the same launch write in X
    auto S = [&](int l) { return (T*)H.levels[l].r; };
    for (int l = l0; l < lt; ++l) {
        HLevel& L = H.levels[l];
        HLevel& C = H.levels[l + 1];
        dim3 g((C.nx + TC - 1) / TC, (C.ny + TC - 1) / TC);
        k_mg_down<T><<<g, CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, F(l), S(l),
                                       (T*)C.f, done);
    }
    TailDesc<T> d{};
    d.nl = nl - lt;
    for (int k = 0; k < d.nl; ++k) {
        HLevel& L = H.levels[lt + k];
        d.nx[k] = L.nx;
        d.ny[k] = L.ny;
        d.diag[k] = (const T*)L.diag;
        d.cE[k] = (const T*)L.cE;
        d.cN[k] = (const T*)L.cN;
    }
    d.f = F(lt);
    d.x = X(lt);
    const size_t tail_lds =
        (size_t)(5 * cells_of(H, lt) + H.levels[lt].nx * H.levels[lt].ny) * sizeof(T);
    k_mg_tail<T><<<1, TB, tail_lds, st>>>(d, coarse_sweeps, done);
    for (int l = lt - 1; l >= l0; --l) {
        HLevel& L = H.levels[l];
        dim3 g((L.nx / 2 + TC - 1) / TC, (L.ny / 2 + TC - 1) / TC);
        k_mg_up<T><<<g, CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, F(l), S(l),
                                     X(l), X(l + 1), done);
    }
    return true;
}

template <typename T>
static void vcycle(Hier& H, const T* r0, T* z0, int nu, int coarse_sweeps, const int* done,
                   hipStream_t st, int l0 = 0) {
    if (nu == 1 && H.fused && vcycle_fused<T>(H, r0, z0, coarse_sweeps, done, st, l0)) return;
    const int nl = (int)H.levels.size();
    auto X = [&](int l) { return l == l0 ? z0 : (T*)H.levels[l].x; };
    auto F = [&](int l) { return l == l0 ? r0 : (const T*)H.levels[l].f; };
    for (int l = l0; l < nl - 1; ++l) {
        HLevel& L = H.levels[l];
        T* x = X(l);
        const T* f = F(l);
        const int half = L.nx * ((L.ny + 1) >> 1);
        for (int s = 0; s < nu; ++s) {
            k_rbgs<T><<<nblk(half), CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, x,
                                                 f, 0, s == 0, done);
            k_rbgs<T><<<nblk(half), CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, x,
                                                 f, 1, 0, done);
        }
        const int n = L.nx * L.ny;
        k_resid<T><<<nblk(n), CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, x, f,
                                           (T*)L.r, done);
        HLevel& C = H.levels[l + 1];
        k_restrict<T><<<nblk(C.nx * C.ny), CB, 0, st>>>(C.nx, C.ny, (T*)L.r, (T*)C.f, done);
    }
    {
        HLevel& C = H.levels[nl - 1];
        k_coarse<T><<<1, CB, 0, st>>>(C.nx, C.ny, (T*)C.diag, (T*)C.cE, (T*)C.cN, X(nl - 1),
                                      F(nl - 1), coarse_sweeps, done);
    }
    for (int l = nl - 2; l >= l0; --l) {
        HLevel& L = H.levels[l];
        HLevel& C = H.levels[l + 1];
        T* x = X(l);
        const T* f = F(l);
        const int n = L.nx * L.ny;
        k_prolong<T><<<nblk(n), CB, 0, st>>>(L.nx, L.ny, (T*)C.x, x, done);
        const int half = L.nx * ((L.ny + 1) >> 1);
        for (int s = 0; s < nu; ++s) {
            k_rbgs<T><<<nblk(half), CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, x,
                                                 f, 1, 0, done);
            k_rbgs<T><<<nblk(half), CB, 0, st>>>(L.nx, L.ny, (T*)L.diag, (T*)L.cE, (T*)L.cN, x,
                                                 f, 0, 0, done);
        }
    }
}

// Learned V-cycle on level 0 of the FP32 hierarchy:
//   x = V r                  (the plain V-cycle, FP32)
//   z = x + N(r - S x)       (one network correction)
// train != 0 keeps the network activations (slot 0) for
// learned_vcycle_backward.
static int learned_vcycle(Hier& H, zj_mpnn* net, const float* r, float* z, int nu,
                          int coarse_sweeps, const int* done, int train, hipStream_t st) {
    HLevel& L = H.levels[0];
    const float *D = (float*)L.diag, *E = (float*)L.cE, *Nn = (float*)L.cN;
    float* t = (float*)L.r;
    const int n = L.nx * L.ny;
    vcycle<float>(H, r, z, nu, coarse_sweeps, done, st);
    k_resid<float><<<nblk(n), CB, 0, st>>>(L.nx, L.ny, D, E, Nn, z, r, t, done);
    return mpnn_forward_f32(net, L.nx, L.ny, t, D, E, Nn, z, 1, train ? 0 : 2, done, st);
}

// Parameter gradients of sum(g * z) for the last training learned_vcycle:
// the network input r - S V r does not depend on the parameters, so this is
// the network's own backward with dz = g.
static int learned_vcycle_backward(Hier& H, zj_mpnn* net, const float* g, int nu,
                                   int coarse_sweeps, hipStream_t st) {
    (void)nu;
    (void)coarse_sweeps;
    HLevel& L = H.levels[0];
    return mpnn_backward_f32(net, L.nx, L.ny, g, (float*)L.diag, nullptr, 0, st);
}

}  // namespace zj

struct zj_chorin {
    zj_solver* fv = nullptr;
    int nx = 0, ny = 0, N = 0;
    double dx = 0, dy = 0, eps = 1e-8;
    zj_chorin_params prm{};
    zj::Hier H64, H32;
    // level-0 FP64 vectors
    double *Qs = nullptr, *Q1 = nullptr, *dQ = nullptr, *rhs = nullptr, *p = nullptr;
    // FV solver cell order (ord[k] = grid cell of FV cell k; inv its inverse)
    int *ord = nullptr, *inv = nullptr;
    double* Qm = nullptr;
    // mixed-precision predictor: FP32 FV solver of the same cells
    zj_solver* fv32 = nullptr;
    float *Qm32 = nullptr, *dQ32 = nullptr;
    double *r = nullptr, *z = nullptr, *d = nullptr, *d2 = nullptr, *q = nullptr;
    // the FP32 input and output of the low-precision preconditioner and the
    // previous residual of flexible CG
    float *r32 = nullptr, *z32 = nullptr;
    double* r_old = nullptr;
    double* part = nullptr;  // reduction partials (4 * RED_MAX)
    double* sc = nullptr;    // CG scalars
    int* fl = nullptr;       // {done, iterations}
    int* host_fl = nullptr;  // pinned copy of fl (inside host_sc)
    double* host_sc = nullptr;
    zj_mpnn* net = nullptr;  // learned preconditioner (may be NULL)
    // iterations of the last solve per preconditioner (ZJ_PRECOND_*): the
    // size of the first chunk of iterations enqueued before the host reads
    // the convergence flag
    int last_iters[4] = {8, 8, 8, 8};
    zj_chorin_stats stats{};
    std::vector<void*> allocs;
};

namespace {

template <typename T>
static int dalloc(zj_chorin* c, T** p, size_t n) {
    void* v = nullptr;
    if (hipMalloc(&v, n * sizeof(T)) != hipSuccess) return ZJ_ENOMEM;
    c->allocs.push_back(v);
    *p = (T*)v;
    return ZJ_OK;
}

static int read_flags(zj_chorin* c, hipStream_t st) {
    // the scalars and the flags are one block: one copy
    HIPCHK(hipMemcpyAsync(c->host_sc, c->sc, (CG_NSLOT + 2) * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return ZJ_OK;
}

// FP64 preconditioned CG on level 0 of the FP64 hierarchy.  pmode:
//   ZJ_PRECOND_MG       the FP64 V-cycle, CG
//   ZJ_PRECOND_MIXED    the FP32 V-cycle of the FP32 hierarchy, flexible CG
//   ZJ_PRECOND_LEARNED  the FP32 V-cycle plus the BF16 network correction,
//                       flexible CG (Polak-Ribiere beta)
// The Krylov vectors, the operator products and the residual stay FP64; the
// low-precision preconditioners receive the residual rounded to FP32 and
// return an FP32 correction.  x holds the initial guess when x0 != 0.
// One iteration is k_spmv_dir (beta, direction, S d), k_axpy2_fold (alpha,
// x and r), k_fin (r.r and the convergence test), the preconditioner and its
// r.z partials; the scalars of beta and alpha are folded by the kernels that
// use them.  Host-synchronous once per chunk of iterations.
static int pcg(zj_chorin* c, const double* rhs, double* x, int x0, double tol, int maxit,
               int pmode, hipStream_t st, int* iters, double* relres, int slot = -1) {
    zj::Hier& H = c->H64;
    zj::HLevel& L = H.levels[0];
    const int n = L.nx * L.ny;
    const int nr = nred(n);
    const double *D = (double*)L.diag, *E = (double*)L.cE, *Nn = (double*)L.cN;
    double *r = c->r, *z = c->z, *q = c->q;
    double* dbuf[2] = {c->d, c->d2};
    const int flex = pmode != ZJ_PRECOND_MG;
    double* r_old = flex ? c->r_old : nullptr;
    const double tol2 = tol * tol;
    // partial sums: r.z (and z.r_old) | p.q | r.r
    double* prz = c->part;
    double* ppq = c->part + 2 * RED_MAX;
    double* prr = c->part + 3 * RED_MAX;
    auto red = [&](double* part, int op) { return Red{part, c->sc, c->fl, tol2, maxit, op}; };
    // z = M r and the r.z partials (flexible: with z.r_old when given)
    auto precond = [&](const double* ro) -> int {
        if (!flex) {
            zj::vcycle<double>(H, r, z, c->prm.nu, c->prm.coarse_sweeps, c->fl, st);
            k_dot<double><<<nr, CB, 0, st>>>(n, z, r, red(prz, FIN_NONE), c->fl);
            return ZJ_OK;
        }
        if (pmode == ZJ_PRECOND_LEARNED) {
            if (int rc = zj::learned_vcycle(c->H32, c->net, c->r32, c->z32, c->prm.nu,
                                            c->prm.coarse_sweeps, c->fl, 0, st))
                return rc;
        } else {
            zj::vcycle<float>(c->H32, c->r32, c->z32, c->prm.nu, c->prm.coarse_sweeps, c->fl,
                              st);
        }
        k_widen_dot<<<nr, CB, 0, st>>>(n, c->z32, z, r, ro, red(prz, FIN_NONE), c->fl);
        return ZJ_OK;
    };
    k_cg_init<double><<<nr, CB, 0, st>>>(L.nx, L.ny, D, E, Nn, x, !x0, rhs, r,
                                         red(prz, FIN_B2));
    k_fin<<<1, RED_MAX, 0, st>>>(red(prz, FIN_B2), nr);
    if (x0) {
        k_dot<double><<<nr, CB, 0, st>>>(n, r, r, red(prr, FIN_RR0), c->fl);
        k_fin<<<1, RED_MAX, 0, st>>>(red(prr, FIN_RR0), nr);
    }
    if (flex) k_cast<float><<<nblk(n), CB, 0, st>>>(n, r, c->r32);
    if (int rc = precond(nullptr)) return rc;
    HIPCHK(hipGetLastError());
    int launched = 0;
    int& last = c->last_iters[slot < 0 ? pmode : slot];
    int chunk = std::max(1, std::min(maxit, last));
    while (true) {
        for (int k = launched; k < launched + chunk; ++k) {
            k_spmv_dir<double><<<nr, CB, 0, st>>>(L.nx, L.ny, D, E, Nn, z, dbuf[(k + 1) & 1],
                                                  dbuf[k & 1], q, prz, nr, flex, k,
                                                  red(ppq, FIN_NONE), c->fl);
            k_axpy2_fold<double><<<nr, CB, 0, st>>>(n, x, r, r_old, flex ? c->r32 : nullptr,
                                                    dbuf[k & 1], q, ppq, nr, k,
                                                    red(prr, FIN_RR), c->fl);
            k_fin<<<1, RED_MAX, 0, st>>>(red(prr, FIN_RR), nr);
            if (int rc = precond(r_old)) return rc;
        }
        HIPCHK(hipGetLastError());
        launched += chunk;
        if (int rc = read_flags(c, st)) return rc;
        if (c->host_fl[0] || launched >= maxit) break;
        chunk = std::min(2, maxit - launched);
    }
    *iters = c->host_fl[1];
    last = std::max(1, *iters);
    double b2 = c->host_sc[CG_B2];
    *relres = b2 > 0 ? std::sqrt(c->host_sc[CG_RR] / b2) : 0.0;
    if (c->host_fl[0] == 2) return ZJ_ESOLVE;
    return ZJ_OK;
}

// FP64 residual r = rhs - S x and its relative norm (host-synchronous).
static int residual64(zj_chorin* c, const double* rhs, const double* x, double* r, double* rel,
                      hipStream_t st) {
    zj::HLevel& L = c->H64.levels[0];
    const int nr = nred(c->N);
    const Red rb{c->part, c->sc, c->fl, 0.0, 1, FIN_B2};
    const Red rr{c->part, c->sc, c->fl, 0.0, 1, FIN_RR0};
    k_cg_init<double><<<nr, CB, 0, st>>>(L.nx, L.ny, (double*)L.diag, (double*)L.cE,
                                         (double*)L.cN, (double*)x, 0, rhs, r, rb);
    k_fin<<<1, RED_MAX, 0, st>>>(rb, nr);
    k_dot<double><<<nr, CB, 0, st>>>(c->N, r, r, rr, nullptr);
    k_fin<<<1, RED_MAX, 0, st>>>(rr, nr);
    HIPCHK(hipGetLastError());
    if (int rc = read_flags(c, st)) return rc;
    double b2 = c->host_sc[CG_B2];
    *rel = b2 > 0 ? std::sqrt(c->host_sc[CG_RR] / b2) : 0.0;
    return ZJ_OK;
}

// Solve S p = rhs on the current hierarchy to the relative residual tol.
// With a low-precision preconditioner (mixed or learned) the FP64 residual
// of the result is recomputed from scratch (rhs - S p, not the recurrence)
// and the solve is redone with the plain FP64 MG-CG when it misses tol.
static int solve(zj_chorin* c, const double* rhs, double* p, int precond, hipStream_t st) {
    zj_chorin_stats& S = c->stats;
    const double tol = c->prm.tol;
    int its = 0;
    double rel = 0.0;
    S.solves += 1;
    if (precond == ZJ_PRECOND_LEARNED_GUESS && c->net) {
        // the learned model as the initial guess of the FP64 MG-CG, with a
        // retry from zero when the solve misses tol (the reference's learned
        // guess, gnn_blueprint/imex_child_solver.py:253-283): x0 = the
        // learned V-cycle of the right-hand side, one network application
        const int n = c->N;
        k_cast<float><<<nblk(n), CB, 0, st>>>(n, rhs, c->r32);
        if (int rc = zj::learned_vcycle(c->H32, c->net, c->r32, c->z32, c->prm.nu,
                                        c->prm.coarse_sweeps, nullptr, 0, st))
            return rc;
        k_widen<<<nblk(n), CB, 0, st>>>(n, c->z32, p);
        int rc = pcg(c, rhs, p, 1, tol, c->prm.maxiter, ZJ_PRECOND_MG, st, &its, &rel,
                     ZJ_PRECOND_LEARNED_GUESS);
        if (rc && rc != ZJ_ESOLVE) return rc;
        S.lp_iterations += its;
        S.lp_solves += 1;
        if (!(rel <= S.lp_relres)) S.lp_relres = rel;
        if (!rc && rel <= tol) {
            S.iterations += its;
            S.relres = std::max(S.relres, rel);
            return ZJ_OK;
        }
        S.fallbacks += 1;  // retry from zero
    }
    if (precond == ZJ_PRECOND_MIXED || (precond == ZJ_PRECOND_LEARNED && c->net)) {
        int rc = pcg(c, rhs, p, 0, tol, c->prm.maxiter, precond, st, &its, &rel);
        if (rc && rc != ZJ_ESOLVE) return rc;
        double true_rel = 0.0;
        if ((rc = residual64(c, rhs, p, c->r_old, &true_rel, st))) return rc;
        S.lp_iterations += its;
        S.lp_solves += 1;
        if (!(true_rel <= S.lp_relres)) S.lp_relres = true_rel;
        if (true_rel <= tol) {
            S.iterations += its;
            S.relres = std::max(S.relres, true_rel);
            return ZJ_OK;
        }
        S.fallbacks += 1;  // redo plainly
    }
    int rc = pcg(c, rhs, p, 0, tol, c->prm.maxiter, ZJ_PRECOND_MG, st, &its, &rel);
    if (rc && rc != ZJ_ESOLVE) return rc;
    S.iterations += its;
    S.relres = std::max(S.relres, rel);
    if (rc == ZJ_ESOLVE || !(rel <= std::max(1e-6, 10.0 * tol))) return ZJ_ESOLVE;
    return ZJ_OK;
}

static int setup_operator(zj_chorin* c, const double* b, const double* h, hipStream_t st) {
    zj::mpnn_invalidate_phi(c->net);
    zj::build_levels(c->H64, c->prm.precond != ZJ_PRECOND_MG ? &c->H32 : nullptr, h, b, c->eps,
                     st);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

static int predict(zj_chorin* c, double dt, const double* Q, double* Qs, hipStream_t st) {
    const int n = c->N;
    if (c->fv32) {  // mixed precision: FP32 fluxes, FP64 state
        k_gather4<float><<<nblk(n), CB, 0, st>>>(n, c->ord, Q, c->Qm32);
        if (int rc = zj_flux_operator(c->fv32, c->Qm32, c->dQ32, st)) return rc;
        k_predict<float><<<nblk(n), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, c->eps, Q,
                                                 c->dQ32, c->inv, Qs);
        HIPCHK(hipGetLastError());
        return ZJ_OK;
    }
    const double* Qf = Q;
    if (c->ord) {
        k_gather4<double><<<nblk(n), CB, 0, st>>>(n, c->ord, Q, c->Qm);
        Qf = c->Qm;
    }
    if (int rc = zj_flux_operator(c->fv, Qf, c->dQ, st)) return rc;
    k_predict<double><<<nblk(n), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, c->eps, Q, c->dQ,
                                              c->inv, Qs);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

static int stage(zj_chorin* c, double dt, const double* Qin, double* Qout, const double* Q0,
                 hipStream_t st) {
    const int n = c->N;
    if (int rc = predict(c, dt, Qin, c->Qs, st)) return rc;
    if (int rc = setup_operator(c, c->Qs, c->Qs + n, st)) return rc;
    k_rhs<double><<<nblk(n), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, c->eps, c->Qs, c->rhs);
    HIPCHK(hipGetLastError());
    if (int rc = solve(c, c->rhs, c->p, c->prm.precond, st)) return rc;
    k_correct<<<nblk(n), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, c->Qs, c->p, Q0, Qout);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // namespace

extern "C" {

int zj_chorin_create(zj_solver* fv, int32_t nx, int32_t ny, double dx, double dy,
                     const zj_chorin_params* prm, zj_chorin** out) {
    if (!out || nx < 2 || ny < 2 || !(dx > 0) || !(dy > 0) || !prm) return ZJ_EINVAL;
    if (prm->tol <= 0 || prm->maxiter < 1 || prm->nu < 1 || prm->coarse_sweeps < 1 ||
        prm->min_coarse < 1 || (prm->time_order != 1 && prm->time_order != 2) ||
        prm->precond < ZJ_PRECOND_MG || prm->precond > ZJ_PRECOND_LEARNED_GUESS)
        return ZJ_EINVAL;
    zj_chorin* c = new (std::nothrow) zj_chorin();
    if (!c) return ZJ_ENOMEM;
    c->fv = fv;
    c->nx = nx;
    c->ny = ny;
    c->N = nx * ny;
    c->dx = dx;
    c->dy = dy;
    c->eps = prm->eps_wet > 0 ? prm->eps_wet : 1e-8;
    c->prm = *prm;
    const size_t N = c->N;
    int rc = c->H64.alloc(nx, ny, dx, dy, prm->min_coarse, sizeof(double));
    if (!rc) rc = c->H32.alloc(nx, ny, dx, dy, prm->min_coarse, sizeof(float));
    if (!rc) rc = dalloc(c, &c->Qs, 5 * N);
    if (!rc) rc = dalloc(c, &c->Q1, 5 * N);
    if (!rc) rc = dalloc(c, &c->dQ, 4 * N);
    if (!rc) rc = dalloc(c, &c->rhs, N);
    if (!rc) rc = dalloc(c, &c->p, N);
    if (!rc) rc = dalloc(c, &c->r, N);
    if (!rc) rc = dalloc(c, &c->z, N);
    if (!rc) rc = dalloc(c, &c->d, N);
    if (!rc) rc = dalloc(c, &c->d2, N);
    if (!rc) rc = dalloc(c, &c->q, N);
    if (!rc) rc = dalloc(c, &c->r32, N);
    if (!rc) rc = dalloc(c, &c->z32, N);
    if (!rc) rc = dalloc(c, &c->r_old, N);
    if (!rc) rc = dalloc(c, &c->part, 4 * RED_MAX);
    // CG scalars and the flags {done, iterations} in one block (the host
    // reads both with one copy)
    if (!rc) rc = dalloc(c, &c->sc, CG_NSLOT + 2);
    if (!rc) c->fl = (int*)(c->sc + CG_NSLOT);
    if (!rc && hipHostMalloc((void**)&c->host_sc, (CG_NSLOT + 2) * sizeof(double)) != hipSuccess)
        rc = ZJ_ENOMEM;
    if (!rc) c->host_fl = (int*)(c->host_sc + CG_NSLOT);
    if (rc) {
        zj_chorin_destroy(c);
        return rc;
    }
    *out = c;
    return ZJ_OK;
}

int zj_chorin_destroy(zj_chorin* c) {
    if (!c) return ZJ_OK;
    (void)hipDeviceSynchronize();
    for (void* p : c->allocs) (void)hipFree(p);
    c->H64.release();
    c->H32.release();
    if (c->host_sc) (void)hipHostFree(c->host_sc);
    delete c;
    return ZJ_OK;
}

int zj_chorin_set_fv_order(zj_chorin* c, const int32_t* order) {
    if (!c) return ZJ_EINVAL;
    if (!order) {
        c->ord = c->inv = nullptr;  // buffers stay allocated until destroy
        return ZJ_OK;
    }
    const int n = c->N;
    std::vector<int> inv(n, -1);
    for (int k = 0; k < n; ++k) {
        const int v = order[k];
        if (v < 0 || v >= n || inv[v] >= 0) return ZJ_EINVAL;
        inv[v] = k;
    }
    int *dord = nullptr, *dinv = nullptr;
    if (dalloc(c, &dord, n) || dalloc(c, &dinv, n)) return ZJ_ENOMEM;
    if (!c->Qm && dalloc(c, &c->Qm, 4 * (size_t)n)) return ZJ_ENOMEM;
    HIPCHK(hipMemcpy(dord, order, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dinv, inv.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    c->ord = dord;
    c->inv = dinv;
    return ZJ_OK;
}

int zj_chorin_set_fv32(zj_chorin* c, zj_solver* fv32) {
    if (!c) return ZJ_EINVAL;
    if (!fv32) {
        c->fv32 = nullptr;
        return ZJ_OK;
    }
    if (zj::solver_dtype(fv32) != ZJ_F32 || zj::solver_cells(fv32) != c->N) return ZJ_EINVAL;
    if (!c->Qm32 && (dalloc(c, &c->Qm32, 4 * (size_t)c->N) || dalloc(c, &c->dQ32, 4 * (size_t)c->N)))
        return ZJ_ENOMEM;
    c->fv32 = fv32;
    return ZJ_OK;
}

int zj_chorin_set_mg_fused(zj_chorin* c, int32_t on) {
    if (!c) return ZJ_EINVAL;
    c->H64.fused = c->H32.fused = on != 0;
    return ZJ_OK;
}

int zj_chorin_set_precond(zj_chorin* c, int32_t precond, zj_mpnn* net) {
    if (!c || precond < ZJ_PRECOND_MG || precond > ZJ_PRECOND_LEARNED_GUESS) return ZJ_EINVAL;
    if ((precond == ZJ_PRECOND_LEARNED || precond == ZJ_PRECOND_LEARNED_GUESS) && !net)
        return ZJ_EINVAL;
    c->prm.precond = precond;
    c->net = net;
    return ZJ_OK;
}

int zj_chorin_setup_operator(zj_chorin* c, const double* b, const double* h, void* stream) {
    if (!c || !b || !h) return ZJ_EINVAL;
    return setup_operator(c, b, h, (hipStream_t)stream);
}

int zj_chorin_apply(zj_chorin* c, const double* x, double* y, void* stream) {
    if (!c || !x || !y) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    zj::HLevel& L = c->H64.levels[0];
    k_spmv<double><<<nred(c->N), CB, 0, st>>>(L.nx, L.ny, (double*)L.diag, (double*)L.cE,
                                              (double*)L.cN, x, y, Red{}, nullptr);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_chorin_vcycle(zj_chorin* c, const void* r, void* z, int32_t dtype, void* stream) {
    if (!c || !r || !z) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == ZJ_F32)
        zj::vcycle<float>(c->H32, (const float*)r, (float*)z, c->prm.nu, c->prm.coarse_sweeps,
                          nullptr, st);
    else
        zj::vcycle<double>(c->H64, (const double*)r, (double*)z, c->prm.nu,
                           c->prm.coarse_sweeps, nullptr, st);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_chorin_learned_vcycle(zj_chorin* c, const float* r, float* z, int32_t train,
                             void* stream) {
    if (!c || !r || !z || !c->net || c->H32.levels.size() < 2) return ZJ_EINVAL;
    return zj::learned_vcycle(c->H32, c->net, r, z, c->prm.nu, c->prm.coarse_sweeps, nullptr,
                              train, (hipStream_t)stream);
}

int zj_chorin_learned_vcycle_backward(zj_chorin* c, const float* g, void* stream) {
    if (!c || !g || !c->net || c->H32.levels.size() < 2) return ZJ_EINVAL;
    return zj::learned_vcycle_backward(c->H32, c->net, g, c->prm.nu, c->prm.coarse_sweeps,
                                       (hipStream_t)stream);
}

int zj_chorin_levels(zj_chorin* c, int32_t* n_levels, int32_t* dims_out, int32_t max_levels) {
    if (!c || !n_levels) return ZJ_EINVAL;
    *n_levels = (int32_t)c->H64.levels.size();
    for (int l = 0; dims_out && l < *n_levels && l < max_levels; ++l) {
        dims_out[2 * l] = c->H64.levels[l].nx;
        dims_out[2 * l + 1] = c->H64.levels[l].ny;
    }
    return ZJ_OK;
}

int zj_chorin_solve(zj_chorin* c, const double* rhs, double* p, void* stream) {
    if (!c || !rhs || !p) return ZJ_EINVAL;
    c->stats = zj_chorin_stats{};
    return solve(c, rhs, p, c->prm.precond, (hipStream_t)stream);
}

int zj_chorin_rhs(zj_chorin* c, double dt, const double* Qs, double* rhs, void* stream) {
    if (!c || !Qs || !rhs || !(dt > 0)) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    k_rhs<double><<<nblk(c->N), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, c->eps, Qs, rhs);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_chorin_correct(zj_chorin* c, double dt, const double* Qs, const double* p,
                      const double* Q0, double* Qout, void* stream) {
    if (!c || !Qs || !p || !Qout || !(dt > 0)) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    k_correct<<<nblk(c->N), CB, 0, st>>>(c->nx, c->ny, c->dx, c->dy, dt, Qs, p, Q0, Qout);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_chorin_predict(zj_chorin* c, double dt, const double* Q, double* Qs, void* stream) {
    if (!c || !Q || !Qs || !(dt > 0)) return ZJ_EINVAL;
    return predict(c, dt, Q, Qs, (hipStream_t)stream);
}

int zj_chorin_stage(zj_chorin* c, double dt, const double* Q, double* Qout, const double* Q0,
                    void* stream) {
    if (!c || !Q || !Qout || !(dt > 0) || Q == Qout) return ZJ_EINVAL;
    c->stats = zj_chorin_stats{};
    return stage(c, dt, Q, Qout, Q0, (hipStream_t)stream);
}

int zj_chorin_step(zj_chorin* c, double dt, const double* Q, double* Qout, void* stream) {
    if (!c || !Q || !Qout || !(dt > 0) || Q == Qout) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    c->stats = zj_chorin_stats{};
    if (c->prm.time_order == 1) return stage(c, dt, Q, Qout, nullptr, st);
    int rc = stage(c, dt, Q, c->Q1, nullptr, st);
    if (rc) return rc;
    return stage(c, dt, c->Q1, Qout, Q, st);
}

int zj_chorin_get_stats(zj_chorin* c, zj_chorin_stats* out) {
    if (!c || !out) return ZJ_EINVAL;
    *out = c->stats;
    return ZJ_OK;
}

}  // extern "C"
