// Device-side numerics of the finite-volume step, shared by every kernel.
//
// Formulas (all per face, state rows b, h, hu, hv):
//  * hinv = sqrt(2) h / sqrt(h^4 + max(h, eps)^4)        (KP desingularised
//    1/h; zoomy_jax/transformation/jax_runtime.py:446-448)
//  * face flux in the reference's (F, Dp, Dm) split, cell update
//    dQ_A -= (F + Dm) |f| / |A|  (zoomy_jax/fvm/solver_jax.py:799-809)
//  * Rusanov: centred flux + path-conservative B = g h db n integral and
//    s_M I dissipation (zoomy_jax/fvm/nonconservative_flux.py:61-186); the
//    positive variant dissipates h by the eta jump and never touches b
//  * HLL with hydrostatic reconstruction (Audusse et al. 2004)
//  * cell-edge nonconservative integral (solver_jax.py:811-825)
//  * limiter candidates (zoomy_jax/fvm/reconstruction_jax.py:603-691)
//
// A gather kernel evaluates each interior face twice (once from each side),
// with the face normal pointing out of the evaluating cell; every formula
// here is antisymmetric under (L <-> R, n -> -n), so the two evaluations are
// the reference's A-side and B-side updates.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace zj {

enum { LIM_NONE = 0, LIM_BJ = 1, LIM_MINMOD = 2, LIM_VK = 3, LIM_VANLEER = 4 };
enum { RS_RUSANOV_NC = 0, RS_RUSANOV = 1, RS_HLL = 2 };
enum { RECON_CONS = 0, RECON_ETA = 1, RECON_XZ = 2, RECON_ETA_HU = 3 };
enum { BC_EXTRAP = 0, BC_WALL = 1, BC_PERIODIC = 2, BC_DIRICHLET = 3 };

constexpr int MAXF = 4;      // faces per cell (quad)
constexpr int BLOCK = 256;   // threads per block (4 wave64)
constexpr int SBLOCK = 256;  // stage kernel: 4 wave64 (3 blocks per CU in FP64, 4 in FP32)

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
// max / min of the limiter bounds, the face-increment extremes and the HLL
// hydrostatic states and wave speeds: one v_max / v_min (IEEE maxNum /
// minNum) where tmax / tmin take a compare and a select per word.  They
// differ from tmax / tmin only for NaN operands (dropped here) and in the
// sign of a zero result; a NaN state still reaches the slopes through the
// least-squares sums and the fluxes through the face velocities.
// (Written as the instructions themselves: through fmax / fmin the
// compiler quiets each operand it cannot prove canonical -- every value
// read from LDS -- with one more v_max first.  The operands here are never
// signalling NaNs.)
__device__ __forceinline__ double bmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double bmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float bmax(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float bmin(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// base[i] through a 32-bit unsigned byte offset, which the compiler turns
// into the SGPR-base + VGPR-offset form of a global load or store (no
// 64-bit address arithmetic per lane) when base is uniform; needs
// i * sizeof(T) < 2^32.
template <typename T>
__device__ __forceinline__ T ldo(const T* base, unsigned i) {
    return *(const T*)((const char*)base + (size_t)(i * (unsigned)sizeof(T)));
}

template <typename T>
__device__ __forceinline__ void sto(T* base, unsigned i, T v) {
    *(T*)((char*)base + (size_t)(i * (unsigned)sizeof(T))) = v;
}

// x clipped to [0, 1] (a limiter value; a NaN gives 0, and the slope it
// scales is then NaN itself)
template <typename T> __device__ __forceinline__ T clip01(T x) {
    return bmin(bmax(x, T(0)), T(1));
}

// FP64 square root, reciprocal and quotient by Newton-Raphson from the
// hardware estimates v_rsq_f64 and v_rcp_f64, for normal finite operands.
// lean_sqrt refines the Goldschmidt pair (g, h) ~ (sqrt x, 1 / (2 sqrt x))
// -- one coupled step, then two corrections of g -- which is the correctly
// rounded square root for x >= 2^-767; smaller, zero, negative and NaN
// operands take sqrt() itself (a branch that no lane normally enters), so
// lean_sqrt equals sqrt() everywhere and only drops the scaling and
// special-value steps from the common path.  lean_rcp (two corrections),
// lean_div (lean_rcp and one residual correction) and lean_rsqrt (two
// corrections) are within an ulp of the correctly rounded results.
__device__ __forceinline__ double lean_sqrt(double x) {
    if (!(x >= 0x1p-767)) return sqrt(x);
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
// The same iteration for operands known to be >= 0 (wave speeds, depths),
// branch-free: 0 (and NaN) pass through; positive operands below 2^-767,
// depths of no physical meaning, lose the last bits instead of taking the
// scaled path.
__device__ __forceinline__ double lean_sqrt_nonneg(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return x > 0.0 ? g : x;
}
__device__ __forceinline__ double lean_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double lean_div(double a, double b) {
    const double r = lean_rcp(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double lean_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return fma(y, fma(-hx * y, y, 0.5), y);
}

// Reciprocal, square root and quotient of the kernels.  FP32: the
// hardware's v_rcp_f32 and v_sqrt_f32 (1 ulp) in place of the correctly
// rounded multi-instruction sequences -- the FP32 kernels are bound by VALU
// issue, and an ulp of FP32 lies far inside what FP32 results are held to.
// FP64: the correctly rounded operations (square root through lean_sqrt);
// with LEAN (the quad and triangle instantiations with the numerics fixed
// at compile time) the reciprocal and the quotient through lean_rcp and
// lean_div, which shortens the latency chains of the FP64 stage kernel.
template <bool LEAN = false> __device__ __forceinline__ double frcp(double x) {
    if constexpr (LEAN) return lean_rcp(x);
    else return 1.0 / x;
}
template <bool LEAN = false> __device__ __forceinline__ float frcp(float x) {
    return __builtin_amdgcn_rcpf(x);
}
__device__ __forceinline__ double fsqrt(double x) { return lean_sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// square root of an operand >= 0: with LEAN branch-free (lean_sqrt_nonneg)
template <bool LEAN = false> __device__ __forceinline__ double fsqrt_nn(double x) {
    if constexpr (LEAN) return lean_sqrt_nonneg(x);
    else return lean_sqrt(x);
}
template <bool LEAN = false> __device__ __forceinline__ float fsqrt_nn(float x) {
    return __builtin_amdgcn_sqrtf(x);
}
template <bool LEAN = false> __device__ __forceinline__ double fdiv(double a, double b) {
    if constexpr (LEAN) return lean_div(a, b);
    else return a / b;
}
template <bool LEAN = false> __device__ __forceinline__ float fdiv(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

template <typename T, bool LEAN = false> __device__ __forceinline__ T hinv_of(T h, T eps) {
    T m = tmax(h, eps);
    T m2 = m * m, h2 = h * h;
    if constexpr (sizeof(T) == 4)
        return T(1.4142135623730951) * h * __builtin_amdgcn_rsqf(h2 * h2 + m2 * m2);
    else if constexpr (LEAN)
        return T(1.4142135623730951) * h * lean_rsqrt(h2 * h2 + m2 * m2);
    else
        return T(1.4142135623730951) * h / lean_sqrt(h2 * h2 + m2 * m2);
}

// Ghost state at a boundary face from the inner state q (Wall mirrors the
// normal momentum, Extrapolation and Dirichlet copy; the prescribed rows of
// a Dirichlet face are set by the caller).
template <typename T>
__device__ __forceinline__ void bc_ghost(int kind, const T* q, T nx, T ny, T* o) {
    o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
    if (kind == BC_WALL) {
        T qn = q[2] * nx + q[3] * ny;
        o[2] = q[2] - T(2) * qn * nx;
        o[3] = q[3] - T(2) * qn * ny;
    }
}

template <typename T>
__device__ __forceinline__ T lim_cand(int lim, T d, T dp, T dn, T ev2) {
    const T eps = T(1e-30);
    if (lim == LIM_NONE) return T(1);
    if (lim == LIM_VK) {
        if (d > eps) {
            T num = dp * dp + ev2 + T(2) * d * dp;
            T den = dp * dp + T(2) * d * d + d * dp + ev2;
            return clip01(num / tmax(den, eps));
        }
        if (d < -eps) {
            T num = dn * dn + ev2 + T(2) * d * dn;
            T den = dn * dn + T(2) * d * d + d * dn + ev2;
            return clip01(-num / tmin(den, -eps));
        }
        return T(1);
    }
    if (lim == LIM_VANLEER) {
        T y;
        if (d > eps) y = dp / d;
        else if (d < -eps) y = dn / d;
        else return T(1);
        if (y <= T(0)) return T(0);
        return y >= T(2) ? T(1) : y * (T(4) - y) * T(0.25);
    }
    // Barth-Jespersen; the reference's minmod is the same clipped ratio
    if (d > eps) return clip01(dp / tmax(d, eps));
    if (d < -eps) return clip01(dn / tmin(d, -eps));
    return T(1);
}

// |u.n| + sqrt(g h) with u = q * hinv (the cell aux)
template <typename T>
__device__ __forceinline__ T max_abs_eig(const T* q, T hinv, T nx, T ny, T g) {
    T un = (q[2] * nx + q[3] * ny) * hinv;
    return fabs(un) + fsqrt(g * tmax(q[1], T(0)));
}

// F + Dm for the evaluating cell (left state qL, neighbour qR, outward n).
// hiL/hiR are the cell aux 1/h used by the fluxes; the Rusanov speed s_M
// takes 1/h of each face state itself (what the reference's periodic-seam
// golden pins: there the ghost is the partner cell, the aux the inner one's).
template <typename T, bool LEAN = false>
__device__ __forceinline__ void face_flux_left(int rs, const T* qL, const T* qR,
                                               T hiL, T hiR, T nx, T ny, T g,
                                               T eps, T* out) {
    if (rs == RS_HLL) {
        T bs = bmax(qL[0], qR[0]);
        T hLs = bmax(T(0), qL[1] + qL[0] - bs);
        T hRs = bmax(T(0), qR[1] + qR[0] - bs);
        T uL = qL[2] * hiL, vL = qL[3] * hiL;
        T uR = qR[2] * hiR, vR = qR[3] * hiR;
        T unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
        T cL = fsqrt_nn<LEAN>(g * hLs), cR = fsqrt_nn<LEAN>(g * hRs);
        T sL = bmin(unL - cL, unR - cR);
        T sR = bmax(unL + cL, unR + cR);
        T pL = T(0.5) * g * hLs * hLs, pR = T(0.5) * g * hRs * hRs;
        T FL0 = hLs * unL, FL1 = hLs * uL * unL + pL * nx, FL2 = hLs * vL * unL + pL * ny;
        T FR0 = hRs * unR, FR1 = hRs * uR * unR + pR * nx, FR2 = hRs * vR * unR + pR * ny;
        T F0, F1, F2;
        if (sL >= T(0)) {
            F0 = FL0; F1 = FL1; F2 = FL2;
        } else if (sR <= T(0)) {
            F0 = FR0; F1 = FR1; F2 = FR2;
        } else {
            T inv = frcp<LEAN>(sR - sL);
            T ss = sL * sR;
            F0 = (sR * FL0 - sL * FR0 + ss * (hRs - hLs)) * inv;
            F1 = (sR * FL1 - sL * FR1 + ss * (hRs * uR - hLs * uL)) * inv;
            F2 = (sR * FL2 - sL * FR2 + ss * (hRs * vR - hLs * vL)) * inv;
        }
        T cA = T(0.5) * g * (qL[1] * qL[1] - hLs * hLs);
        out[0] = T(0);
        out[1] = F0;
        out[2] = F1 + cA * nx;
        out[3] = F2 + cA * ny;
        return;
    }
    // Rusanov (centred flux + fluctuation with s_M dissipation)
    T unL = (qL[2] * nx + qL[3] * ny) * hiL;
    T unR = (qR[2] * nx + qR[3] * ny) * hiR;
    T pL = T(0.5) * g * qL[1] * qL[1], pR = T(0.5) * g * qR[1] * qR[1];
    T s = tmax(fabs((qL[2] * nx + qL[3] * ny) * hinv_of(qL[1], eps)) + fsqrt(g * tmax(qL[1], T(0))),
               fabs((qR[2] * nx + qR[3] * ny) * hinv_of(qR[1], eps)) + fsqrt(g * tmax(qR[1], T(0))));
    T d0 = qR[0] - qL[0], d1 = qR[1] - qL[1], d2 = qR[2] - qL[2], d3 = qR[3] - qL[3];
    T Bd = g * (T(0.5) * (qL[1] + qR[1])) * d0;
    T diss0, diss1;
    if (rs == RS_RUSANOV) {
        diss0 = T(0);
        diss1 = s * (d1 + d0);
    } else {
        diss0 = s * d0;
        diss1 = s * d1;
    }
    out[0] = T(0.5) * (T(0) + T(0)) + T(0.5) * (T(0) - diss0);
    out[1] = T(0.5) * ((qL[2] * nx + qL[3] * ny) + (qR[2] * nx + qR[3] * ny)) + T(0.5) * (T(0) - diss1);
    out[2] = T(0.5) * ((qL[2] * unL + pL * nx) + (qR[2] * unR + pR * nx)) + T(0.5) * (Bd * nx - s * d2);
    out[3] = T(0.5) * ((qL[3] * unL + pL * ny) + (qR[3] * unR + pR * ny)) + T(0.5) * (Bd * ny - s * d3);
}

// The same face numerics as face_flux_left, kept apart as the reference's
// stored triple (blocks.py:185-238): conservative flux F and the two
// fluctuations Dp (B side) and Dm (A side), with F + Dm the A-side update
// and -F + Dp the B-side one.  hiL / hiR: the cell aux 1/h of A and B.
template <typename T>
__device__ __forceinline__ void face_split(int rs, const T* qL, const T* qR, T hiL, T hiR,
                                           T nx, T ny, T g, T eps, T* F, T* Dp, T* Dm) {
    if (rs == RS_HLL) {
        T bs = bmax(qL[0], qR[0]);
        T hLs = bmax(T(0), qL[1] + qL[0] - bs);
        T hRs = bmax(T(0), qR[1] + qR[0] - bs);
        T uL = qL[2] * hiL, vL = qL[3] * hiL;
        T uR = qR[2] * hiR, vR = qR[3] * hiR;
        T unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
        T cL = fsqrt(g * hLs), cR = fsqrt(g * hRs);
        T sL = bmin(unL - cL, unR - cR);
        T sR = bmax(unL + cL, unR + cR);
        T pL = T(0.5) * g * hLs * hLs, pR = T(0.5) * g * hRs * hRs;
        T FL[3] = {hLs * unL, hLs * uL * unL + pL * nx, hLs * vL * unL + pL * ny};
        T FR[3] = {hRs * unR, hRs * uR * unR + pR * nx, hRs * vR * unR + pR * ny};
        T UL[3] = {hLs, hLs * uL, hLs * vL};
        T UR[3] = {hRs, hRs * uR, hRs * vR};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (sL >= T(0)) F[k + 1] = FL[k];
            else if (sR <= T(0)) F[k + 1] = FR[k];
            else F[k + 1] = fdiv(sR * FL[k] - sL * FR[k] + sL * sR * (UR[k] - UL[k]), sR - sL);
        }
        F[0] = T(0);
        // hydrostatic pressure corrections g/2 (h^2 - h*^2) n
        T cA = T(0.5) * g * (qL[1] * qL[1] - hLs * hLs);
        T cB = T(0.5) * g * (qR[1] * qR[1] - hRs * hRs);
        Dm[0] = Dm[1] = T(0);
        Dm[2] = cA * nx;
        Dm[3] = cA * ny;
        Dp[0] = Dp[1] = T(0);
        Dp[2] = -cB * nx;
        Dp[3] = -cB * ny;
        return;
    }
    T unL = (qL[2] * nx + qL[3] * ny) * hiL;
    T unR = (qR[2] * nx + qR[3] * ny) * hiR;
    T pL = T(0.5) * g * qL[1] * qL[1], pR = T(0.5) * g * qR[1] * qR[1];
    T s = tmax(fabs((qL[2] * nx + qL[3] * ny) * hinv_of(qL[1], eps)) + fsqrt(g * tmax(qL[1], T(0))),
               fabs((qR[2] * nx + qR[3] * ny) * hinv_of(qR[1], eps)) + fsqrt(g * tmax(qR[1], T(0))));
    T d[4] = {qR[0] - qL[0], qR[1] - qL[1], qR[2] - qL[2], qR[3] - qL[3]};
    T Bd = g * (T(0.5) * (qL[1] + qR[1])) * d[0];
    F[0] = T(0);
    F[1] = T(0.5) * ((qL[2] * nx + qL[3] * ny) + (qR[2] * nx + qR[3] * ny));
    F[2] = T(0.5) * ((qL[2] * unL + pL * nx) + (qR[2] * unR + pR * nx));
    F[3] = T(0.5) * ((qL[3] * unL + pL * ny) + (qR[3] * unR + pR * ny));
    T diss[4] = {s * d[0], s * d[1], s * d[2], s * d[3]};
    if (rs == RS_RUSANOV) {
        diss[0] = T(0);
        diss[1] = s * (d[1] + d[0]);
    }
    const T B[4] = {T(0), T(0), Bd * nx, Bd * ny};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        Dm[v] = T(0.5) * (B[v] - diss[v]);
        Dp[v] = T(0.5) * (B[v] + diss[v]);
    }
}

// Cell-edge nonconservative integral g (h_c + h_e)/2 (b_e - b_c) n (q rows).
template <typename T>
__device__ __forceinline__ void cell_edge(const T* qc, const T* qe, T nx, T ny, T g,
                                          T* out) {
    T k = g * (T(0.5) * (qc[1] + qe[1])) * (qe[0] - qc[0]);
    out[2] = k * nx;
    out[3] = k * ny;
}

// Block-wide reductions over BLOCK threads (wave64 shuffles + LDS).
template <typename T, typename Op, int NT = BLOCK>
__device__ __forceinline__ T block_reduce(T v, Op op, T* lds) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < NT / 64; ++i) v = op(v, lds[i]);
    }
    return v;  // valid on thread 0
}

// Barth-Jespersen / minmod limiter of one variable over all faces with one
// division: min_f (dp / d_f) over d_f > eps equals dp / max_f d_f (and
// dn / min_f d_f over d_f < -eps), exactly, because correctly rounded
// division is monotone in the divisor; of the two ratios (both >= 0, as
// dp >= 0 >= dn) the smaller is picked by cross-multiplication,
// dn / dmin < dp / dmax  <=>  dn dmax > dp dmin  (dmax > 0 > dmin),
// which can pick the other one only where the two agree to rounding.
// Returns false when no face increment exceeds eps (no limiting).
template <typename T, bool LEAN = false>
__device__ __forceinline__ bool min_face_ratio(T dmax_pos, T dmin_neg, T dp, T dn, T& y) {
    const T eps = T(1e-30);
    const bool a1 = dmax_pos > eps, a2 = dmin_neg < -eps;
    if (!a1 && !a2) return false;
    const bool use2 = a2 && (!a1 || dn * dmax_pos > dp * dmin_neg);
    const T num = use2 ? dn : dp, den = use2 ? dmin_neg : dmax_pos;
    y = fdiv<LEAN>(num, den);
    return true;
}

template <typename T, bool LEAN = false>
__device__ __forceinline__ T bj_phi(T dmax_pos, T dmin_neg, T dp, T dn) {
    T y;
    return min_face_ratio<T, LEAN>(dmax_pos, dmin_neg, dp, dn, y) ? clip01(y) : T(1);
}

// van Leer over all faces with one division: the per-face candidate
// y (4 - y) / 4 (0 for y <= 0, 1 for y >= 2) is non-decreasing in the face
// ratio y, so its minimum over the faces is its value at the smallest ratio,
// which min_face_ratio gives.
template <typename T, bool LEAN = false>
__device__ __forceinline__ T vl_phi(T dmax_pos, T dmin_neg, T dp, T dn) {
    T y;
    if (!min_face_ratio<T, LEAN>(dmax_pos, dmin_neg, dp, dn, y)) return T(1);
    if (y <= T(0)) return T(0);
    return y >= T(2) ? T(1) : y * (T(4) - y) * T(0.25);
}

// Blocks are dispatched round-robin over the 8 XCDs of the chip; remap the
// block index so that every XCD works on one contiguous range of cells (its
// own L2 then holds the neighbours its cells gather).
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int x = b & 7, k = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace zj
