// MI355X finite-volume time step: kernels and the C ABI (include/zoomy_hip.h).
//
// The cells are cut into patches of up to 256 consecutive ids (16 x 16
// tiles of a Morton-ordered quad mesh); one workgroup stages a patch with its
// face ring and stencil ring in LDS and runs a whole SSP-RK2 stage on it
// (k_stage): LSQ gradients, limiter bounds and limiters of the own and
// face-ring cells, then reconstruction, Riemann flux and cell-edge term of
// every face of every own cell, gathered per cell so that the update is
// atomic-free and bit-reproducible (the face-scatter of the reference,
// solver_jax.py:806-809 and solver2/blocks.py:241-287, becomes a gather over
// the face-of-cell table that blocks.py:261-263 names).  Face geometry is
// recomputed from the staged vertices instead of streamed from HBM.  An
// SSP-RK2 step is two launches (stage 2 adds the RK average, the explicit
// source, the troubled flag and the per-wave dt minima) and one small
// reduction.  The post_step aux refresh is the kernel's aux-only
// instantiation; separate kernels: the dt law (k_dt), the compensated mass
// (k_mass_*) and the local implicit source (k_imex).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/zoomy_hip.h"
#include "zj_device.hpp"
#include "zj_chorin.hpp"
#include "zj_patches.hpp"

using namespace zj;

// Per-face loops are unrolled over MAXF with an early exit so that per-face
// private arrays stay in registers (a runtime trip count would put them in
// scratch memory).
#define ZJ_FOR_FACES(k) _Pragma("unroll") for (int k = 0; k < MAXF; ++k) if (k < m.mf)

namespace {

struct DevMesh {
    int dim, N, nverts, mf, mx, nbf;
    const int* cv;        // (mf, N)
    const int* nbr;       // (mf, N)
    const int* extra;     // (mx, N)
    const double2* verts; // (nverts)
    const double* cen;    // (2, N)
    const double* rin;    // (N)
    const int* bf_kind;   // (nbf)
    const int* bf_pcell;  // (nbf) periodic partner cell or -1
    const int* bf_mask;   // (nbf) Dirichlet rows (bit v)
    const double* bf_val; // (4, nbf) Dirichlet values
};

// Ghost state of boundary face bf seen from the face state q: the periodic
// partner cell, or the kind's ghost with the prescribed rows of a Dirichlet
// face set to their values (the face value is the ghost: tests/refs/
// bc_inflow of the reference holds to round-off only with that).
template <typename T>
__device__ __forceinline__ void ghost_state(const DevMesh& m, const T* __restrict__ Qin,
                                            int bf, const T* q, T nx, T ny, T* o) {
    const int kind = m.bf_kind[bf];
    if (kind == BC_PERIODIC) {
        const int pc = m.bf_pcell[bf];
        o[0] = Qin[pc]; o[1] = Qin[(size_t)m.N + pc];
        o[2] = Qin[2 * (size_t)m.N + pc]; o[3] = Qin[3 * (size_t)m.N + pc];
        return;
    }
    bc_ghost(kind, q, nx, ny, o);
    if (kind == BC_DIRICHLET) {
        const int mask = m.bf_mask[bf];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (mask >> v & 1) o[v] = T(m.bf_val[(size_t)v * m.nbf + bf]);
    }
}

struct Opts {
    double g, cfl, eps, dt_max, manning;
    int order, lim, rs, recon;
    int pos;           // eta reconstruction: 1 = Zhang-Shu deviation cap
    double front_tol;  // eta reconstruction: front pre-detector (<= 0: off)
    int stage_aux;     // 1: the second RK stage's fluxes read the step-start 1/h
};

// depth below which the conservative-momentum eta reconstruction demotes a
// cell to first order and zeroes a face momentum
constexpr double ETA_HU_EPS = 1e-3;

struct Geo {
    double cx, cy, vol;
    double fx[MAXF], fy[MAXF], nx[MAXF], ny[MAXF], len[MAXF];
};

__device__ __forceinline__ void load_geo(const DevMesh& m, int c, Geo& G) {
    // every field is written on both branches (keeps G in registers)
    const int N = m.N;
    G.cx = m.cen[c];
    G.cy = m.cen[N + c];
    double2 v[MAXF];
#pragma unroll
    for (int k = 0; k < MAXF; ++k) v[k] = k < m.mf ? m.verts[m.cv[k * N + c]] : make_double2(0.0, 0.0);
    if (m.dim == 1) {
        G.vol = v[1].x - v[0].x;
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            G.fx[k] = v[k].x;
            G.fy[k] = 0.0;
            G.nx[k] = k == 0 ? -1.0 : 1.0;
            G.ny[k] = 0.0;
            G.len[k] = 1.0;
        }
    } else {
        double area = 0.0;
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            const double2 a = v[k];
            const double2 b = (k + 1 < m.mf) ? v[k + 1 < MAXF ? k + 1 : 0] : v[0];
            const double ex = b.x - a.x, ey = b.y - a.y;
            const double L = sqrt(ex * ex + ey * ey);
            const bool on = k < m.mf;
            G.len[k] = on ? L : 0.0;
            G.nx[k] = on ? ey / L : 0.0;
            G.ny[k] = on ? -ex / L : 0.0;
            G.fx[k] = 0.5 * (a.x + b.x);
            G.fy[k] = 0.5 * (a.y + b.y);
            area += on ? a.x * b.y - b.x * a.y : 0.0;
        }
        G.vol = 0.5 * area;
    }
}

template <typename T>
__device__ __forceinline__ void load_q(const T* __restrict__ Q, int N, int c, T* q) {
#pragma unroll
    for (int r = 0; r < 4; ++r) q[r] = ldo(Q + (size_t)r * N, c);
}

// Reconstruction variables: conservative (b, h, hu, hv), (b, eta, u, v),
// or (b, eta, hu, hv).
template <typename T>
__device__ __forceinline__ void to_w(int recon, const T* q, T hi, T* w) {
    if (recon == RECON_ETA) {
        w[0] = q[0];
        w[1] = q[1] + q[0];
        w[2] = q[2] * hi;
        w[3] = q[3] * hi;
    } else if (recon == RECON_ETA_HU) {
        w[0] = q[0]; w[1] = q[1] + q[0]; w[2] = q[2]; w[3] = q[3];
    } else {
        w[0] = q[0]; w[1] = q[1]; w[2] = q[2]; w[3] = q[3];
    }
}

// Face value from a cell (state q, aux hi, limited slopes s[8]) at offset r.
template <typename T>
__device__ __forceinline__ void recon_face(int recon, const T* q, T hi, const T* s,
                                           T rx, T ry, T* out) {
    if (recon == RECON_ETA) {
        T b = q[0] + s[0] * rx + s[1] * ry;
        T eta = (q[1] + q[0]) + s[2] * rx + s[3] * ry;
        T u = q[2] * hi + s[4] * rx + s[5] * ry;
        T v = q[3] * hi + s[6] * rx + s[7] * ry;
        eta = tmax(eta, b);  // hydrostatic face clip, reconstruction_jax.py:1269-1272
        T h = eta - b;
        out[0] = b; out[1] = h; out[2] = u * h; out[3] = v * h;
    } else if (recon == RECON_ETA_HU) {
        T b = q[0] + s[0] * rx + s[1] * ry;
        T eta = (q[1] + q[0]) + s[2] * rx + s[3] * ry;
        eta = tmax(eta, b);
        T h = eta - b;
        const bool dry = h < T(ETA_HU_EPS);
        out[0] = b; out[1] = h;
        out[2] = dry ? T(0) : q[2] + s[4] * rx + s[5] * ry;
        out[3] = dry ? T(0) : q[3] + s[6] * rx + s[7] * ry;
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] = q[v] + s[2 * v] * rx + s[2 * v + 1] * ry;
    }
}

// The eta reconstruction from a staged state whose momentum rows already
// hold the velocities (u, v) = (hu, hv) / h -- the products to_w and
// recon_face form -- so that a stencil row costs no conversion.
template <typename T>
__device__ __forceinline__ void to_w_prim(const T* p, T* w) {
    w[0] = p[0];
    w[1] = p[1] + p[0];
    w[2] = p[2];
    w[3] = p[3];
}
template <typename T>
__device__ __forceinline__ void recon_face_prim(const T* p, const T* s, T rx, T ry, T* out) {
    T b = p[0] + s[0] * rx + s[1] * ry;
    T eta = (p[1] + p[0]) + s[2] * rx + s[3] * ry;
    T u = p[2] + s[4] * rx + s[5] * ry;
    T v = p[3] + s[6] * rx + s[7] * ry;
    eta = tmax(eta, b);  // hydrostatic face clip, reconstruction_jax.py:1269-1272
    T h = eta - b;
    out[0] = b; out[1] = h; out[2] = u * h; out[3] = v * h;
}

// Per (slot, cell): the outward unit normal and r_f = min(r_cell, r_nbr)
// (r_cell on boundary slots) -- what the dt law reads, precomputed once per
// mesh so that k_dt streams them instead of gathering vertices.
__global__ __launch_bounds__(BLOCK) void k_slot_dt_geo(DevMesh m, double2* __restrict__ sn,
                                                       double* __restrict__ sr) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= m.N) return;
    const int N = m.N;
    Geo G;
    load_geo(m, c, G);
    const double rc = m.rin[c];
    ZJ_FOR_FACES(k) {
        const int nb = m.nbr[k * N + c];
        sn[(size_t)k * N + c] = make_double2(G.nx[k], G.ny[k]);
        sr[(size_t)k * N + c] = nb >= 0 ? fmin(rc, m.rin[nb]) : rc;
    }
}

// dt of a state (compute_timestep): per-cell min over faces, block min to
// part[block]; k_reduce_part folds the partials (a single address taking
// one atomic per block serialises them: 16k blocks cost ~0.2 ms).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_dt(DevMesh m, Opts o, const T* __restrict__ Q,
                                              const double2* __restrict__ sn,
                                              const double* __restrict__ sr,
                                              double* __restrict__ part) {
    __shared__ double lds[BLOCK / 64];
    const int c = xcd_block() * BLOCK + threadIdx.x;
    double loc = INFINITY;
    if (c < m.N) {
        const int N = m.N;
        T q[4];
        load_q(Q, N, c, q);
        const T hn = hinv_of(q[1], T(o.eps));
        ZJ_FOR_FACES(k) {
            const double2 n = sn[(size_t)k * N + c];
            const double rf = sr[(size_t)k * N + c];
            const double a = (double)max_abs_eig(q, hn, T(n.x), T(n.y), T(o.g));
            double l = o.cfl * 2.0 * rf / ((double)m.dim * a);
            if (l != l) l = 0.0;
            loc = fmin(loc, l);
        }
    }
    double s = block_reduce(loc, [](double a, double b) { return fmin(a, b); }, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// rows x N state between the caller's cell order and a renumbered one:
// scatter == 0: out[r, k] = in[r, ord[k]]; scatter != 0: out[r, ord[k]] =
// in[r, k].
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_permute_rows(int N, int rows, const int* __restrict__ ord,
                                                        const T* __restrict__ in,
                                                        T* __restrict__ out, int scatter) {
    const long i = (long)xcd_block() * BLOCK + threadIdx.x;
    if (i >= (long)rows * N) return;
    const int r = (int)(i / N), k = (int)(i - (long)r * N);
    const size_t o = (size_t)r * N;
    if (scatter)
        out[o + ord[k]] = in[o + k];
    else
        out[o + k] = in[o + ord[k]];
}

__global__ void k_dt_init(unsigned long long* red, int* troubled) {
    if (red) *red = (unsigned long long)__double_as_longlong(INFINITY);
    if (troubled) *troubled = 0;
}

__global__ void k_dt_finalize(unsigned long long* red, double* dt, const double* t,
                              double t_end, double dt_max) {
    double d = __longlong_as_double((long long)*red);
    if (dt_max > 0.0) d = fmin(d, dt_max);
    if (t_end >= 0.0 && t) d = fmin(d, t_end - *t);
    *dt = d;
    *red = (unsigned long long)__double_as_longlong(INFINITY);
}

__global__ void k_t_advance(double* t, const double* dt) { *t += *dt; }

// dt of compute_timestep capped by a host remainder (t_rem < 0: no cap).
__global__ void k_dt_finalize_rem(unsigned long long* red, double* dt, double t_rem,
                                  double dt_max) {
    double d = __longlong_as_double((long long)*red);
    if (dt_max > 0.0) d = fmin(d, dt_max);
    if (t_rem >= 0.0) d = fmin(d, t_rem);
    *dt = d;
}

// The march's per-step head: t += the previous step's dt (advance != 0),
// then dt of this step from the reduction scalar (capped by dt_max and
// t_end - t; the scalar is reset), then the MOOD bookkeeping of the
// previous step (stats += (1, troubled cells) when it had troubled cells;
// prev != 0) and the troubled count cleared for this step's candidate.
__global__ void k_step_head(unsigned long long* red, double* dt, double* t, double t_end,
                            double dt_max, int advance, int* cnt, int* stats, int prev) {
    if (advance) *t += *dt;
    double d = __longlong_as_double((long long)*red);
    if (dt_max > 0.0) d = fmin(d, dt_max);
    if (t_end >= 0.0) d = fmin(d, t_end - *t);
    *dt = d;
    *red = (unsigned long long)__double_as_longlong(INFINITY);
    if (cnt) {
        if (prev && stats && *cnt != 0) {
            stats[0] += 1;
            stats[1] += *cnt;
        }
        *cnt = 0;
    }
}

// Folds the per-wave dt minima of a stage launch into the reduction scalar.
__global__ __launch_bounds__(1024) void k_reduce_part(const double* __restrict__ part, int n,
                                                      unsigned long long* __restrict__ dt_red,
                                                      const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ double lds[1024 / 64];
    double v = INFINITY;
    for (int i = blockIdx.x * 1024 + threadIdx.x; i < n; i += gridDim.x * 1024)
        v = fmin(v, part[i]);
    auto mn = [](double a, double b) { return fmin(a, b); };
    v = block_reduce<double, decltype(mn), 1024>(v, mn, lds);
    if (threadIdx.x == 0) atomicMin(dt_red, (unsigned long long)__double_as_longlong(v));
}

// MOOD: the redo's dt replaces the candidate's (only when a redo runs).
__global__ void k_dt_reset_if(const int* __restrict__ cnt, unsigned long long* __restrict__ dt_red) {
    if (*cnt != 0) *dt_red = (unsigned long long)__double_as_longlong(INFINITY);
}

// MOOD bookkeeping: stats[0] += steps redone, stats[1] += cells demoted.
__global__ void k_mood_finish(const int* __restrict__ cnt, int* __restrict__ stats) {
    if (*cnt != 0) {
        stats[0] += 1;
        stats[1] += *cnt;
    }
}

// Compensated (Neumaier) mass: per-thread partials, per-block (sum, comp)
// pairs folded by a fixed tree, then one more tree over the blocks --
// deterministic for a fixed grid.
__device__ __forceinline__ void neumaier(double& s, double& comp, double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) comp += (s - t) + x;
    else comp += (x - t) + s;
    s = t;
}

__device__ __forceinline__ double cell_volume(const DevMesh& m, const double2* v);

// Fixed-order pairwise fold of NT compensated (sum, compensation) pairs in
// LDS into lds[0] (deterministic: the tree does not depend on timing).
template <int NT>
__device__ __forceinline__ void fold_pairs(double2* lds) {
#pragma unroll 1
    for (int w = NT / 2; w > 0; w >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < w) {
            const double2 a = lds[threadIdx.x], b = lds[threadIdx.x + w];
            double S = a.x, C = a.y;
            neumaier(S, C, b.x);
            lds[threadIdx.x] = make_double2(S, C + b.y);
        }
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_mass_partial(DevMesh m, const T* __restrict__ Q,
                                                        double2* __restrict__ part) {
    __shared__ double2 lds[BLOCK];
    double s = 0.0, comp = 0.0;
    for (int c = blockIdx.x * BLOCK + threadIdx.x; c < m.N; c += gridDim.x * BLOCK) {
        double2 v[MAXF];
#pragma unroll
        for (int k = 0; k < MAXF; ++k)
            v[k] = k < m.mf ? m.verts[m.cv[k * m.N + c]] : make_double2(0.0, 0.0);
        neumaier(s, comp, (double)Q[m.N + c] * cell_volume(m, v));
    }
    lds[threadIdx.x] = make_double2(s, comp);
    fold_pairs<BLOCK>(lds);
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(1024) void k_mass_final(const double2* part, int n, double* out) {
    __shared__ double2 lds[1024];
    lds[threadIdx.x] = threadIdx.x < n ? part[threadIdx.x] : make_double2(0.0, 0.0);
    fold_pairs<1024>(lds);
    if (threadIdx.x == 0) *out = lds[0].x + lds[0].y;
}

// Local implicit Manning stage (solver_imex_jax.py:268-342).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_imex(DevMesh m, Opts o, const double* __restrict__ dtp,
                                                int maxiter, const T* __restrict__ Qe,
                                                T* __restrict__ Qo) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= m.N) return;
    const int N = m.N;
    T q[4];
    load_q(Qe, N, c, q);
    Qo[c] = q[0];
    Qo[N + c] = q[1];
    if (o.manning <= 0.0) {
        Qo[2 * N + c] = q[2];
        Qo[3 * N + c] = q[3];
        return;
    }
    const T dt = T(*dtp);
    const T hinv = hinv_of(q[1], T(o.eps));
    const T k = T(o.g * o.manning * o.manning) * cbrt(hinv);
    T qx = q[2], qy = q[3];
    const T ex = qx, ey = qy;
    for (int it = 0; it < maxiter; ++it) {
        T u = qx * hinv, v = qy * hinv;
        T sp = sqrt(u * u + v * v);
        T Sx = -k * u * sp, Sy = -k * v * sp;
        T jxx = 0, jxy = 0, jyy = 0;
        if (sp > T(0)) {
            jxx = -k * hinv * (sp + u * u / sp);
            jxy = -k * hinv * (u * v / sp);
            jyy = -k * hinv * (sp + v * v / sp);
        }
        T a11 = T(1) - dt * jxx, a12 = -dt * jxy, a21 = -dt * jxy, a22 = T(1) - dt * jyy;
        T rx = qx - ex - dt * Sx, ry = qy - ey - dt * Sy;
        T det = a11 * a22 - a12 * a21;
        qx += (-rx * a22 + ry * a12) / det;
        qy += (-ry * a11 + rx * a21) / det;
    }
    Qo[2 * N + c] = qx;
    Qo[3 * N + c] = qy;
}

// ─────────────────────────────────────────────────────────────────────────
// LDS-staged stage kernel: one block per patch of up to BLOCK own cells.
// Phase 1 stages the state and centre of every cell the patch touches and
// the vertices of its own and face-ring cells in LDS (coalesced where the
// own cells are contiguous ids; a mesh whose cell order cuts into thin
// strips gets patches of spatial clusters, with their own cells listed
// explicitly: expl); phase 2 computes the limited slopes of own and
// face-ring cells from LDS; phase 3 computes the fluxes of the own cells.
// No slope array round-trips through HBM and no gather chains global ->
// global remain (boundary faces read their BC kind from global memory).
// ─────────────────────────────────────────────────────────────────────────
struct DevPatch {
    int P, npatch, ML, MS, MV;
    int expl;          // own cells listed first in load_ids (else ids p P + i)
    const int4* meta;  // per patch: template, loaded cells, vertices, slope cells | own << 16
    const int* load_ids;
    const int* vert_ids;
    const uint16_t* tpl_nbr;
    const uint32_t* tpl_extra2;  // extra-stencil local ids, packed in pairs
    const uint16_t* tpl_vert;
    const double2* cen2;
};

constexpr uint16_t LNONE = 0xFFFF;

// The template rows of local cell ls that a pass reads from global memory:
// face neighbours and vertices (MAXF each) and the first batch of NB packed
// extra-stencil pairs.  Loaded ahead of the barrier before the pass, so
// that their latency overlaps the wait.
template <int NB>
struct TplRow {
    uint16_t nbr[MAXF], vert[MAXF];
    uint32_t ex[NB > 0 ? NB : 1];
    __device__ __forceinline__ void load(const uint16_t* tn, const uint16_t* tv,
                                         const uint32_t* te, int mf, int mxh, int MS, int ls) {
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            nbr[k] = k < mf ? ldo(tn, k * MS + ls) : LNONE;
            vert[k] = k < mf ? ldo(tv, k * MS + ls) : 0;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) ex[u] = u < mxh ? ldo(te, u * MS + ls) : ~0u;
    }
};


// Face centres of a cell with local vertex ids vl, relative to org (the
// frame of the staged centres), in the precision T of the stage: formed in
// FP64 from the staged FP64 vertices, then rounded once.
template <typename T>
__device__ __forceinline__ void face_centres(const DevMesh& m, const double2* Vs,
                                             const uint16_t* vl, double2 org, double2* v,
                                             T* fx, T* fy) {
#pragma unroll
    for (int k = 0; k < MAXF; ++k) v[k] = k < m.mf ? Vs[vl[k]] : make_double2(0.0, 0.0);
    if (m.dim == 1) {
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            fx[k] = T(v[k].x - org.x);
            fy[k] = T(0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            const double2 b = (k + 1 < m.mf) ? v[k + 1 < MAXF ? k + 1 : 0] : v[0];
            fx[k] = T(0.5 * (v[k].x + b.x) - org.x);
            fy[k] = T(0.5 * (v[k].y + b.y) - org.y);
        }
    }
}

// Outward unit normal and length of slot k in the precision T of the stage.
// The edge vector is formed in FP64 from the global vertices, so the two
// cells of a face (which run the edge in opposite directions) get exactly
// negated normals and the same length.
template <typename T, bool LEAN = false>
__device__ __forceinline__ void face_normal(const DevMesh& m, const double2* v, int k, T& nx,
                                            T& ny, T& len) {
    if (m.dim == 1) {
        nx = k == 0 ? T(-1) : T(1);
        ny = T(0);
        len = T(1);
        return;
    }
    const double2 a = v[k];
    const double2 b = ((k + 1 < m.mf) ? v[k + 1 < MAXF ? k + 1 : 0] : v[0]);
    const T ex = T(b.x - a.x), ey = T(b.y - a.y);
    const T l2 = ex * ex + ey * ey;
    T il;
    if constexpr (sizeof(T) == 4)
        il = rsqrtf(l2);
    else if constexpr (LEAN)
        il = lean_rsqrt(l2);
    else
        il = rsqrt(l2);
    len = l2 * il;
    nx = ey * il;
    ny = -ex * il;
}

__device__ __forceinline__ double cell_volume(const DevMesh& m, const double2* v) {
    if (m.dim == 1) return v[1].x - v[0].x;
    double area = 0.0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        if (k >= m.mf) continue;
        const double2 a = v[k];
        const double2 b = ((k + 1 < m.mf) ? v[k + 1 < MAXF ? k + 1 : 0] : v[0]);
        area += a.x * b.y - b.x * a.y;
    }
    return 0.5 * area;
}

// Xing-Zhang-Shu factor of a cell with mean depth hbar and depth slope
// (sx, sy): theta = clip(hbar / max(hbar - min_f h_f, 1e-14), 0, 1) when the
// face-midpoint depth h_f = hbar + s . (x_f - x_c) is negative on some face,
// else 1 (reconstruction_jax.py:840-850).
template <typename T, typename T2>
__device__ __forceinline__ T pp_theta(T hbar, T sx, T sy, const T* fx, const T* fy, T2 cc,
                                      int mf) {
    T hmin = T(INFINITY);
#pragma unroll
    for (int k = 0; k < MAXF; ++k)
        if (k < mf) hmin = tmin(hmin, hbar + (sx * (fx[k] - cc.x) + sy * (fy[k] - cc.y)));
    const T den = tmax(hbar - hmin, T(1e-14));
    return hmin < T(0) ? clip01(hbar / den) : T(1);
}

// The 2-D least-squares gradient from the normal-matrix sums of the
// neighbour rows (S) and of the ghost rows that enter gradient component d
// (G[d]): inverse rows (c0, c1) of system 0 (x component) and (c2, c3) of
// system 1 (y component).  These helpers and face_increment spell out their
// fused multiply-adds, so that every pass that computes a cell's slope (the
// own-cell and the face-ring passes of the stage kernel) rounds it alike.
template <typename T, bool LEAN = false>
__device__ __forceinline__ void lsq_inverse_rows(T Sxx, T Sxy, T Syy, const T* Gxx,
                                                 const T* Gxy, const T* Gyy, T& c0, T& c1,
                                                 T& c2, T& c3) {
    const T axx = Sxx + Gxx[0], axy = Sxy + Gxy[0], ayy = Syy + Gyy[0];
    const T di0 = frcp<LEAN>(fma(axx, ayy, -(axy * axy)));
    const T bxx = Sxx + Gxx[1], bxy = Sxy + Gxy[1], byy = Syy + Gyy[1];
    const T di1 = frcp<LEAN>(fma(bxx, byy, -(bxy * bxy)));
    c0 = ayy * di0;
    c1 = -axy * di0;
    c2 = -bxy * di1;
    c3 = bxx * di1;
}

template <typename T>
__device__ __forceinline__ void lsq_solve(T c0, T c1, T c2, T c3, T bx0, T by0, T bx1, T by1,
                                          T& gx, T& gy) {
    gx = fma(c0, bx0, c1 * by0);
    gy = fma(c2, bx1, c3 * by1);
}

// Increment of a slope (gx, gy) over the offset (rx, ry).
template <typename T>
__device__ __forceinline__ T face_increment(T gx, T gy, T rx, T ry) {
    return fma(gx, rx, gy * ry);
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// The staged state is structure-of-arrays in pairs, (b, h) for the ML
// loaded cells followed by (hu, hv): consecutive lanes then read
// consecutive 2-word pairs, free of LDS bank conflicts.
template <typename T>
__device__ __forceinline__ void lds_q4(const typename Vec2<T>::type* Qs, int ML, int i, T* q) {
    const auto a = Qs[i];
    const auto b = Qs[ML + i];
    q[0] = a.x; q[1] = a.y; q[2] = b.x; q[3] = b.y;
}

// Outputs of the reconstruction-only instantiation (AUXONLY = 2): face
// states (4, F) in the reference's face order, written through the per-slot
// table cf = 2 f + side (side 1: the cell is the face's neighbour B), and the
// limited cell gradients (4, dim, N) in conservative variables (or NULL).
template <typename T>
struct ReconOut {
    T* QL;
    T* QR;
    T* lim;
    const int* cf;
    int F;
};

// Ensemble of independent instances stored as one mesh (ENS = 1): the
// instance of every cell, each instance's dt, and two dt partials per wave
// (the minimum over the cells of the wave's first instance, and over the
// cells of a second instance starting inside the wave; instances hold more
// than 64 cells, so a wave meets at most two).
struct EnsArgs {
    const int* inst;
    const double* dts;
    double* part2;
};

// SPEC = 2 fixes only the cell shape of 2-D triangle meshes (3 faces).
// SPEC = 1 compiles the quad-mesh path of the 2-D workload with every
// run-time switch fixed (dim 2, 4 faces, Barth-Jespersen / minmod limiter,
// HLL flux, conservative variables), SPEC = 3 the same numerics on triangle
// meshes, SPEC = 4 the wet/dry numerics on triangle meshes (van Leer on the
// eta reconstruction, HLL; the positivity guards stay run-time switches);
// SPEC = 0 reads them from the mesh and the options.  Every instantiation
// evaluates the same formulas.  AUXONLY = 1
// is the post_step aux refresh alone: staging and the own cells' gradients;
// AUXONLY = 2 the reconstruction object alone (face states, no fluxes).
template <typename T, int SPEC, int AUXONLY = 0, int ENS = 0>
__global__ __launch_bounds__(SBLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4))) void k_stage(
    DevMesh m_in, DevPatch pt, Opts o_in, const T* __restrict__ Qin, const T* Q0, T* Qout,
    const double* __restrict__ dtp, T* __restrict__ aux_out, int mode, int final_stage,
    int write_b, int use_slopes, int* __restrict__ troubled, double* __restrict__ dt_part,
    const uint8_t* __restrict__ o1_mask, uint8_t* __restrict__ bad_out,
    const int* __restrict__ gate, ReconOut<T> ro, EnsArgs ea) {
    // gate: a MOOD redo launch runs only when the candidate had troubled cells
    if (gate && *gate == 0) return;
    DevMesh m = m_in;
    Opts o = o_in;
    if (SPEC == 1 || SPEC == 3) {
        m.dim = 2;
        m.mf = SPEC == 1 ? 4 : 3;
        o.lim = LIM_MINMOD;
        o.rs = RS_HLL;
        o.recon = RECON_CONS;
    } else if (SPEC == 2) {
        // triangle meshes: the face count fixed, the numerics at run time
        m.dim = 2;
        m.mf = 3;
    } else if (SPEC == 4) {
        m.dim = 2;
        m.mf = 3;
        o.lim = LIM_VANLEER;
        o.rs = RS_HLL;
        o.recon = RECON_ETA;
    }
    // the post_step aux rows are gradients of the conservative state, whatever
    // the reconstruction variables (solver_jax.py:284-337)
    if (AUXONLY == 1) o.recon = RECON_CONS;
    using T2 = typename Vec2<T>::type;
    // the lean FP64 reciprocal and quotient (zj_device.hpp) in the
    // instantiations with the numerics fixed at compile time
    constexpr bool LEAN = SPEC == 1 || SPEC == 3;
    // the wet/dry instantiation stages the velocities (u, v) in place of the
    // momenta when it reconstructs (its eta variables convert every stencil
    // row otherwise); the face pass reads the own cell's momenta from Qin
    constexpr bool PRIM = SPEC == 4 && AUXONLY == 0;
    const bool prim = PRIM && use_slopes;
    // extra-stencil ids fetched per batch on triangle meshes (a vertex
    // stencil of up to 12 cells); the batch of the run-time numerics
    // (SPEC = 2) is smaller, for its register budget
    constexpr int MXB = (SPEC == 3 || SPEC == 4) ? 12 : SPEC == 2 ? 6 : SPEC == 1 ? 4 : 0;
    // dynamic LDS, 16-byte aligned (pairs of doubles)
    extern __shared__ double2 smem_v2[];
    double* smem_d = reinterpret_cast<double*>(smem_v2);
    const int p = xcd_block();
    const int tid = threadIdx.x;
    const int N = m.N;
    const int ML = pt.ML, MS = pt.MS, MV = pt.MV;
    const int4 md = pt.meta[p];
    const int tpl = md.x, nload = md.y, nvert = md.z;
    const int nslope = md.w & 0xFFFF, nown = md.w >> 16;
    const int c0 = p * pt.P;  // own cells: the ids c0 .. c0 + nown - 1, or
                              // with expl the first nown entries of lids
    const bool ex = pt.expl != 0;
    T2* Cs = (T2*)smem_d;                   // ML centres (relative to org)
    double2* Vs = (double2*)(Cs + ML);      // MV vertices
    T* Rs = (T*)(Vs + MV);                  // MS inradius (final stage)
    T2* Qs = (T2*)(Rs + MS);                // 2 x ML state pairs
    T* Hs = (T*)(Qs + 2 * ML);              // ML hinv
    T2* Ss = (T2*)(Hs + ML);                // 4 x MS limited slope pairs (x, y)
    const T eps = T(o.eps), g = T(o.g);
    const int* lids = pt.load_ids + (size_t)p * ML;
    // FP32 stages the centres as offsets from the patch's first own cell:
    // their differences (the LSQ geometry and the face offsets) keep what
    // FP32 slopes can use of the FP64 centres, in half the LDS; FP64 stages
    // them as they are (org = 0).  Vertices stay FP64 in the global frame,
    // so both cells of a face compute the same normal and length.
    const double2 org = sizeof(T) == 4 ? pt.cen2[ex ? lids[0] : c0] : make_double2(0.0, 0.0);
    const uint16_t* tn = pt.tpl_nbr + (size_t)tpl * m.mf * MS;
    // word j2 of slope cell ls: extra-stencil entries 2 j2 (low half) and
    // 2 j2 + 1 (high half)
    const int mxh = (m.mx + 1) >> 1;
    const uint32_t* te = pt.tpl_extra2 + (size_t)tpl * mxh * MS;
    const uint16_t* tv = pt.tpl_vert + (size_t)tpl * m.mf * MS;

    // ── phase 1: stage cells and vertices ────────────────────────────
    // SL slots per thread and pass (triangle patches: one pass for up to
    // 768 loaded cells or vertices), every global load of a pass issued
    // before the first LDS store; a slot past the end reads the first own
    // cell / vertex 0 and stores nothing.
    {
        constexpr int SL = SPEC == 1 ? 2 : 3;
        const int* vids = pt.vert_ids + (size_t)p * MV;
        // the state rows (uniform bases: the loads take 32-bit offsets)
        const T* Qin1 = Qin + N;
        const T* Qin2 = Qin + 2 * (size_t)N;
        const T* Qin3 = Qin + 3 * (size_t)N;
        const int nit = nload > nvert ? nload : nvert;
        for (int i0 = tid; i0 < nit; i0 += SL * SBLOCK) {
            int gc[SL], gv[SL];
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const int i = i0 + u * SBLOCK;
                gc[u] = (i < nown && !ex) ? c0 + i : (i < nload ? ldo(lids, i) : ex ? lids[0] : c0);
                gv[u] = i < nvert ? ldo(vids, i) : 0;
            }
            T2 a[SL], b[SL];
            double2 cc[SL], vx[SL];
            double rr[SL];
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const int gi = gc[u];
                a[u].x = ldo(Qin, gi);
                a[u].y = ldo(Qin1, gi);
                b[u].x = ldo(Qin2, gi);
                b[u].y = ldo(Qin3, gi);
                cc[u] = ldo(pt.cen2, gi);
                // inradii: the slope cells' (own + face ring) enter the dt law
                rr[u] = (final_stage && i0 + u * SBLOCK < nslope) ? ldo(m.rin, gi) : 0.0;
                vx[u] = ldo(m.verts, gv[u]);
            }
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const int i = i0 + u * SBLOCK;
                if (i < nload) {
                    Qs[i] = a[u];
                    // with conservative variables only the fluxes read 1/h:
                    // own and face-ring (slope) cells
                    if (!LEAN || i < nslope) {
                        const T hv = hinv_of<T, LEAN>(a[u].y, eps);
                        Hs[i] = hv;
                        if (prim) {
                            b[u].x = b[u].x * hv;
                            b[u].y = b[u].y * hv;
                        }
                    }
                    Qs[ML + i] = b[u];
                    T2 cl;
                    cl.x = T(cc[u].x - org.x);
                    cl.y = T(cc[u].y - org.y);
                    Cs[i] = cl;
                    if (final_stage && i < nslope) Rs[i] = T(rr[u]);
                }
                if (i < nvert) Vs[i] = vx[u];
            }
        }
    }
    // the first slope pass's template rows, in flight during the barrier
    // (the quad instantiation's templates are shared by most patches and
    // cache-resident: in FP64 it loads them in the pass, for its register
    // budget)
    constexpr bool PRE_TPL = !(SPEC == 1 && sizeof(T) == 8);
    const int nsl = AUXONLY == 1 ? nown : nslope;
    TplRow<MXB / 2> row;
    if (PRE_TPL && use_slopes && tid < nsl) row.load(tn, tv, te, m.mf, mxh, MS, tid);
    __syncthreads();

    // ── phase 2: limited slopes of own + face-ring cells ─────────────
    // With the numerics fixed to conservative variables and the Barth-
    // Jespersen / minmod limiter (SPEC 1, 3) the variables' slopes are
    // independent: in FP64 (latency-bound at 3 waves per SIMD) the own cells
    // take one 4-variable pass, and the face-ring cells (64 of a 16 x 16
    // quad patch, which as a second pass would keep one wave busy and three
    // waiting) one variable per wave.  FP32 is bound by VALU issue, and
    // there the per-variable pass's repeated geometry costs more than the
    // idle waves.
    constexpr bool SPLIT = (SPEC == 1 || SPEC == 3) && AUXONLY == 0 && sizeof(T) == 8;
    if (use_slopes) {
        // AUXONLY (post_step aux rows only, launched as mode 3): the own
        // cells' gradients, no limiter, no slopes, no face pass
        const int nsl4 = SPLIT ? (nsl < nown ? nsl : nown) : nsl;
        for (int ls = tid; ls < nsl4; ls += SBLOCK) {
            if (!PRE_TPL || ls != tid) row.load(tn, tv, te, m.mf, mxh, MS, ls);
            const int gi = (ls < nown && !ex) ? c0 + ls : ldo(lids, ls);
            const uint16_t* nls = row.nbr;
            const uint32_t* pre = row.ex;
            T q[4];
            lds_q4(Qs, ML, ls, q);
            const T hic = Hs[ls];
            T w[4];
            if (PRIM) to_w_prim(q, w);
            else to_w(o.recon, q, hic, w);
            double2 v[MAXF];
            T fx[MAXF], fy[MAXF];
            face_centres(m, Vs, row.vert, org, v, fx, fy);
            const T2 cc = Cs[ls];
            T Sxx = 0, Sxy = 0, Syy = 0;
            T Gxx[2] = {0, 0}, Gxy[2] = {0, 0}, Gyy[2] = {0, 0};
            T bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
            T gbx[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            T gby[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            T umax[4], umin[4];
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) { umax[vv] = w[vv]; umin[vv] = w[vv]; }
            ZJ_FOR_FACES(k) {
                const uint16_t nl = nls[k];
                if (nl != LNONE) {
                    T qn[4], wn[4];
                    lds_q4(Qs, ML, nl, qn);
                    if (PRIM) to_w_prim(qn, wn);
                    else to_w(o.recon, qn, Hs[nl], wn);
                    const T2 cn = Cs[nl];
                    const T rx = cn.x - cc.x, ry = cn.y - cc.y;
                    Sxx += rx * rx; Sxy += rx * ry; Syy += ry * ry;
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        T dv = wn[vv] - w[vv];
                        bx[vv] += T(rx) * dv;
                        by[vv] += T(ry) * dv;
                        if (AUXONLY != 1) {  // the aux rows need no bounds
                            umax[vv] = bmax(umax[vv], wn[vv]);
                            umin[vv] = bmin(umin[vv], wn[vv]);
                        }
                    }
                } else {
                    T nx, ny, len;
                    face_normal<T, LEAN>(m, v, k, nx, ny, len);
                    const int bf = -1 - m.nbr[k * N + gi];
                    const int kind = m.bf_kind[bf];
                    T qg[4], wg[4];
                    if (PRIM) {
                        T qc[4];  // the cell's conservative state
                        load_q(Qin, N, gi, qc);
                        ghost_state(m, Qin, bf, qc, T(nx), T(ny), qg);
                    } else {
                        ghost_state(m, Qin, bf, q, T(nx), T(ny), qg);
                    }
                    to_w(o.recon, qg, hinv_of<T, LEAN>(qg[1], eps), wg);
                    const T rx = T(2) * (fx[k] - cc.x), ry = T(2) * (fy[k] - cc.y);
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        T bv = (kind == BC_PERIODIC) ? wg[vv] : T(2) * wg[vv] - w[vv];
                        if (AUXONLY != 1) {
                            umax[vv] = bmax(umax[vv], bv);
                            umin[vv] = bmin(umin[vv], bv);
                        }
                    }
                    const T an[2] = {fabs(nx), fabs(ny)};
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        if (an[d] > 0.5) {
                            Gxx[d] += rx * rx; Gxy[d] += rx * ry; Gyy[d] += ry * ry;
#pragma unroll
                            for (int vv = 0; vv < 4; ++vv) {
                                T dv = wg[vv] - w[vv];
                                gbx[d][vv] += T(rx) * dv;
                                gby[d][vv] += T(ry) * dv;
                            }
                        }
                    }
                }
            }
            auto extra_row = [&](uint16_t el) {
                T qn[4], wn[4];
                lds_q4(Qs, ML, el, qn);
                if (PRIM) to_w_prim(qn, wn);
                else to_w(o.recon, qn, Hs[el], wn);
                const T2 cn = Cs[el];
                const T rx = cn.x - cc.x, ry = cn.y - cc.y;
                Sxx += rx * rx; Sxy += rx * ry; Syy += ry * ry;
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    T dv = wn[vv] - w[vv];
                    bx[vv] += T(rx) * dv;
                    by[vv] += T(ry) * dv;
                }
            };
            if (MXB > 0) {
                // the ids of a batch of MXB entries are fetched before the
                // first is used: one round trip per batch, not per entry
                for (int b2 = 0; b2 < mxh; b2 += MXB / 2) {
                    uint32_t els[MXB > 0 ? MXB / 2 : 1];
#pragma unroll
                    for (int u = 0; u < MXB / 2; ++u)
                        els[u] = b2 == 0 ? pre[u] : b2 + u < mxh ? ldo(te, (b2 + u) * MS + ls) : ~0u;
#pragma unroll
                    for (int j = 0; j < MXB; ++j) {
                        const uint16_t el = (uint16_t)(els[j >> 1] >> (16 * (j & 1)));
                        if (el != LNONE) extra_row(el);
                    }
                }
            } else {
                for (int j2 = 0; j2 < mxh; ++j2) {
                    const uint32_t pair = te[j2 * MS + ls];
                    const uint16_t a = (uint16_t)(pair & 0xFFFFu), b = (uint16_t)(pair >> 16);
                    if (a == LNONE) break;
                    extra_row(a);
                    if (b == LNONE) break;
                    extra_row(b);
                }
            }
            T gx[4], gy[4], ax[4], ay[4];
            if (m.dim == 1) {
                const T inv = T(1) / (Sxx + Gxx[0]);
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    gx[vv] = inv * (bx[vv] + gbx[0][vv]); gy[vv] = T(0);
                    ax[vv] = inv * bx[vv]; ay[vv] = T(0);
                }
            } else {
                T c0, c1, c2, c3;
                lsq_inverse_rows<T, LEAN>(Sxx, Sxy, Syy, Gxx, Gxy, Gyy, c0, c1, c2, c3);
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    lsq_solve(c0, c1, c2, c3, bx[vv] + gbx[0][vv], by[vv] + gby[0][vv],
                              bx[vv] + gbx[1][vv], by[vv] + gby[1][vv], gx[vv], gy[vv]);
                    ax[vv] = c0 * bx[vv] + c1 * by[vv];
                    ay[vv] = c2 * bx[vv] + c3 * by[vv];
                }
            }
            if (aux_out && ls < nown) {
                // post_step aux rows of the stage input (conservative only)
                int row = 0;
                for (int d = 0; d < m.dim; ++d) {
                    aux_out[(size_t)(row++) * N + gi] = d == 0 ? ax[2] : ay[3];
                    aux_out[(size_t)(row++) * N + gi] = d == 0 ? ax[1] : ay[1];
                }
                for (int d = 0; d < m.dim; ++d)
                    aux_out[(size_t)(row++) * N + gi] = d == 0 ? ax[0] : ay[0];
                aux_out[(size_t)row * N + gi] = hic;
            }
            if (AUXONLY == 1) continue;
            T phi[4] = {1, 1, 1, 1};
            if (o.lim == LIM_BJ || o.lim == LIM_MINMOD || o.lim == LIM_VANLEER) {
                // the extreme face increments, then one division per variable
                T dpos[4] = {0, 0, 0, 0}, dneg[4] = {0, 0, 0, 0};
                ZJ_FOR_FACES(k) {
                    const T rx = T(fx[k] - cc.x), ry = T(fy[k] - cc.y);
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        const T d = face_increment(gx[vv], gy[vv], rx, ry);
                        dpos[vv] = bmax(dpos[vv], d);
                        dneg[vv] = bmin(dneg[vv], d);
                    }
                }
#pragma unroll
                for (int vv = 0; vv < 4; ++vv)
                    phi[vv] = o.lim == LIM_VANLEER
                                  ? vl_phi<T, LEAN>(dpos[vv], dneg[vv], umax[vv] - w[vv], umin[vv] - w[vv])
                                  : bj_phi<T, LEAN>(dpos[vv], dneg[vv], umax[vv] - w[vv], umin[vv] - w[vv]);
            } else if (o.lim != LIM_NONE) {
                const double vol = cell_volume(m, v);
                const T ev2 = T(m.dim == 1 ? vol * vol : fabs(vol));
                ZJ_FOR_FACES(k) {
                    const T rx = T(fx[k] - cc.x), ry = T(fy[k] - cc.y);
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        const T d = face_increment(gx[vv], gy[vv], rx, ry);
                        phi[vv] = tmin(phi[vv], lim_cand(o.lim, d, umax[vv] - w[vv], umin[vv] - w[vv], ev2));
                    }
                }
            }
            const bool eta_rows = o.recon == RECON_ETA || o.recon == RECON_ETA_HU;
            if (eta_rows) {
                phi[0] = tmin(phi[0], phi[1]);
                phi[2] = tmin(phi[2], phi[1]);
                phi[3] = tmin(phi[3], phi[1]);
            }
            // MOOD (o1_mask) or the front pre-detector: first order here
            bool first = (o1_mask && o1_mask[gi]) ||
                         (o.recon == RECON_ETA_HU && q[1] < T(ETA_HU_EPS));
            if (o.recon == RECON_XZ || (eta_rows && (o.pos || o.front_tol > 0.0))) {
                // Xing-Zhang-Shu theta of the limited depth slope
                // (reconstruction_jax.py:840-850, :924-937, :1211-1244)
                T sx, sy;
                if (o.recon == RECON_XZ) {
                    sx = phi[1] * gx[1];
                    sy = phi[1] * gy[1];
                } else {
                    sx = phi[1] * gx[1] - phi[0] * gx[0];
                    sy = phi[1] * gy[1] - phi[0] * gy[0];
                }
                const T th = pp_theta(q[1], sx, sy, fx, fy, cc, m.mf);
                if (o.recon == RECON_XZ) {
                    phi[1] *= th;
                    phi[2] *= th;
                    phi[3] *= th;
                } else {
                    if (o.front_tol > 0.0 && (double)th < o.front_tol) first = true;
                    if (o.pos) {
                        // eta slope := phi_b grad b + theta sh, phi_eta = 1
                        gx[1] = phi[0] * gx[0] + th * sx;
                        gy[1] = phi[0] * gy[0] + th * sy;
                        phi[1] = T(1);
                        phi[2] *= th;
                        phi[3] *= th;
                    }
                }
            }
            if (first) {
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) phi[vv] = T(0);
            }
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                T pv = clip01(phi[vv]);
                T2 sv;
                sv.x = pv * gx[vv];
                sv.y = pv * gy[vv];
                Ss[vv * MS + ls] = sv;
            }
            if (AUXONLY == 2 && ro.lim && ls < nown) {
                // limited gradient pushed back to conservative rows, the
                // JVP of the inverse map at the cell (reconstruction_jax.py
                // :1313-1366): dh = deta - db, d(hu) = h du + u dh
                T sx[4], sy[4];
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    const T pv = clip01(phi[vv]);
                    sx[vv] = pv * gx[vv];
                    sy[vv] = pv * gy[vv];
                }
                if (eta_rows) {
                    const T dhx = sx[1] - sx[0], dhy = sy[1] - sy[0];
                    if (o.recon == RECON_ETA) {
                        const T hc = w[1] - w[0];
#pragma unroll
                        for (int vv = 2; vv < 4; ++vv) {
                            sx[vv] = hc * sx[vv] + w[vv] * dhx;
                            sy[vv] = hc * sy[vv] + w[vv] * dhy;
                        }
                    }
                    sx[1] = dhx;
                    sy[1] = dhy;
                }
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    ro.lim[(size_t)(vv * m.dim) * N + gi] = sx[vv];
                    if (m.dim == 2) ro.lim[(size_t)(vv * 2 + 1) * N + gi] = sy[vv];
                }
            }
        }
        if (SPLIT) {
            // face-ring cells: wave vs computes variable vs of 64 cells per
            // round; the arithmetic of each variable is the 4-variable
            // pass's, term for term
            const int vs = __builtin_amdgcn_readfirstlane(tid >> 6);
            const int half = vs >> 1, hi = vs & 1;
            auto pick = [&](int i) {
                const T2 pr = Qs[half * ML + i];
                return hi ? pr.y : pr.x;
            };
            for (int r0 = nown; r0 < nsl; r0 += 64) {
                const int ls = r0 + (tid & 63);
                if (ls >= nsl) break;
                TplRow<MXB / 2> rw;
                rw.load(tn, tv, te, m.mf, mxh, MS, ls);
                const int gi = ldo(lids, ls);
                const T w = pick(ls);
                double2 v[MAXF];
                T fx[MAXF], fy[MAXF];
                face_centres(m, Vs, rw.vert, org, v, fx, fy);
                const T2 cc = Cs[ls];
                T Sxx = 0, Sxy = 0, Syy = 0;
                T Gxx[2] = {0, 0}, Gxy[2] = {0, 0}, Gyy[2] = {0, 0};
                T bx = 0, by = 0, gbx[2] = {0, 0}, gby[2] = {0, 0};
                T umax = w, umin = w;
                ZJ_FOR_FACES(k) {
                    const uint16_t nl = rw.nbr[k];
                    if (nl != LNONE) {
                        const T wn = pick(nl);
                        const T2 cn = Cs[nl];
                        const T rx = cn.x - cc.x, ry = cn.y - cc.y;
                        Sxx += rx * rx; Sxy += rx * ry; Syy += ry * ry;
                        const T dv = wn - w;
                        bx += T(rx) * dv;
                        by += T(ry) * dv;
                        umax = bmax(umax, wn);
                        umin = bmin(umin, wn);
                    } else {
                        T nx, ny, len;
                        face_normal<T, LEAN>(m, v, k, nx, ny, len);
                        const int bf = -1 - m.nbr[k * N + gi];
                        const int kind = m.bf_kind[bf];
                        T q[4], qg[4];
                        lds_q4(Qs, ML, ls, q);
                        ghost_state(m, Qin, bf, q, T(nx), T(ny), qg);
                        const T wg = vs == 0 ? qg[0] : vs == 1 ? qg[1] : vs == 2 ? qg[2] : qg[3];
                        const T rx = T(2) * (fx[k] - cc.x), ry = T(2) * (fy[k] - cc.y);
                        const T bv = (kind == BC_PERIODIC) ? wg : T(2) * wg - w;
                        umax = bmax(umax, bv);
                        umin = bmin(umin, bv);
                        const T an[2] = {fabs(nx), fabs(ny)};
#pragma unroll
                        for (int d = 0; d < 2; ++d) {
                            if (an[d] > 0.5) {
                                Gxx[d] += rx * rx; Gxy[d] += rx * ry; Gyy[d] += ry * ry;
                                const T dv = wg - w;
                                gbx[d] += T(rx) * dv;
                                gby[d] += T(ry) * dv;
                            }
                        }
                    }
                }
                auto extra_row = [&](uint16_t el) {
                    const T wn = pick(el);
                    const T2 cn = Cs[el];
                    const T rx = cn.x - cc.x, ry = cn.y - cc.y;
                    Sxx += rx * rx; Sxy += rx * ry; Syy += ry * ry;
                    const T dv = wn - w;
                    bx += T(rx) * dv;
                    by += T(ry) * dv;
                };
                for (int b2 = 0; b2 < mxh; b2 += MXB / 2) {
                    uint32_t els[MXB > 0 ? MXB / 2 : 1];
#pragma unroll
                    for (int u = 0; u < MXB / 2; ++u)
                        els[u] = b2 == 0 ? rw.ex[u] : b2 + u < mxh ? ldo(te, (b2 + u) * MS + ls) : ~0u;
#pragma unroll
                    for (int j = 0; j < MXB; ++j) {
                        const uint16_t el = (uint16_t)(els[j >> 1] >> (16 * (j & 1)));
                        if (el != LNONE) extra_row(el);
                    }
                }
                T c0, c1, c2, c3, gx, gy;
                lsq_inverse_rows<T, LEAN>(Sxx, Sxy, Syy, Gxx, Gxy, Gyy, c0, c1, c2, c3);
                lsq_solve(c0, c1, c2, c3, bx + gbx[0], by + gby[0], bx + gbx[1], by + gby[1],
                          gx, gy);
                T dpos = 0, dneg = 0;
                ZJ_FOR_FACES(k) {
                    const T rx = T(fx[k] - cc.x), ry = T(fy[k] - cc.y);
                    const T d = face_increment(gx, gy, rx, ry);
                    dpos = bmax(dpos, d);
                    dneg = bmin(dneg, d);
                }
                T phi = bj_phi<T, LEAN>(dpos, dneg, umax - w, umin - w);
                if (o1_mask && o1_mask[gi]) phi = T(0);
                const T pv = clip01(phi);
                T2 sv;
                sv.x = pv * gx;
                sv.y = pv * gy;
                Ss[vs * MS + ls] = sv;
            }
        }
    }
    // the face pass's global inputs (own cell id, step state and dt of the
    // RK average, template rows), in flight during the barrier; in-place
    // stage 2 reads only the own cell of Q0, which this thread writes last
    int gi3 = 0;
    T q0[4] = {0, 0, 0, 0};
    T qc3[4] = {0, 0, 0, 0};  // staged velocities: the own cell's conservative state
    T dt3 = T(0);
    TplRow<0> row3;
    if (AUXONLY != 1 && tid < nown) {
        gi3 = ex ? ldo(lids, tid) : c0 + tid;
        if (PRE_TPL) row3.load(tn, tv, te, m.mf, mxh, MS, tid);
        if (AUXONLY == 0) {
            if (prim) load_q(Qin, N, gi3, qc3);
            if (mode == 1) load_q(Q0, N, gi3, q0);
            dt3 = mode == 2 ? T(0) : T(ENS ? ea.dts[ea.inst[gi3]] : *dtp);
        }
    }
    if (use_slopes) __syncthreads();
    if (AUXONLY == 1) return;
    if (!PRE_TPL && tid < nown) row3.load(tn, tv, te, m.mf, mxh, MS, tid);
    if (AUXONLY == 2) {
        // the reconstruction object: each own cell's face values into the
        // face arrays; a boundary face's Q_R is the ghost of the cell state
        // (the bf values of halo_bc, blocks.py:129-157)
        if (tid < nown) {
            const int gi = gi3;
            T q[4];
            lds_q4(Qs, ML, tid, q);
            const T hic = Hs[tid];
            T sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (use_slopes) {
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    const T2 sv = Ss[vv * MS + tid];
                    sc[2 * vv] = sv.x;
                    sc[2 * vv + 1] = sv.y;
                }
            }
            double2 v[MAXF];
            T fx[MAXF], fy[MAXF];
            face_centres(m, Vs, row3.vert, org, v, fx, fy);
            const T2 cc = Cs[tid];
            ZJ_FOR_FACES(k) {
                T qf[4] = {q[0], q[1], q[2], q[3]};
                if (use_slopes)
                    recon_face(o.recon, q, hic, sc, T(fx[k] - cc.x), T(fy[k] - cc.y), qf);
                const int f2 = ro.cf[(size_t)k * N + gi];
                const int f = f2 >> 1;
                T* dst = (f2 & 1) ? ro.QR : ro.QL;
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) dst[(size_t)vv * ro.F + f] = qf[vv];
                const int nb = m.nbr[(size_t)k * N + gi];
                if (nb < 0) {
                    T nxd, nyd, len;
                    face_normal<T, LEAN>(m, v, k, nxd, nyd, len);
                    T qg[4];
                    ghost_state(m, Qin, -1 - nb, q, T(nxd), T(nyd), qg);
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) ro.QR[(size_t)vv * ro.F + f] = qg[vv];
                }
            }
        }
        return;
    }

    // ── phase 3: fluxes of the own cells ─────────────────────────────
    int bad = 0;
    double best_r = INFINITY, best_a = 1.0;  // min over faces of r / a
    T out_keep[4] = {0, 0, 0, 0};
    int gi_keep = -1;
    const int ls = tid;
    if (ls < nown) {
        const int gi = gi3;
        const T dt = dt3;
        const uint16_t* nls = row3.nbr;
        T q[4];
        lds_q4(Qs, ML, ls, q);  // with prim: (b, h, u, v) until the update
        const T hic = Hs[ls];
        // the fluxes' 1/h: this stage's state, or in a frozen second stage
        // the step state's (the face reconstruction keeps this stage's)
        const bool frz = SPEC != 1 && SPEC != 3 && o.stage_aux && mode == 1;
        const T hif = frz ? hinv_of<T, LEAN>(q0[1], eps) : hic;
        T sc[8];
        if (use_slopes) {
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const T2 sv = Ss[vv * MS + ls];
                sc[2 * vv] = sv.x;
                sc[2 * vv + 1] = sv.y;
            }
        }
        double2 v[MAXF];
        T fx[MAXF], fy[MAXF];
        face_centres(m, Vs, row3.vert, org, v, fx, fy);
        const T2 cc = Cs[ls];
        T nxs[MAXF], nys[MAXF];
        T acc[4] = {0, 0, 0, 0};
        ZJ_FOR_FACES(k) {
            T nxd, nyd, len;
            face_normal<T, LEAN>(m, v, k, nxd, nyd, len);
            nxs[k] = nxd;
            nys[k] = nyd;
            const T nx = T(nxd), ny = T(nyd);
            T qL[4], qR[4];
            if (prim) recon_face_prim(q, sc, T(fx[k] - cc.x), T(fy[k] - cc.y), qL);
            else if (use_slopes) recon_face(o.recon, q, hic, sc, T(fx[k] - cc.x), T(fy[k] - cc.y), qL);
            else { qL[0] = q[0]; qL[1] = q[1]; qL[2] = q[2]; qL[3] = q[3]; }
            const uint16_t nl = nls[k];
            T hiR = hic, hfR = hif;
            if (nl != LNONE) {
                T qn[4];
                lds_q4(Qs, ML, nl, qn);
                hiR = Hs[nl];
                hfR = hiR;
                if (frz) hfR = hinv_of<T, LEAN>(Q0[(size_t)N + ((nl < nown && !ex) ? c0 + nl : lids[nl])], eps);
                if (use_slopes) {
                    T sn[8];
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        const T2 sv = Ss[vv * MS + nl];
                        sn[2 * vv] = sv.x;
                        sn[2 * vv + 1] = sv.y;
                    }
                    const T2 cn = Cs[nl];
                    if (prim) recon_face_prim(qn, sn, T(fx[k] - cn.x), T(fy[k] - cn.y), qR);
                    else recon_face(o.recon, qn, hiR, sn, T(fx[k] - cn.x), T(fy[k] - cn.y), qR);
                } else {
                    qR[0] = qn[0]; qR[1] = qn[1]; qR[2] = qn[2]; qR[3] = qn[3];
                }
            } else {
                const int bf = -1 - m.nbr[k * N + gi];
                ghost_state(m, Qin, bf, qL, nx, ny, qR);
            }
            T f[4];
            face_flux_left<T, LEAN>(o.rs, qL, qR, hif, hfR, nx, ny, g, T(o.eps), f);
            const T L = T(len);
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) acc[vv] += f[vv] * L;
            if (use_slopes) {
                T ce[4];
                cell_edge(q, qL, nx, ny, g, ce);
                acc[2] += ce[2] * L;
                acc[3] += ce[3] * L;
            }
        }
        const T ivol = frcp<LEAN>(T(cell_volume(m, v)));
        if (prim) {
            // the update reads the conservative momenta
            q[2] = qc3[2];
            q[3] = qc3[3];
        }
        T out[4];
        if (mode == 2) {
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) out[vv] = -acc[vv] * ivol;
        } else {
            if (mode == 1) {
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) out[vv] = T(0.5) * (q0[vv] + (q[vv] - dt * (acc[vv] * ivol)));
            } else {
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) out[vv] = q[vv] - dt * (acc[vv] * ivol);
            }
            if (final_stage && o.manning > 0.0) {
                const T h0 = mode == 1 ? q0[1] : q[1];
                const T hs = hinv_of<T, LEAN>(h0, eps);
                const T u = out[2] * hs, vv2 = out[3] * hs;
                const T sp = sqrt(u * u + vv2 * vv2);
                const T kk = T(o.g * o.manning * o.manning) * cbrt(hs);
                out[2] += dt * (-kk * u * sp);
                out[3] += dt * (-kk * vv2 * sp);
            }
        }
        if (final_stage && mode != 2) {
            bad = (out[1] < T(0)) || !isfinite(out[0]) || !isfinite(out[1]) ||
                  !isfinite(out[2]) || !isfinite(out[3]);
            if (bad_out) bad_out[gi] = (uint8_t)bad;
            if (dt_part || (ENS && ea.part2)) {
                // min over faces of r_f / a_f without a division per face
                // (in the stage precision; the products of the comparison
                // are exact in FP64 for FP32 operands)
                const T hn = hinv_of<T, LEAN>(out[1], eps);
                const T c = fsqrt_nn<LEAN>(g * bmax(out[1], T(0)));
                const T rc = Rs[ls];
                ZJ_FOR_FACES(k) {
                    const uint16_t nl = nls[k];
                    const T rf = nl != LNONE ? bmin(rc, Rs[nl]) : rc;
                    const T a = fabs((out[2] * nxs[k] + out[3] * nys[k]) * hn) + c;
                    if (!(a == a)) { best_r = 0.0; best_a = 1.0; }
                    else if ((double)rf * best_a < best_r * (double)a) { best_r = rf; best_a = a; }
                }
            }
        }
        out_keep[0] = out[0]; out_keep[1] = out[1]; out_keep[2] = out[2]; out_keep[3] = out[3];
        gi_keep = gi;
    }
    // Wave-level reductions, no block barrier: the troubled count is added
    // only by waves that found one, the dt minimum of every wave goes to a
    // per-wave slot that k_reduce_part folds (deterministic, no contention).
    if (final_stage && mode != 2 && (troubled || dt_part)) {
        double loc_dt = best_r == INFINITY ? INFINITY
                                           : fdiv<LEAN>(o.cfl * 2.0 * best_r, (double)m.dim * best_a);
        if (loc_dt != loc_dt) loc_dt = 0.0;
        for (int off = 32; off > 0; off >>= 1) {
            bad += __shfl_xor(bad, off, 64);
            loc_dt = fmin(loc_dt, __shfl_xor(loc_dt, off, 64));
        }
        if ((tid & 63) == 0) {
            if (troubled && bad) atomicAdd(troubled, bad);
            if (dt_part) dt_part[(size_t)p * (SBLOCK / 64) + (tid >> 6)] = loc_dt;
        }
    }
    if (ENS && final_stage && mode != 2 && ea.part2) {
        double loc_dt = best_r == INFINITY ? INFINITY
                                           : fdiv<LEAN>(o.cfl * 2.0 * best_r, (double)m.dim * best_a);
        if (loc_dt != loc_dt) loc_dt = 0.0;
        // own cells are a prefix of the patch: lane 0 of a wave with cells
        // has one, and names the wave's first instance
        const int my = gi_keep >= 0 ? ea.inst[gi_keep] : -1;
        const int i0 = __shfl(my, 0, 64);
        double m0 = (my >= 0 && my == i0) ? loc_dt : INFINITY;
        double m1 = (my >= 0 && my != i0) ? loc_dt : INFINITY;
        for (int off = 32; off > 0; off >>= 1) {
            m0 = fmin(m0, __shfl_xor(m0, off, 64));
            m1 = fmin(m1, __shfl_xor(m1, off, 64));
        }
        if ((tid & 63) == 0) {
            const size_t w = (size_t)p * (SBLOCK / 64) + (tid >> 6);
            ea.part2[2 * w] = m0;
            ea.part2[2 * w + 1] = m1;
        }
    }
    if (gi_keep >= 0) {
        if (write_b) sto(Qout, gi_keep, out_keep[0]);
        sto(Qout + N, gi_keep, out_keep[1]);
        sto(Qout + 2 * (size_t)N, gi_keep, out_keep[2]);
        sto(Qout + 3 * (size_t)N, gi_keep, out_keep[3]);
    }
}

}  // namespace

// ═════════════════════════════════════════════════════════════════════════
// C ABI
// ═════════════════════════════════════════════════════════════════════════
struct zj_mesh {
    DevMesh d;
    DevPatch pt;
    double2* slot_n = nullptr;  // (mf, N) outward normals (dt law)
    double* slot_r = nullptr;   // (mf, N) min(r_cell, r_nbr)
    int n_templates = 0;
    std::vector<void*> allocs;
    int* bf_cell = nullptr;     // (nbf) inner cell of each boundary face
    int* bf_slot = nullptr;     // (nbf) its slot in that cell
    // reference-layout face list (zj_mesh_set_faces), for the per-face
    // entry points
    int nfaces = 0;
    int* fc = nullptr;          // (2, F) owner A, neighbour B (-1: boundary)
    int* fs = nullptr;          // (2, F) slot of the face in A, in B
    int* cell_face = nullptr;   // (mf, N) 2 f + side of each cell slot
    // When the automatic layout cuts the patches from clusters of the Morton
    // order (a strip-like cell order, such as RCM), the march runs on this
    // mesh: the same cells renumbered in the cluster order, whose patches
    // are consecutive ids (own cells implicit, own-cell loads and stores in
    // runs).  ren_order[k] is the caller's id of renumbered cell k.
    zj_mesh* ren = nullptr;
    int* ren_order = nullptr;
};

// centres (2 words) + state and hinv (5 words) per loaded cell, inradius (1
// word) + limited slopes (8 words) per slope cell, FP64 vertices
static size_t stage_lds_bytes(const DevPatch& pt, size_t esz) {
    return (size_t)pt.ML * 7 * esz + (size_t)pt.MV * 16 + (size_t)pt.MS * 9 * esz;
}
// FP64 working set of one patch: two blocks per CU (160 KB of LDS)
static constexpr size_t LDS_LIMIT = 80 * 1024;

// The patch decomposition zj_mesh_create builds: patches of the largest
// size (BLOCK, or the caller's patch_cells) whose FP64 working set fits the
// LDS budget, cut from consecutive ids or -- when that order drags in more
// than 3x a patch's size of ring cells (orders whose id runs are thin
// strips, such as RCM) and clusters do better, or when asked -- from
// clusters of the Morton order of the cell centres.
static int plan_patches(const zj_mesh_desc* d, zj::PatchHost& ph,
                        std::vector<int32_t>* order_out = nullptr) {
    const int N = d->n_cells, mf = d->max_faces, mx = d->max_extra;
    std::vector<int32_t> ex_dummy(N, -1);
    const int32_t* exh = mx ? d->cell_extra : ex_dummy.data();
    int P0 = d->patch_cells > 0 ? d->patch_cells : BLOCK;
    if (P0 < 32 || P0 > BLOCK || (P0 & (P0 - 1))) return ZJ_EINVAL;
    if (d->patch_order < 0 || d->patch_order > 2) return ZJ_EINVAL;
    if (d->patch_order == 2 && d->dim != 2) return ZJ_EUNSUP;
    auto fit = [&](const int32_t* order, zj::PatchHost& out) -> int {
        for (int P = P0;; P /= 2) {
            if (zj::build_patches(N, mf, mx ? mx : 1, d->n_verts, d->cell_vertices, d->cell_nbr,
                                  exh, P, order, out))
                return ZJ_EINVAL;
            DevPatch probe{};
            probe.ML = out.max_load; probe.MS = out.max_slope; probe.MV = out.max_vert;
            if (stage_lds_bytes(probe, 8) <= LDS_LIMIT) return ZJ_OK;
            if (P <= 32) return ZJ_EUNSUP;
        }
    };
    std::vector<int32_t> order;
    int rc = ZJ_OK;
    if (d->patch_order != 2) rc = fit(nullptr, ph);
    const bool strips = rc != ZJ_OK || ph.max_load > 3 * ph.P;
    if (d->patch_order == 2 || (d->patch_order == 0 && d->dim == 2 && strips)) {
        order = zj::morton_cluster_order(N, d->cell_center);
        zj::PatchHost pc;
        const int rcc = fit(order.data(), pc);
        if (rcc == ZJ_OK && (d->patch_order == 2 || rc != ZJ_OK || pc.P > ph.P ||
                             (pc.P == ph.P && 5 * pc.max_load < 4 * ph.max_load))) {
            ph = std::move(pc);
            rc = ZJ_OK;
            if (order_out) *order_out = order;
        } else if (d->patch_order == 2) {
            rc = rcc;
        }
    }
    if (rc != ZJ_OK) return rc;
    if (zj::validate_patches(N, mf, mx ? mx : 1, d->n_verts, d->cell_nbr, exh, d->cell_vertices,
                             ph))
        return ZJ_EINVAL;
    return ZJ_OK;
}

struct zj_solver {
    zj_mesh* m;
    zj_params p;
    Opts o;
    void* Q1;                   // (4, N) stage state
    unsigned long long* dt_red; // reduction scalar
    double* dt_part;            // per-wave dt minima of a stage launch
    int* troubled;              // scratch counter
    double2* mass_part;         // per-block mass partials
    double* dt_tmp;             // device dt
    int mass_blocks;
    // a-posteriori MOOD (off by default)
    int mood = 0;
    void* Qc = nullptr;             // (4, N) candidate state
    uint8_t* mood_mask = nullptr;   // per-cell first-order flags
    int* mood_cnt = nullptr;        // troubled cells of the candidate
    int* mood_stats = nullptr;      // {steps redone, cells demoted}
    // (4, N) second-stage output of an in-place step with the frozen stage
    // aux: that stage reads the neighbours' step-start depth from Q0, which
    // other workgroups of the same launch would be overwriting
    void* Qfrz = nullptr;
    // the march on the mesh's renumbered twin (zj_mesh::ren): a solver of
    // the same numerics and options on it, and the (4, N) state in its order
    zj_solver* ren = nullptr;
    void* Qp = nullptr;
    // set by zj_march around its steps: the stage-2 launches leave their dt
    // minima, which one fold at the end of the step reduces (with MOOD: the
    // redo's when a redo ran -- it overwrites the candidate's -- so no reset
    // is needed), and the MOOD bookkeeping moves into the next step's head
    int defer_dt = 0;
};

namespace zj {
int solver_dtype(const zj_solver* s) { return s ? s->p.dtype : -1; }
int solver_cells(const zj_solver* s) { return s ? s->m->d.N : -1; }
int solver_dim(const zj_solver* s) { return s ? s->m->d.dim : -1; }
zj_params solver_params(const zj_solver* s) { return s->p; }
}  // namespace zj

#define HIPCHK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "zoomy_hip: %s failed: %s\n", #x,         \
                    hipGetErrorString(e_));                           \
            return ZJ_EHIP;                                           \
        }                                                             \
    } while (0)

static int dev_upload(zj_mesh* m, const void* src, size_t bytes, void** dst) {
    if (bytes == 0) bytes = 8;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return ZJ_ENOMEM;
    m->allocs.push_back(p);
    if (src) HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(p, 0, bytes));
    *dst = p;
    return ZJ_OK;
}

static inline int nblocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// Dynamic LDS above 64 KB is granted per kernel: every stage-kernel
// instantiation of element type T may stage the mesh's patch working set.
template <typename T>
static int grant_stage_lds(size_t lds) {
    if (lds <= 64 * 1024) return ZJ_OK;
    const void* ks[] = {(const void*)k_stage<T, 0>,       (const void*)k_stage<T, 1>,
                        (const void*)k_stage<T, 2>,       (const void*)k_stage<T, 3>,
                        (const void*)k_stage<T, 4>,
                        (const void*)k_stage<T, 0, 1>,    (const void*)k_stage<T, 1, 1>,
                        (const void*)k_stage<T, 3, 1>,    (const void*)k_stage<T, 0, 2>,
                        (const void*)k_stage<T, 0, 0, 1>, (const void*)k_stage<T, 2, 0, 1>,
                        (const void*)k_stage<T, 3, 0, 1>};
    for (const void* k : ks)
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
            return ZJ_EHIP;
    return ZJ_OK;
}

extern "C" {

const char* zj_version(void) { return "zoomy_hip 0.1 gfx950"; }

int zj_patch_plan(const zj_mesh_desc* d, int32_t* stats_out) {
    if (!d || !stats_out || d->n_cells <= 0 || d->dim < 1 || d->dim > 2 || d->max_faces < 2 ||
        d->max_faces > MAXF || d->max_extra < 0)
        return ZJ_EINVAL;
    zj::PatchHost ph;
    const int rc = plan_patches(d, ph);
    if (rc != ZJ_OK) return rc;
    stats_out[0] = ph.P;
    stats_out[1] = ph.npatch;
    stats_out[2] = ph.ntpl;
    stats_out[3] = ph.max_load;
    stats_out[4] = ph.max_slope;
    stats_out[5] = ph.max_vert;
    stats_out[6] = ph.expl;
    return ZJ_OK;
}

// The mesh of d with its cells renumbered: new cell k is the caller's cell
// order[k]; neighbour and stencil ids, the boundary faces' inner cells and
// the per-cell geometry follow, and the patches are consecutive ids.
static int build_renumbered(const zj_mesh_desc* d, const std::vector<int32_t>& order,
                            zj_mesh* m) {
    const int N = d->n_cells, mf = d->max_faces, mx = d->max_extra;
    std::vector<int32_t> inv(N);
    for (int k = 0; k < N; ++k) inv[order[k]] = k;
    std::vector<int32_t> cv((size_t)mf * N), nbr((size_t)mf * N), ex((size_t)mx * N);
    std::vector<double> cen((size_t)2 * N), rin(N);
    for (int k = 0; k < N; ++k) {
        const int c = order[k];
        for (int s = 0; s < mf; ++s) {
            cv[(size_t)s * N + k] = d->cell_vertices[(size_t)s * N + c];
            const int nb = d->cell_nbr[(size_t)s * N + c];
            nbr[(size_t)s * N + k] = nb >= 0 ? inv[nb] : nb;
        }
        for (int j = 0; j < mx; ++j) {
            const int e = d->cell_extra[(size_t)j * N + c];
            ex[(size_t)j * N + k] = e >= 0 ? inv[e] : e;
        }
        cen[k] = d->cell_center[c];
        cen[(size_t)N + k] = d->cell_center[(size_t)N + c];
        rin[k] = d->cell_inradius[c];
    }
    std::vector<int32_t> bcell(d->n_bf > 0 ? d->n_bf : 1);
    for (int b = 0; b < d->n_bf; ++b) bcell[b] = inv[d->bface_cell[b]];
    zj_mesh_desc d2 = *d;
    d2.cell_vertices = cv.data();
    d2.cell_nbr = nbr.data();
    d2.cell_extra = mx ? ex.data() : nullptr;
    d2.cell_center = cen.data();
    d2.cell_inradius = rin.data();
    d2.bface_cell = bcell.data();
    d2.patch_order = 1;  // consecutive ids: the clusters themselves
    int rc = zj_mesh_create(&d2, &m->ren);
    if (rc) return rc;
    void* p = nullptr;
    if (dev_upload(m, order.data(), sizeof(int32_t) * N, &p)) return ZJ_ENOMEM;
    m->ren_order = (int*)p;
    return ZJ_OK;
}

int zj_mesh_create(const zj_mesh_desc* d, zj_mesh** out) {
    if (!d || !out) return ZJ_EINVAL;
    if (d->dim < 1 || d->dim > 2 || d->n_cells <= 0 || d->max_faces < 2 ||
        d->max_faces > MAXF || d->max_extra < 0 || d->n_bf < 0)
        return ZJ_EINVAL;
    if (d->dim == 1 && d->max_faces != 2) return ZJ_EINVAL;
    // the kernels index rows of (rows, N) arrays with 32-bit offsets (up to
    // 8 rows: the state, the aux rows, the slot tables)
    if ((int64_t)d->n_cells * 8 > (int64_t)INT32_MAX || (int64_t)d->n_cells * d->max_extra >
                                                             (int64_t)INT32_MAX)
        return ZJ_EUNSUP;
    const int N = d->n_cells, mf = d->max_faces, mx = d->max_extra;
    // host-side index validation: every gather in the kernels stays in bounds
    for (size_t i = 0; i < (size_t)mf * N; ++i) {
        int v = d->cell_vertices[i];
        if (v < 0 || v >= d->n_verts) return ZJ_EINVAL;
        int nb = d->cell_nbr[i];
        if (nb >= N || (nb < 0 && -1 - nb >= d->n_bf)) return ZJ_EINVAL;
    }
    for (size_t i = 0; i < (size_t)mx * N; ++i)
        if (d->cell_extra[i] >= N) return ZJ_EINVAL;
    std::vector<int> pcell(d->n_bf > 0 ? d->n_bf : 1, -1);
    for (int b = 0; b < d->n_bf; ++b) {
        // boundary face b sits in slot bface_slot[b] of cell bface_cell[b]
        const int bc = d->bface_cell[b], bs = d->bface_slot[b];
        if (bc < 0 || bc >= N || bs < 0 || bs >= mf ||
            d->cell_nbr[(size_t)bs * N + bc] != -1 - b)
            return ZJ_EINVAL;
        int k = d->bface_kind[b];
        if (k < 0 || k > 3) return ZJ_EINVAL;
        if (k == ZJ_BC_DIRICHLET && (!d->bface_mask || !d->bface_value ||
                                     (d->bface_mask[b] & ~15)))
            return ZJ_EINVAL;
        if (k == ZJ_BC_PERIODIC) {
            int pb = d->bface_partner ? d->bface_partner[b] : -1;
            if (pb < 0 || pb >= d->n_bf) return ZJ_EINVAL;
            pcell[b] = d->bface_cell[pb];
            if (pcell[b] < 0 || pcell[b] >= N) return ZJ_EINVAL;
        }
    }
    zj_mesh* m = new (std::nothrow) zj_mesh();
    if (!m) return ZJ_ENOMEM;
    std::vector<int32_t> cluster_order;
    m->d.dim = d->dim; m->d.N = N; m->d.nverts = d->n_verts;
    m->d.mf = mf; m->d.mx = mx; m->d.nbf = d->n_bf;
    void* p;
    int rc = 0;
    rc |= dev_upload(m, d->cell_vertices, sizeof(int) * mf * N, &p); m->d.cv = (const int*)p;
    rc |= dev_upload(m, d->cell_nbr, sizeof(int) * mf * N, &p); m->d.nbr = (const int*)p;
    rc |= dev_upload(m, mx ? d->cell_extra : nullptr, sizeof(int) * mx * N, &p); m->d.extra = (const int*)p;
    rc |= dev_upload(m, d->verts, sizeof(double) * 2 * d->n_verts, &p); m->d.verts = (const double2*)p;
    rc |= dev_upload(m, d->cell_center, sizeof(double) * 2 * N, &p); m->d.cen = (const double*)p;
    rc |= dev_upload(m, d->cell_inradius, sizeof(double) * N, &p); m->d.rin = (const double*)p;
    rc |= dev_upload(m, d->n_bf ? d->bface_kind : nullptr, sizeof(int) * d->n_bf, &p); m->d.bf_kind = (const int*)p;
    rc |= dev_upload(m, pcell.data(), sizeof(int) * pcell.size(), &p); m->d.bf_pcell = (const int*)p;
    {
        // Dirichlet rows (zero masks when the mesh has none)
        const int nb = d->n_bf > 0 ? d->n_bf : 1;
        std::vector<int> mask(nb, 0);
        std::vector<double> val((size_t)4 * nb, 0.0);
        if (d->bface_mask && d->bface_value)
            for (int b = 0; b < d->n_bf; ++b) {
                mask[b] = d->bface_kind[b] == ZJ_BC_DIRICHLET ? d->bface_mask[b] : 0;
                for (int v = 0; v < 4; ++v)
                    val[(size_t)v * nb + b] = d->bface_value[(size_t)v * d->n_bf + b];
            }
        rc |= dev_upload(m, mask.data(), sizeof(int) * mask.size(), &p); m->d.bf_mask = (const int*)p;
        rc |= dev_upload(m, d->n_bf ? d->bface_cell : nullptr, sizeof(int) * d->n_bf, &p);
        m->bf_cell = (int*)p;
        rc |= dev_upload(m, d->n_bf ? d->bface_slot : nullptr, sizeof(int) * d->n_bf, &p);
        m->bf_slot = (int*)p;
        rc |= dev_upload(m, val.data(), sizeof(double) * val.size(), &p); m->d.bf_val = (const double*)p;
    }
    if (rc) {
        zj_mesh_destroy(m);
        return ZJ_ENOMEM;
    }
    // patch decomposition for the LDS-staged stage kernel
    {
        std::vector<int32_t> ex_dummy(N, -1);
        zj::PatchHost ph;
        const int prc = plan_patches(d, ph, &cluster_order);
        if (prc != ZJ_OK) {
            zj_mesh_destroy(m);
            return prc;
        }
        DevPatch& pt = m->pt;
        pt.P = ph.P; pt.npatch = ph.npatch; pt.ML = ph.max_load; pt.MS = ph.max_slope;
        pt.MV = ph.max_vert; pt.expl = ph.expl;
        std::vector<double> cen2((size_t)2 * N);
        for (int c = 0; c < N; ++c) {
            cen2[2 * (size_t)c] = d->cell_center[c];
            cen2[2 * (size_t)c + 1] = d->cell_center[N + c];
        }
        rc = 0;
        std::vector<int32_t> meta((size_t)4 * ph.npatch);
        for (int q = 0; q < ph.npatch; ++q) {
            const int t = ph.tpl_id[q];
            meta[4 * (size_t)q] = t;
            meta[4 * (size_t)q + 1] = ph.n_load[q];
            meta[4 * (size_t)q + 2] = ph.n_vert[q];
            meta[4 * (size_t)q + 3] = ph.tpl_nslope[t] | (ph.tpl_nown[t] << 16);
        }
        rc |= dev_upload(m, meta.data(), sizeof(int32_t) * meta.size(), &p); pt.meta = (const int4*)p;
        rc |= dev_upload(m, ph.load_ids.data(), sizeof(int) * ph.load_ids.size(), &p); pt.load_ids = (const int*)p;
        rc |= dev_upload(m, ph.vert_ids.data(), sizeof(int) * ph.vert_ids.size(), &p); pt.vert_ids = (const int*)p;
        rc |= dev_upload(m, ph.tpl_nbr.data(), 2 * ph.tpl_nbr.size(), &p); pt.tpl_nbr = (const uint16_t*)p;
        {
            // extra-stencil ids packed in pairs (one 32-bit load per two)
            const int S = ph.max_slope, mxh = (ph.mx + 1) / 2;
            std::vector<uint32_t> te2((size_t)ph.ntpl * mxh * S);
            for (int t = 0; t < ph.ntpl; ++t)
                for (int j2 = 0; j2 < mxh; ++j2)
                    for (int ls = 0; ls < S; ++ls) {
                        const uint32_t lo = ph.tpl_extra[((size_t)t * ph.mx + 2 * j2) * S + ls];
                        const uint32_t hi = 2 * j2 + 1 < ph.mx
                                                ? ph.tpl_extra[((size_t)t * ph.mx + 2 * j2 + 1) * S + ls]
                                                : 0xFFFFu;
                        te2[((size_t)t * mxh + j2) * S + ls] = lo | hi << 16;
                    }
            rc |= dev_upload(m, te2.data(), 4 * te2.size(), &p);
            pt.tpl_extra2 = (const uint32_t*)p;
        }
        rc |= dev_upload(m, ph.tpl_vert.data(), 2 * ph.tpl_vert.size(), &p); pt.tpl_vert = (const uint16_t*)p;
        rc |= dev_upload(m, cen2.data(), sizeof(double) * cen2.size(), &p); pt.cen2 = (const double2*)p;
        // the kernels index m.mx stencil rows: keep at least one (all -1)
        if (mx == 0) {
            rc |= dev_upload(m, ex_dummy.data(), sizeof(int) * N, &p);
            m->d.extra = (const int*)p;
            m->d.mx = 1;
        }
        if (rc) {
            zj_mesh_destroy(m);
            return ZJ_ENOMEM;
        }
        m->n_templates = ph.ntpl;
    }
    {
        void* q = nullptr;
        rc = dev_upload(m, nullptr, sizeof(double2) * (size_t)mf * N, &q);
        m->slot_n = (double2*)q;
        rc |= dev_upload(m, nullptr, sizeof(double) * (size_t)mf * N, &q);
        m->slot_r = (double*)q;
        if (rc) {
            zj_mesh_destroy(m);
            return ZJ_ENOMEM;
        }
        hipLaunchKernelGGL(k_slot_dt_geo, dim3(nblocks(N)), dim3(BLOCK), 0, 0, m->d, m->slot_n,
                           m->slot_r);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            zj_mesh_destroy(m);
            return ZJ_EHIP;
        }
    }
    if (d->patch_order == 0 && m->pt.expl && cluster_order.size() == (size_t)N) {
        if (int rrc = build_renumbered(d, cluster_order, m)) {
            zj_mesh_destroy(m);
            return rrc;
        }
    }
    *out = m;
    return ZJ_OK;
}

int zj_mesh_destroy(zj_mesh* m) {
    if (!m) return ZJ_OK;
    if (m->ren) zj_mesh_destroy(m->ren);
    for (void* p : m->allocs) (void)hipFree(p);
    for (void* p : {(void*)m->fc, (void*)m->fs, (void*)m->cell_face})
        if (p) (void)hipFree(p);
    delete m;
    return ZJ_OK;
}

int zj_solver_create(zj_mesh* m, const zj_params* p, zj_solver** out) {
    if (!m || !p || !out) return ZJ_EINVAL;
    if (p->dtype != ZJ_F64 && p->dtype != ZJ_F32) return ZJ_EINVAL;
    if (p->limiter < 0 || p->limiter > 4 || p->riemann < 0 || p->riemann > 2 ||
        p->recon < 0 || p->recon > 3 || p->order < 1 || p->order > 2)
        return ZJ_EINVAL;
    zj_solver* s = new (std::nothrow) zj_solver();
    if (!s) return ZJ_ENOMEM;
    s->m = m;
    s->p = *p;
    s->o = Opts{p->g, p->cfl, p->eps_wet, p->dt_max, p->manning_n,
                p->order, p->limiter, p->riemann, p->recon, 0, -1.0, 0};
    const size_t esz = p->dtype == ZJ_F64 ? 8 : 4;
    const int N = m->d.N;
    s->mass_blocks = std::min(1024, nblocks(N));
    if (hipMalloc(&s->Q1, esz * 4 * N) != hipSuccess ||
        hipMalloc((void**)&s->dt_red, 8) != hipSuccess ||
        hipMalloc((void**)&s->dt_part, sizeof(double) * (size_t)m->pt.npatch * (SBLOCK / 64)) !=
            hipSuccess ||
        hipMalloc((void**)&s->troubled, 8) != hipSuccess ||
        hipMalloc((void**)&s->dt_tmp, 8) != hipSuccess ||
        hipMalloc((void**)&s->mass_part, sizeof(double2) * s->mass_blocks) != hipSuccess) {
        zj_solver_destroy(s);
        return ZJ_ENOMEM;
    }
    const int grc = esz == 8 ? grant_stage_lds<double>(stage_lds_bytes(m->pt, 8))
                             : grant_stage_lds<float>(stage_lds_bytes(m->pt, 4));
    if (grc != ZJ_OK) {
        zj_solver_destroy(s);
        return grc;
    }
    if (m->ren) {
        int rc = zj_solver_create(m->ren, p, &s->ren);
        if (!rc && hipMalloc(&s->Qp, esz * 4 * N) != hipSuccess) rc = ZJ_ENOMEM;
        if (rc) {
            zj_solver_destroy(s);
            return rc;
        }
    }
    *out = s;
    return ZJ_OK;
}

int zj_solver_destroy(zj_solver* s) {
    if (!s) return ZJ_OK;
    (void)hipFree(s->Q1);
    (void)hipFree(s->dt_red);
    (void)hipFree(s->dt_part);
    (void)hipFree(s->troubled);
    (void)hipFree(s->dt_tmp);
    (void)hipFree(s->mass_part);
    if (s->Qc) (void)hipFree(s->Qc);
    if (s->mood_mask) (void)hipFree(s->mood_mask);
    if (s->mood_cnt) (void)hipFree(s->mood_cnt);
    if (s->mood_stats) (void)hipFree(s->mood_stats);
    if (s->Qfrz) (void)hipFree(s->Qfrz);
    if (s->ren) zj_solver_destroy(s->ren);
    if (s->Qp) (void)hipFree(s->Qp);
    delete s;
    return ZJ_OK;
}

}  // extern "C"

namespace {

// blocks of the dt fold: one per 4096 partials, at most 64 (each folds its
// share and takes the minimum with one atomicMin -- order-independent, so
// the result does not depend on the block count)
static inline int reduce_blocks(int n) { return std::max(1, std::min(64, (n + 4095) / 4096)); }

// dt of Q into s->dt_red (already reset): block minima, then the fold.
// dt_part holds npatch * 4 >= N / 64 slots, more than the N / 256 blocks.
template <typename T>
void launch_dt(zj_solver* s, const T* Q, hipStream_t st) {
    const DevMesh& m = s->m->d;
    const int nb = (m.N + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_dt<T>, dim3(nb), dim3(BLOCK), 0, st, m, s->o, Q,
                       (const double2*)s->m->slot_n, (const double*)s->m->slot_r, s->dt_part);
    hipLaunchKernelGGL(k_reduce_part, dim3(reduce_blocks(nb)), dim3(1024), 0, st,
                       (const double*)s->dt_part, nb, s->dt_red, (const int*)nullptr);
}

template <typename T>
void launch_stage(zj_solver* s, const T* Qin, const T* Q0, T* Qout, const double* dt, T* aux,
                  int mode, int final_stage, int write_b, int use_slopes, int* troubled,
                  unsigned long long* dt_red, hipStream_t st,
                  const uint8_t* o1_mask = nullptr, uint8_t* bad_out = nullptr,
                  const int* gate = nullptr, ReconOut<T> ro = ReconOut<T>{},
                  const EnsArgs* ens = nullptr) {
    const zj_mesh* M = s->m;
    // the aux-only pass stores no limited slopes: its LDS ends before them
    const size_t lds = mode == 3 ? stage_lds_bytes(M->pt, sizeof(T)) -
                                       (size_t)M->pt.MS * 8 * sizeof(T)
                                 : stage_lds_bytes(M->pt, sizeof(T));
    const bool fixed = (s->o.lim == LIM_BJ || s->o.lim == LIM_MINMOD) && s->o.rs == RS_HLL &&
                       s->o.recon == RECON_CONS && !s->o.stage_aux;
    const bool tri = M->d.dim == 2 && M->d.mf == 3;
    const bool spec = M->d.dim == 2 && M->d.mf == 4 && fixed;
    const bool spec_tri = tri && fixed;
    const bool wetdry_tri = tri && s->o.lim == LIM_VANLEER && s->o.rs == RS_HLL &&
                            s->o.recon == RECON_ETA;
    auto kern = ens ? (spec_tri ? k_stage<T, 3, 0, 1> : tri ? k_stage<T, 2, 0, 1> : k_stage<T, 0, 0, 1>)
                : mode == 4  ? k_stage<T, 0, 2>
                : mode == 3  ? (spec ? k_stage<T, 1, 1> : spec_tri ? k_stage<T, 3, 1> : k_stage<T, 0, 1>)
                : spec       ? k_stage<T, 1>
                : spec_tri   ? k_stage<T, 3>
                : wetdry_tri ? k_stage<T, 4>
                : tri        ? k_stage<T, 2>
                             : k_stage<T, 0>;
    hipLaunchKernelGGL(kern, dim3(M->pt.npatch), dim3(SBLOCK),
                       lds, st, M->d, M->pt, s->o, Qin, Q0, Qout, dt, aux, mode, final_stage,
                       write_b, use_slopes, troubled, dt_red ? s->dt_part : (double*)nullptr,
                       o1_mask, bad_out, gate, ro, ens ? *ens : EnsArgs{});
    if (dt_red && !s->defer_dt)
        hipLaunchKernelGGL(k_reduce_part, dim3(reduce_blocks(M->pt.npatch * (SBLOCK / 64))),
                           dim3(1024), 0, st, (const double*)s->dt_part,
                           M->pt.npatch * (SBLOCK / 64), dt_red, gate);
}

template <typename T>
int step_impl(zj_solver* s, const double* dt, const T* Q, T* Qout, int* troubled,
              unsigned long long* dt_red, hipStream_t st, T* aux_of_Q = nullptr,
              bool copy_b = true, bool copy_b_out = true) {
    const DevMesh& m = s->m->d;
    T* Q1 = (T*)s->Q1;
    const int write_b = s->o.rs == RS_RUSANOV_NC ? 1 : 0;
    // solvers that never change b leave the stage buffer's b row alone: it
    // is copied once per call (once per march: copy_b on the first step),
    // and into an output buffer unless it already holds it (copy_b_out)
    if (!write_b && copy_b)
        HIPCHK(hipMemcpyAsync(Q1, Q, sizeof(T) * m.N, hipMemcpyDeviceToDevice, st));
    if (!write_b && Qout != Q && copy_b_out)
        HIPCHK(hipMemcpyAsync(Qout, Q, sizeof(T) * m.N, hipMemcpyDeviceToDevice, st));
    if (s->o.order < 2) {
        launch_stage<T>(s, Q, (const T*)nullptr, Q1, dt, (T*)nullptr, 0, 1, write_b, 0, troubled,
                        dt_red, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(Qout, Q1, sizeof(T) * 4 * m.N, hipMemcpyDeviceToDevice, st));
        return ZJ_OK;
    }
    // the aux rows are gradients of the conservative state: the slope pass
    // gives them when it reconstructs conservative variables (cons, xz)
    const bool fuse_aux = aux_of_Q && (s->o.recon == RECON_CONS || s->o.recon == RECON_XZ);
    launch_stage<T>(s, Q, (const T*)nullptr, Q1, dt, fuse_aux ? aux_of_Q : (T*)nullptr, 0, 0,
                    write_b, 1, (int*)nullptr, (unsigned long long*)nullptr, st);
    if (!s->mood) {
        // with the frozen stage aux, stage 2 gathers the neighbours' step-start
        // depth from Q: an in-place step writes it to scratch first
        const bool stage_scratch = s->o.stage_aux && (const void*)Qout == (const void*)Q;
        T* Q2 = stage_scratch ? (T*)s->Qfrz : Qout;
        if (stage_scratch && !Q2) return ZJ_EINVAL;
        launch_stage<T>(s, (const T*)Q1, Q, Q2, dt, (T*)nullptr, 1, 1, write_b, 1, troubled,
                        dt_red, st);
        if (stage_scratch)
            HIPCHK(hipMemcpyAsync(Qout + (write_b ? 0 : m.N), Q2 + (write_b ? 0 : m.N),
                                  sizeof(T) * (write_b ? 4 : 3) * m.N, hipMemcpyDeviceToDevice,
                                  st));
        HIPCHK(hipGetLastError());
        return ZJ_OK;
    }
    // A-posteriori MOOD (solver_jax.py:1055-1071): the candidate goes to the
    // output (to Qc when the step runs in place: the redo needs Q) with its
    // troubled mask; a redo with the flagged cells at first order runs on
    // the device only when the candidate had troubled cells (the redo
    // kernels read the count and return at once when it is 0).
    const bool in_place = (const void*)Qout == (const void*)Q;
    T* Qc = in_place ? (T*)s->Qc : Qout;
    if (!s->defer_dt)
        hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, (unsigned long long*)nullptr,
                           s->mood_cnt);
    launch_stage<T>(s, (const T*)Q1, Q, Qc, dt, (T*)nullptr, 1, 1, write_b, 1, s->mood_cnt, dt_red,
                    st, (const uint8_t*)nullptr, s->mood_mask);
    if (dt_red && !s->defer_dt)
        hipLaunchKernelGGL(k_dt_reset_if, dim3(1), dim3(1), 0, st, (const int*)s->mood_cnt, dt_red);
    launch_stage<T>(s, Q, (const T*)nullptr, Q1, dt, (T*)nullptr, 0, 0, write_b, 1, (int*)nullptr,
                    (unsigned long long*)nullptr, st, s->mood_mask, (uint8_t*)nullptr,
                    s->mood_cnt);
    launch_stage<T>(s, (const T*)Q1, Q, Qc, dt, (T*)nullptr, 1, 1, write_b, 1, troubled, dt_red, st,
                    s->mood_mask, (uint8_t*)nullptr, s->mood_cnt);
    if (!s->defer_dt)
        hipLaunchKernelGGL(k_mood_finish, dim3(1), dim3(1), 0, st, (const int*)s->mood_cnt,
                           s->mood_stats);
    if (in_place)
        HIPCHK(hipMemcpyAsync(Qout + (write_b ? 0 : m.N), Qc + (write_b ? 0 : m.N),
                              sizeof(T) * (write_b ? 4 : 3) * m.N, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

template <typename T>
int flux_impl(zj_solver* s, const T* Q, T* dQ, hipStream_t st) {
    launch_stage<T>(s, Q, (const T*)nullptr, dQ, (const double*)nullptr, (T*)nullptr, 2, 0, 1,
                    s->o.order >= 2 ? 1 : 0, (int*)nullptr, (unsigned long long*)nullptr, st);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // namespace

extern "C" {

int zj_flux_operator(zj_solver* s, const void* Q, void* dQ, void* stream) {
    if (!s || !Q || !dQ || Q == dQ) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    return s->p.dtype == ZJ_F64 ? flux_impl(s, (const double*)Q, (double*)dQ, st)
                                : flux_impl(s, (const float*)Q, (float*)dQ, st);
}

int zj_compute_dt(zj_solver* s, const void* Q, double t_remaining, double* dt_out, void* stream) {
    if (!s || !Q || !dt_out) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const DevMesh& m = s->m->d;
    hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, s->dt_red, (int*)nullptr);
    if (s->p.dtype == ZJ_F64)
        launch_dt<double>(s, (const double*)Q, st);
    else
        launch_dt<float>(s, (const float*)Q, st);
    // t_remaining is applied through a host scalar: finalize with t = 0 and
    // t_end = t_remaining.
    if (t_remaining >= 0.0) {
        HIPCHK(hipMemsetAsync(s->dt_tmp, 0, 8, st));
        hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, st, s->dt_red, dt_out,
                           (const double*)s->dt_tmp, t_remaining, s->p.dt_max);
    } else {
        hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, st, s->dt_red, dt_out,
                           (const double*)nullptr, -1.0, s->p.dt_max);
    }
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_step(zj_solver* s, const double* dt, const void* Q, void* Qout, int32_t* troubled_out,
            void* stream) {
    if (!s || !dt || !Q || !Qout) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (troubled_out)
        hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, s->dt_red, troubled_out);
    return s->p.dtype == ZJ_F64
               ? step_impl(s, dt, (const double*)Q, (double*)Qout, troubled_out,
                           (unsigned long long*)nullptr, st)
               : step_impl(s, dt, (const float*)Q, (float*)Qout, troubled_out,
                           (unsigned long long*)nullptr, st);
}

int zj_solver_set_mood(zj_solver* s, int32_t on) {
    if (!s) return ZJ_EINVAL;
    if (s->ren)
        if (int rc = zj_solver_set_mood(s->ren, on)) return rc;
    if (on && !s->Qc) {
        const size_t N = s->m->d.N;
        const size_t esz = s->p.dtype == ZJ_F64 ? 8 : 4;
        // all four buffers or none: a partial set is freed again
        void *qc = nullptr, *mask = nullptr, *cnt = nullptr, *stats = nullptr;
        if (hipMalloc(&qc, esz * 4 * N) != hipSuccess || hipMalloc(&mask, N) != hipSuccess ||
            hipMalloc(&cnt, sizeof(int)) != hipSuccess ||
            hipMalloc(&stats, 2 * sizeof(int)) != hipSuccess ||
            hipMemset(stats, 0, 2 * sizeof(int)) != hipSuccess) {
            for (void* q : {qc, mask, cnt, stats})
                if (q) (void)hipFree(q);
            return ZJ_ENOMEM;
        }
        s->Qc = qc;
        s->mood_mask = (uint8_t*)mask;
        s->mood_cnt = (int*)cnt;
        s->mood_stats = (int*)stats;
    }
    s->mood = on ? 1 : 0;
    return ZJ_OK;
}

int zj_solver_set_positivity(zj_solver* s, int32_t method, double front_tol) {
    if (!s || method < ZJ_POSITIVITY_NONE || method > ZJ_POSITIVITY_ZHANG_SHU ||
        !(front_tol <= 1.0) || front_tol != front_tol)
        return ZJ_EINVAL;
    s->o.pos = method;
    s->o.front_tol = front_tol > 0.0 ? front_tol : -1.0;
    if (s->ren) return zj_solver_set_positivity(s->ren, method, front_tol);
    return ZJ_OK;
}

int zj_solver_set_stage_aux(zj_solver* s, int32_t mode) {
    if (!s || mode < ZJ_STAGE_AUX_REFRESH || mode > ZJ_STAGE_AUX_FROZEN) return ZJ_EINVAL;
    if (s->ren)
        if (int rc = zj_solver_set_stage_aux(s->ren, mode)) return rc;
    if (mode == ZJ_STAGE_AUX_FROZEN && !s->Qfrz) {
        const size_t esz = s->p.dtype == ZJ_F64 ? 8 : 4;
        if (hipMalloc(&s->Qfrz, esz * 4 * (size_t)s->m->d.N) != hipSuccess) {
            s->Qfrz = nullptr;
            return ZJ_ENOMEM;
        }
    }
    s->o.stage_aux = mode;
    return ZJ_OK;
}

int zj_solver_mood_stats(zj_solver* s, int32_t* out2) {
    if (!s || !out2) return ZJ_EINVAL;
    out2[0] = out2[1] = 0;
    if (s->mood_stats) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(out2, s->mood_stats, 2 * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (s->ren) {
        // the steps the march ran on the renumbered twin
        int32_t r[2];
        if (int rc = zj_solver_mood_stats(s->ren, r)) return rc;
        out2[0] += r[0];
        out2[1] += r[1];
    }
    return ZJ_OK;
}

int zj_update_aux(zj_solver* s, const void* Q, void* Qaux, void* stream) {
    if (!s || !Q || !Qaux) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const int write_b = s->o.rs == RS_RUSANOV_NC ? 1 : 0;
    // the patch kernel's gradient pass alone (mode 3): staged neighbours
    // instead of per-cell gathers, conservative variables for every
    // reconstruction
    if (s->p.dtype == ZJ_F64)
        launch_stage<double>(s, (const double*)Q, (const double*)nullptr, (double*)nullptr,
                             (const double*)nullptr, (double*)Qaux, 3, 0, write_b, 1,
                             (int*)nullptr, (unsigned long long*)nullptr, st);
    else
        launch_stage<float>(s, (const float*)Q, (const float*)nullptr, (float*)nullptr,
                            (const double*)nullptr, (float*)Qaux, 3, 0, write_b, 1,
                            (int*)nullptr, (unsigned long long*)nullptr, st);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_march(zj_solver* s, void* Q, void* Qaux, int32_t n_steps, double t_end, double* t_dev,
             double* dt_dev, int32_t* troubled_out, void* stream) {
    if (!s || !Q || !t_dev || !dt_dev || n_steps < 0) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const DevMesh& m = s->m->d;
    if (s->ren) {
        // the march on the renumbered twin: the state in its order on the way
        // in, back in the caller's order on the way out; the aux rows of the
        // final state from the caller's mesh
        const int nb = nblocks(4 * m.N);
        const int* ord = s->m->ren_order;
        if (s->p.dtype == ZJ_F64)
            hipLaunchKernelGGL(k_permute_rows<double>, dim3(nb), dim3(BLOCK), 0, st, m.N, 4, ord,
                               (const double*)Q, (double*)s->Qp, 0);
        else
            hipLaunchKernelGGL(k_permute_rows<float>, dim3(nb), dim3(BLOCK), 0, st, m.N, 4, ord,
                               (const float*)Q, (float*)s->Qp, 0);
        HIPCHK(hipGetLastError());
        int rc = zj_march(s->ren, s->Qp, nullptr, n_steps, t_end, t_dev, dt_dev, troubled_out,
                          stream);
        if (rc) return rc;
        if (s->p.dtype == ZJ_F64)
            hipLaunchKernelGGL(k_permute_rows<double>, dim3(nb), dim3(BLOCK), 0, st, m.N, 4, ord,
                               (const double*)s->Qp, (double*)Q, 1);
        else
            hipLaunchKernelGGL(k_permute_rows<float>, dim3(nb), dim3(BLOCK), 0, st, m.N, 4, ord,
                               (const float*)s->Qp, (float*)Q, 1);
        HIPCHK(hipGetLastError());
        if (Qaux && n_steps > 0) return zj_update_aux(s, Q, Qaux, stream);
        return ZJ_OK;
    }
    hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, s->dt_red, troubled_out);
    if (s->p.dtype == ZJ_F64)
        launch_dt<double>(s, (const double*)Q, st);
    else
        launch_dt<float>(s, (const float*)Q, st);
    HIPCHK(hipGetLastError());
    // post_step aux refresh: fused into the first slope pass of the next
    // step (same gather) when the reconstruction is conservative and the
    // step is order 2; the state after the last step gets its own pass.
    const bool fuse = Qaux && s->o.order >= 2 &&
                      (s->o.recon == RECON_CONS || s->o.recon == RECON_XZ);
    // with MOOD the steps alternate between Q and the solver's Qc (the
    // candidate and its redo are written out of place), so no step copies
    // its candidate back; an odd step count ends with one copy into Q
    const bool ping = s->mood && s->o.order >= 2;
    const size_t esz = s->p.dtype == ZJ_F64 ? 8 : 4;
    const int write_b = s->o.rs == RS_RUSANOV_NC ? 1 : 0;
    if (ping && !write_b && n_steps > 0)
        HIPCHK(hipMemcpyAsync(s->Qc, Q, esz * m.N, hipMemcpyDeviceToDevice, st));
    void* buf[2] = {Q, ping ? s->Qc : Q};
    const bool mood2 = s->mood && s->o.order >= 2;
    const int npart = s->m->pt.npatch * (SBLOCK / 64);
    for (int k = 0; k < n_steps; ++k) {
        void* in = buf[k & 1];
        void* out = buf[(k + 1) & 1];
        hipLaunchKernelGGL(k_step_head, dim3(1), dim3(1), 0, st, s->dt_red, dt_dev, t_dev, t_end,
                           s->p.dt_max, k > 0 ? 1 : 0, mood2 ? s->mood_cnt : (int*)nullptr,
                           s->mood_stats, k > 0 ? 1 : 0);
        s->defer_dt = 1;
        int rc = s->p.dtype == ZJ_F64
                     ? step_impl(s, dt_dev, (const double*)in, (double*)out, troubled_out,
                                 s->dt_red, st, fuse && k > 0 ? (double*)Qaux : (double*)nullptr,
                                 k == 0, !ping)
                     : step_impl(s, dt_dev, (const float*)in, (float*)out, troubled_out,
                                 s->dt_red, st, fuse && k > 0 ? (float*)Qaux : (float*)nullptr,
                                 k == 0, !ping);
        s->defer_dt = 0;
        if (rc) return rc;
        // the step's dt minima: the last stage-2 launch's (the redo's, when
        // a MOOD redo ran), for the next step's head
        hipLaunchKernelGGL(k_reduce_part, dim3(reduce_blocks(npart)), dim3(1024), 0, st,
                           (const double*)s->dt_part, npart, s->dt_red, (const int*)nullptr);
        if (Qaux && !fuse) {
            rc = zj_update_aux(s, out, Qaux, stream);
            if (rc) return rc;
        }
    }
    // the MOOD bookkeeping of the last step
    if (mood2 && n_steps > 0)
        hipLaunchKernelGGL(k_mood_finish, dim3(1), dim3(1), 0, st, (const int*)s->mood_cnt,
                           s->mood_stats);
    if (ping && (n_steps & 1))
        HIPCHK(hipMemcpyAsync((char*)Q + (write_b ? 0 : esz * m.N),
                              (const char*)s->Qc + (write_b ? 0 : esz * m.N),
                              esz * (write_b ? 4 : 3) * m.N, hipMemcpyDeviceToDevice, st));
    if (n_steps > 0)
        hipLaunchKernelGGL(k_t_advance, dim3(1), dim3(1), 0, st, t_dev, (const double*)dt_dev);
    if (Qaux && fuse && n_steps > 0) {
        int rc = zj_update_aux(s, Q, Qaux, stream);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

// Average duration (ms) of each kernel of a step, timed with HIP events on
// the given stream over `reps` launches each: ms_out[0] stage 1 (slopes +
// fluxes -> stage state), [1] stage 2 (slopes + fluxes + RK average + source
// + troubled + dt of the new state), [2] standalone dt, [3] standalone aux
// refresh, [4] stage 1 with the fused aux refresh.  Q is read only; the
// outputs go to solver scratch.
}  // extern "C"

template <typename T>
static int time_kernels_impl(zj_solver* s, const T* q, const double* dt, int reps, float* ms_out,
                             hipStream_t st) {
    const DevMesh& m = s->m->d;
    const int nb = nblocks(m.N);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    T* out = nullptr;
    T* aux = nullptr;
    if (hipMalloc((void**)&out, sizeof(T) * 4 * m.N) != hipSuccess) return ZJ_ENOMEM;
    if (hipMalloc((void**)&aux, sizeof(T) * (3 * m.dim + 1) * m.N) != hipSuccess) {
        (void)hipFree(out);
        return ZJ_ENOMEM;
    }
    const int write_b = s->o.rs == RS_RUSANOV_NC ? 1 : 0;
    const int slopes = s->o.order >= 2 ? 1 : 0;
    for (int which = 0; which < 5; ++which) {
        for (int pass = 0; pass < 2; ++pass) {  // pass 0 warms up
            const int n = pass ? reps : 2;
            if (pass) HIPCHK(hipEventRecord(e0, st));
            for (int r = 0; r < n; ++r) {
                if (which == 0)
                    launch_stage<T>(s, q, (const T*)nullptr, out, dt, (T*)nullptr, 0, 0, write_b,
                                    slopes, (int*)nullptr, (unsigned long long*)nullptr, st);
                else if (which == 1)
                    launch_stage<T>(s, q, q, out, dt, (T*)nullptr, 1, 1, write_b, slopes,
                                    s->troubled, s->dt_red, st);
                else if (which == 2)
                    launch_dt<T>(s, q, st);
                else if (which == 3)
                    (void)zj_update_aux(s, q, aux, st);
                else
                    launch_stage<T>(s, q, (const T*)nullptr, out, dt, aux, 0, 0, write_b, slopes,
                                    (int*)nullptr, (unsigned long long*)nullptr, st);
            }
            if (pass) {
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                ms_out[which] = ms / (float)reps;
            }
        }
    }
    HIPCHK(hipGetLastError());
    (void)hipFree(out);
    (void)hipFree(aux);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ZJ_OK;
}

extern "C" int zj_time_kernels(zj_solver* s, const void* Q, const double* dt, int32_t reps,
                               float* ms_out, void* stream) {
    if (!s || !Q || !dt || !ms_out || reps <= 0) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    return s->p.dtype == ZJ_F64
               ? time_kernels_impl(s, (const double*)Q, dt, reps, ms_out, st)
               : time_kernels_impl(s, (const float*)Q, dt, reps, ms_out, st);
}

extern "C" {

int zj_mass(zj_solver* s, const void* Q, double* mass_out, void* stream) {
    if (!s || !Q || !mass_out) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const DevMesh& m = s->m->d;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_mass_partial<double>, dim3(s->mass_blocks), dim3(BLOCK), 0, st, m,
                           (const double*)Q, s->mass_part);
    else
        hipLaunchKernelGGL(k_mass_partial<float>, dim3(s->mass_blocks), dim3(BLOCK), 0, st, m,
                           (const float*)Q, s->mass_part);
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(1024), 0, st, (const double2*)s->mass_part,
                       s->mass_blocks, mass_out);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_imex_local(zj_solver* s, const double* dt, int32_t maxiter, const void* Qexp, void* Qout,
                  void* stream) {
    if (!s || !dt || !Qexp || !Qout || maxiter < 0) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const DevMesh& m = s->m->d;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_imex<double>, dim3(nblocks(m.N)), dim3(BLOCK), 0, st, m, s->o, dt,
                           maxiter, (const double*)Qexp, (double*)Qout);
    else
        hipLaunchKernelGGL(k_imex<float>, dim3(nblocks(m.N)), dim3(BLOCK), 0, st, m, s->o, dt,
                           maxiter, (const float*)Qexp, (float*)Qout);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // extern "C"

// ─────────────────────────────────────────────────────────────────────────
// Allocation-free dt: the kernels of zj_compute_dt with their per-block
// partials and reduction word in a file-scope device array instead of the
// solver's scratch.  Partials of up to DT_NOALLOC_MAX_BLOCKS blocks of BLOCK
// cells (16,777,216 cells), then the 8-byte reduction word: 512 KiB + 256 B
// that every GPU the library runs on holds until the process ends.  Every
// call on a device shares that device's array; calls that overlap on one
// device race.
// ─────────────────────────────────────────────────────────────────────────
namespace {
constexpr int DT_NOALLOC_MAX_BLOCKS = 65536;
constexpr int DT_NOALLOC_WORDS = DT_NOALLOC_MAX_BLOCKS + 32;
}  // namespace

static __device__ __align__(256) double scratch[DT_NOALLOC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[MAX_DEVICES];  // device address of `scratch`, per device

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the array's name is only a host-side
// stand-in: never pass it to a kernel or a copy).
int scratch_on_current_device(double** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return ZJ_EUNSUP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ZJ_EHIP;
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return ZJ_OK;
}
}  // namespace

extern "C" int zj_compute_dt_noalloc(zj_solver* s, const void* Q, double t_remaining,
                                     double* dt_out, void* stream) {
    if (!s || !Q || !dt_out) return ZJ_EINVAL;
    const DevMesh& m = s->m->d;
    const int nb = nblocks(m.N);
    if (nb > DT_NOALLOC_MAX_BLOCKS) {
        fprintf(stderr,
                "zoomy_hip: zj_compute_dt_noalloc: %d cells exceed the %d cells its static "
                "scratch holds; use zj_compute_dt\n",
                m.N, DT_NOALLOC_MAX_BLOCKS * BLOCK);
        return ZJ_EINVAL;
    }
    double* part = nullptr;
    const int rc = scratch_on_current_device(&part);
    if (rc) return rc;
    unsigned long long* red = (unsigned long long*)(part + DT_NOALLOC_MAX_BLOCKS);
    hipStream_t st = (hipStream_t)stream;
    // the reduction word keeps the last call's value: reset it first
    hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, red, (int*)nullptr);
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_dt<double>, dim3(nb), dim3(BLOCK), 0, st, m, s->o, (const double*)Q,
                           (const double2*)s->m->slot_n, (const double*)s->m->slot_r, part);
    else
        hipLaunchKernelGGL(k_dt<float>, dim3(nb), dim3(BLOCK), 0, st, m, s->o, (const float*)Q,
                           (const double2*)s->m->slot_n, (const double*)s->m->slot_r, part);
    hipLaunchKernelGGL(k_reduce_part, dim3(reduce_blocks(nb)), dim3(1024), 0, st,
                       (const double*)part, nb, red, (const int*)nullptr);
    hipLaunchKernelGGL(k_dt_finalize_rem, dim3(1), dim3(1), 0, st, red, dt_out, t_remaining,
                       s->p.dt_max);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

// ─────────────────────────────────────────────────────────────────────────
// The reference's block contract (solver2/blocks.py) as separate kernels on
// stored per-face arrays: dt_pass (per-face wave-speed bounds), reduce_dt,
// halo_bc (boundary ghost values), reconstruct (the reconstruction object),
// flux_pass (F, Dp, Dm per face) and gather_update (per-cell gather over the
// face-of-cell table, troubled flags).  The fused stage kernel computes the
// same step without storing the face arrays; these serve callers that drive
// the blocks themselves.  Faces are in the caller's reference-layout order
// (zj_mesh_set_faces).
// ─────────────────────────────────────────────────────────────────────────
namespace {

// Per-face bounds of the SWE eigenvalues {0, u.n - c, u.n, u.n + c} (the
// bed row's 0 included) over both cell states, 1/h from the aux row when
// given; a boundary face uses its owner on both sides (context.py:111-115).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_face_speeds(DevMesh m, Opts o, const T* __restrict__ Q,
                                                       const T* __restrict__ hinv_aux,
                                                       const int* __restrict__ fc,
                                                       const int* __restrict__ fs, int F,
                                                       const double2* __restrict__ sn,
                                                       T* __restrict__ lo, T* __restrict__ hi) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= F) return;
    const int N = m.N;
    const int A = fc[f];
    const int B = fc[F + f] >= 0 ? fc[F + f] : A;
    const double2 n = sn[(size_t)fs[f] * N + A];
    T l = T(0), h = T(0);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int c = side ? B : A;
        T q[4];
        load_q(Q, N, c, q);
        const T hv = hinv_aux ? hinv_aux[c] : hinv_of(q[1], T(o.eps));
        const T un = (q[2] * T(n.x) + q[3] * T(n.y)) * hv;
        const T cs = sqrt(T(o.g) * tmax(q[1], T(0)));
        l = tmin(l, un - cs);
        h = tmax(h, un + cs);
    }
    lo[f] = l;
    hi[f] = h;
}

// reduce_dt's per-face limit fac * 2 r_f / max(|lo|, |hi|) (fac = CFL /
// (d (2k + 1))), block minima to part[]; a wave-free face gives +inf, a
// non-finite speed 0 (so the dt guard raises).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_face_dt(const T* __restrict__ lo,
                                                   const T* __restrict__ hi,
                                                   const double* __restrict__ rf, int F,
                                                   double fac, double* __restrict__ part) {
    __shared__ double lds[BLOCK / 64];
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    double loc = INFINITY;
    if (f < F) {
        const double a = fmax(fabs((double)lo[f]), fabs((double)hi[f]));
        loc = fac * 2.0 * rf[f] / a;
        if (loc != loc || a != a) loc = 0.0;
    }
    double v = block_reduce(loc, [](double x, double y) { return fmin(x, y); }, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// halo_bc: the ghost of each boundary face's cell state (the periodic
// partner's state at a seam, the prescribed rows of a Dirichlet face).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_bc_values(DevMesh m, const T* __restrict__ Q,
                                                     const int* __restrict__ bcell,
                                                     const int* __restrict__ bslot,
                                                     const double2* __restrict__ sn,
                                                     T* __restrict__ bfv) {
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if (b >= m.nbf) return;
    const int c = bcell[b];
    const double2 n = sn[(size_t)bslot[b] * m.N + c];
    T q[4], g[4];
    load_q(Q, m.N, c, q);
    ghost_state(m, Q, b, q, T(n.x), T(n.y), g);
#pragma unroll
    for (int v = 0; v < 4; ++v) bfv[(size_t)v * m.nbf + b] = g[v];
}

// flux_pass: one evaluation per face of the stored face states; a boundary
// face's Q_R is recomputed as BC(Q_L) (the partner cell's state at a
// periodic seam), blocks.py:207-229.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_flux_pass(DevMesh m, Opts o, const T* __restrict__ Q,
                                                     const T* __restrict__ hinv_aux,
                                                     const T* __restrict__ QL,
                                                     const T* __restrict__ QR,
                                                     const int* __restrict__ fc,
                                                     const int* __restrict__ fs, int F,
                                                     const double2* __restrict__ sn,
                                                     T* __restrict__ Fo, T* __restrict__ Dpo,
                                                     T* __restrict__ Dmo) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= F) return;
    const int N = m.N;
    const int A = fc[f], B = fc[F + f], sA = fs[f];
    const double2 n = sn[(size_t)sA * N + A];
    T qL[4], qR[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) qL[v] = QL[(size_t)v * F + f];
    if (B >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) qR[v] = QR[(size_t)v * F + f];
    } else {
        ghost_state(m, Q, -1 - m.nbr[(size_t)sA * N + A], qL, T(n.x), T(n.y), qR);
    }
    const T eps = T(o.eps);
    const T hA = hinv_aux ? hinv_aux[A] : hinv_of(Q[(size_t)N + A], eps);
    const T hB = B < 0 ? hA : (hinv_aux ? hinv_aux[B] : hinv_of(Q[(size_t)N + B], eps));
    T Fx[4], Dp[4], Dm[4];
    face_split(o.rs, qL, qR, hA, hB, T(n.x), T(n.y), T(o.g), eps, Fx, Dp, Dm);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        Fo[(size_t)v * F + f] = Fx[v];
        Dpo[(size_t)v * F + f] = Dp[v];
        Dmo[(size_t)v * F + f] = Dm[v];
    }
}

// The runtime's numerical_flux / numerical_fluctuations slots: the
// solver's Riemann solver on n given face states.  The hinv of each side is
// the last aux row of the face's cell (qauxL / qauxR, (3 dim + 1, n)), or
// the hinv of the face state when no aux is given; normals (dim, n).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_face_numerics(
    Opts o, int dim, int n, const T* __restrict__ qL, const T* __restrict__ qR,
    const T* __restrict__ qauxL, const T* __restrict__ qauxR, const double* __restrict__ nrm,
    T* __restrict__ Fo, T* __restrict__ Dpo, T* __restrict__ Dmo) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= n) return;
    T a[4], b[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        a[v] = qL[(size_t)v * n + f];
        b[v] = qR[(size_t)v * n + f];
    }
    const T eps = T(o.eps);
    const size_t hrow = (size_t)(3 * dim) * n;
    const T hA = qauxL ? qauxL[hrow + f] : hinv_of(a[1], eps);
    const T hB = qauxR ? qauxR[hrow + f] : hinv_of(b[1], eps);
    const T nx = T(nrm[f]), ny = dim == 2 ? T(nrm[(size_t)n + f]) : T(0);
    T Fx[4], Dp[4], Dm[4];
    face_split(o.rs, a, b, hA, hB, nx, ny, T(o.g), eps, Fx, Dp, Dm);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (Fo) Fo[(size_t)v * n + f] = Fx[v];
        if (Dpo) Dpo[(size_t)v * n + f] = Dp[v];
        if (Dmo) Dmo[(size_t)v * n + f] = Dm[v];
    }
}

// gather_update: Q_cand = Q0 + a dt (-div(F + D) + cell_term + source) per
// cell over its faces (blocks.py:241-287); the cell-interior term is the
// given array or, from a limited gradient, -B(Q0) . grad with B[q_d, b] =
// g h (march.py:144-149); the source is the explicit Manning friction of Q0
// (1/h from the aux row when given).  troubled[c] = h < 0 or non-finite,
// and their count is added to *count.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gather_update(
    DevMesh m, Opts o, const T* __restrict__ Q0, const T* __restrict__ Fo,
    const T* __restrict__ Dpo, const T* __restrict__ Dmo, const int* __restrict__ cface, int F,
    const double* __restrict__ dtp, double a, const T* __restrict__ cell_term,
    const T* __restrict__ lim_grad, int source, const T* __restrict__ hinv_aux,
    T* __restrict__ Qc, uint8_t* __restrict__ troubled, int* __restrict__ count) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    int bad = 0;
    if (c < m.N) {
        const int N = m.N;
        Geo G;
        load_geo(m, c, G);
        T q[4];
        load_q(Q0, N, c, q);
        T dQ[4] = {0, 0, 0, 0};
        ZJ_FOR_FACES(k) {
            const int f2 = cface[(size_t)k * N + c];
            const int f = f2 >> 1;
            const T r = T(G.len[k] / G.vol);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T Fv = Fo[(size_t)v * F + f];
                dQ[v] -= ((f2 & 1) ? (-Fv + Dpo[(size_t)v * F + f])
                                   : (Fv + Dmo[(size_t)v * F + f])) * r;
            }
        }
        if (cell_term) {
#pragma unroll
            for (int v = 0; v < 4; ++v) dQ[v] += cell_term[(size_t)v * N + c];
        } else if (lim_grad) {
            const T gh = T(o.g) * q[1];
            dQ[2] -= gh * lim_grad[(size_t)0 * N + c];
            if (m.dim == 2) dQ[3] -= gh * lim_grad[(size_t)1 * N + c];
        }
        const T hv = hinv_aux ? hinv_aux[c] : hinv_of(q[1], T(o.eps));
        if (source && o.manning > 0.0) {
            const T u = q[2] * hv, w = q[3] * hv;
            const T sp = sqrt(u * u + w * w);
            const T kk = T(o.g * o.manning * o.manning) * cbrt(hv);
            dQ[2] += -kk * u * sp;
            dQ[3] += -kk * w * sp;
        }
        const T adt = T(a * *dtp);
        T out[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            out[v] = q[v] + adt * dQ[v];
            Qc[(size_t)v * N + c] = out[v];
        }
        bad = (out[1] < T(0)) || !isfinite(out[0]) || !isfinite(out[1]) ||
              !isfinite(out[2]) || !isfinite(out[3]);
        if (troubled) troubled[c] = (uint8_t)bad;
    }
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off, 64);
    if (count && (threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

static inline int fblocks(int n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace

extern "C" {

int zj_mesh_set_faces(zj_mesh* m, int32_t n_faces, const int32_t* face_cells,
                      const int32_t* face_slot) {
    if (!m || n_faces <= 0 || !face_cells || !face_slot) return ZJ_EINVAL;
    const int N = m->d.N, mf = m->d.mf, F = n_faces;
    std::vector<int32_t> nbr((size_t)mf * N);
    HIPCHK(hipMemcpy(nbr.data(), m->d.nbr, sizeof(int32_t) * nbr.size(), hipMemcpyDeviceToHost));
    // every (cell, slot) must be named by exactly one face side, consistent
    // with the mesh's neighbour table: the gathers then stay in bounds
    std::vector<int32_t> cf((size_t)mf * N, -1);
    for (int f = 0; f < F; ++f) {
        const int A = face_cells[f], B = face_cells[F + f];
        const int sA = face_slot[f], sB = face_slot[F + f];
        if (A < 0 || A >= N || sA < 0 || sA >= mf || B >= N) return ZJ_EINVAL;
        const int nA = nbr[(size_t)sA * N + A];
        if (B >= 0) {
            if (sB < 0 || sB >= mf || nA != B || nbr[(size_t)sB * N + B] != A) return ZJ_EINVAL;
        } else if (nA >= 0) {
            return ZJ_EINVAL;
        }
        int32_t& ca = cf[(size_t)sA * N + A];
        if (ca >= 0) return ZJ_EINVAL;
        ca = 2 * f;
        if (B >= 0) {
            int32_t& cb = cf[(size_t)sB * N + B];
            if (cb >= 0) return ZJ_EINVAL;
            cb = 2 * f + 1;
        }
    }
    for (size_t i = 0; i < cf.size(); ++i)
        if (cf[i] < 0) return ZJ_EINVAL;
    int *fc = nullptr, *fs = nullptr, *cfd = nullptr;
    if (hipMalloc((void**)&fc, sizeof(int) * 2 * F) != hipSuccess ||
        hipMalloc((void**)&fs, sizeof(int) * 2 * F) != hipSuccess ||
        hipMalloc((void**)&cfd, sizeof(int) * cf.size()) != hipSuccess) {
        for (void* q : {(void*)fc, (void*)fs, (void*)cfd})
            if (q) (void)hipFree(q);
        return ZJ_ENOMEM;
    }
    if (hipMemcpy(fc, face_cells, sizeof(int) * 2 * F, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fs, face_slot, sizeof(int) * 2 * F, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(cfd, cf.data(), sizeof(int) * cf.size(), hipMemcpyHostToDevice) != hipSuccess) {
        for (void* q : {(void*)fc, (void*)fs, (void*)cfd}) (void)hipFree(q);
        return ZJ_EHIP;
    }
    for (void* q : {(void*)m->fc, (void*)m->fs, (void*)m->cell_face})
        if (q) (void)hipFree(q);
    m->fc = fc;
    m->fs = fs;
    m->cell_face = cfd;
    m->nfaces = F;
    return ZJ_OK;
}

int zj_face_speeds(zj_solver* s, const void* Q, const void* hinv_aux, void* lo, void* hi,
                   void* stream) {
    if (!s || !Q || !lo || !hi) return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (!M->fc) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_face_speeds<double>, dim3(fblocks(M->nfaces)), dim3(BLOCK), 0, st,
                           M->d, s->o, (const double*)Q, (const double*)hinv_aux, M->fc, M->fs,
                           M->nfaces, (const double2*)M->slot_n, (double*)lo, (double*)hi);
    else
        hipLaunchKernelGGL(k_face_speeds<float>, dim3(fblocks(M->nfaces)), dim3(BLOCK), 0, st,
                           M->d, s->o, (const float*)Q, (const float*)hinv_aux, M->fc, M->fs,
                           M->nfaces, (const double2*)M->slot_n, (float*)lo, (float*)hi);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_reduce_dt(zj_solver* s, const void* lo, const void* hi, const double* inradius_f,
                 double cfl, int32_t dimension, int32_t degree, double dt_max, double clamp,
                 double* dt_out, void* stream) {
    if (!s || !lo || !hi || !inradius_f || !dt_out || dimension < 1 || degree < 0 ||
        !(cfl > 0.0))
        return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (!M->fc) return ZJ_EINVAL;
    const int F = M->nfaces, nb = fblocks(F);
    // dt_part holds npatch * 4 >= N / 64 slots; F <= 2 N + nbf faces need
    // F / 256 of them
    if (nb > M->pt.npatch * (SBLOCK / 64)) return ZJ_EUNSUP;
    hipStream_t st = (hipStream_t)stream;
    const double fac = cfl / ((double)dimension * (double)(2 * degree + 1));
    hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, s->dt_red, (int*)nullptr);
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_face_dt<double>, dim3(nb), dim3(BLOCK), 0, st, (const double*)lo,
                           (const double*)hi, inradius_f, F, fac, s->dt_part);
    else
        hipLaunchKernelGGL(k_face_dt<float>, dim3(nb), dim3(BLOCK), 0, st, (const float*)lo,
                           (const float*)hi, inradius_f, F, fac, s->dt_part);
    hipLaunchKernelGGL(k_reduce_part, dim3(reduce_blocks(nb)), dim3(1024), 0, st,
                       (const double*)s->dt_part, nb, s->dt_red, (const int*)nullptr);
    hipLaunchKernelGGL(k_dt_finalize_rem, dim3(1), dim3(1), 0, st, s->dt_red, dt_out, clamp,
                       dt_max);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_bc_values(zj_solver* s, const void* Q, void* bf_values, void* stream) {
    if (!s || !Q || !bf_values) return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (M->d.nbf == 0) return ZJ_OK;
    hipStream_t st = (hipStream_t)stream;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_bc_values<double>, dim3(fblocks(M->d.nbf)), dim3(BLOCK), 0, st, M->d,
                           (const double*)Q, M->bf_cell, M->bf_slot, (const double2*)M->slot_n,
                           (double*)bf_values);
    else
        hipLaunchKernelGGL(k_bc_values<float>, dim3(fblocks(M->d.nbf)), dim3(BLOCK), 0, st, M->d,
                           (const float*)Q, M->bf_cell, M->bf_slot, (const double2*)M->slot_n,
                           (float*)bf_values);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_reconstruct(zj_solver* s, const void* Q, const uint8_t* force_o1, int32_t order1,
                   void* QL, void* QR, void* lim_grad, void* stream) {
    if (!s || !Q || !QL || !QR) return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (!M->fc) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const int slopes = (s->o.order >= 2 && !order1) ? 1 : 0;
    if (s->p.dtype == ZJ_F64) {
        ReconOut<double> ro{(double*)QL, (double*)QR, slopes ? (double*)lim_grad : nullptr,
                            M->cell_face, M->nfaces};
        launch_stage<double>(s, (const double*)Q, (const double*)nullptr, (double*)nullptr,
                             (const double*)nullptr, (double*)nullptr, 4, 0, 0, slopes,
                             (int*)nullptr, (unsigned long long*)nullptr, st, force_o1,
                             (uint8_t*)nullptr, (const int*)nullptr, ro);
    } else {
        ReconOut<float> ro{(float*)QL, (float*)QR, slopes ? (float*)lim_grad : nullptr,
                           M->cell_face, M->nfaces};
        launch_stage<float>(s, (const float*)Q, (const float*)nullptr, (float*)nullptr,
                            (const double*)nullptr, (float*)nullptr, 4, 0, 0, slopes,
                            (int*)nullptr, (unsigned long long*)nullptr, st, force_o1,
                            (uint8_t*)nullptr, (const int*)nullptr, ro);
    }
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_flux_pass(zj_solver* s, const void* Q, const void* hinv_aux, const void* QL,
                 const void* QR, void* Fface, void* Dp, void* Dm, void* stream) {
    if (!s || !Q || !QL || !QR || !Fface || !Dp || !Dm) return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (!M->fc) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_flux_pass<double>, dim3(fblocks(M->nfaces)), dim3(BLOCK), 0, st, M->d,
                           s->o, (const double*)Q, (const double*)hinv_aux, (const double*)QL,
                           (const double*)QR, M->fc, M->fs, M->nfaces, (const double2*)M->slot_n,
                           (double*)Fface, (double*)Dp, (double*)Dm);
    else
        hipLaunchKernelGGL(k_flux_pass<float>, dim3(fblocks(M->nfaces)), dim3(BLOCK), 0, st, M->d,
                           s->o, (const float*)Q, (const float*)hinv_aux, (const float*)QL,
                           (const float*)QR, M->fc, M->fs, M->nfaces, (const double2*)M->slot_n,
                           (float*)Fface, (float*)Dp, (float*)Dm);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_face_numerics(zj_solver* s, int32_t n, const void* qL, const void* qR,
                     const void* qauxL, const void* qauxR, const double* normals, void* Fface,
                     void* Dp, void* Dm, void* stream) {
    if (!s || n < 0 || !qL || !qR || !normals || (!Fface && !Dp && !Dm) || (!Dp != !Dm))
        return ZJ_EINVAL;
    if (n == 0) return ZJ_OK;
    hipStream_t st = (hipStream_t)stream;
    const int dim = s->m->d.dim;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_face_numerics<double>, dim3(fblocks(n)), dim3(BLOCK), 0, st, s->o,
                           dim, n, (const double*)qL, (const double*)qR, (const double*)qauxL,
                           (const double*)qauxR, normals, (double*)Fface, (double*)Dp,
                           (double*)Dm);
    else
        hipLaunchKernelGGL(k_face_numerics<float>, dim3(fblocks(n)), dim3(BLOCK), 0, st, s->o,
                           dim, n, (const float*)qL, (const float*)qR, (const float*)qauxL,
                           (const float*)qauxR, normals, (float*)Fface, (float*)Dp, (float*)Dm);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_gather_update(zj_solver* s, const void* Q0, const void* Fface, const void* Dp,
                     const void* Dm, const double* dt, double a_stage, const void* cell_term,
                     const void* lim_grad, int32_t source, const void* hinv_aux, void* Qcand,
                     uint8_t* troubled, int32_t* count, void* stream) {
    if (!s || !Q0 || !Fface || !Dp || !Dm || !dt || !Qcand || Qcand == Q0) return ZJ_EINVAL;
    const zj_mesh* M = s->m;
    if (!M->fc) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_gather_update<double>, dim3(fblocks(M->d.N)), dim3(BLOCK), 0, st,
                           M->d, s->o, (const double*)Q0, (const double*)Fface, (const double*)Dp,
                           (const double*)Dm, M->cell_face, M->nfaces, dt, a_stage,
                           (const double*)cell_term, (const double*)lim_grad, source,
                           (const double*)hinv_aux, (double*)Qcand, troubled, count);
    else
        hipLaunchKernelGGL(k_gather_update<float>, dim3(fblocks(M->d.N)), dim3(BLOCK), 0, st,
                           M->d, s->o, (const float*)Q0, (const float*)Fface, (const float*)Dp,
                           (const float*)Dm, M->cell_face, M->nfaces, dt, a_stage,
                           (const float*)cell_term, (const float*)lim_grad, source,
                           (const float*)hinv_aux, (float*)Qcand, troubled, count);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_flux_operator_o1(zj_solver* s, const void* Q, const uint8_t* force_o1, void* dQ,
                        void* stream) {
    if (!s || !Q || !dQ || Q == dQ) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const int slopes = s->o.order >= 2 ? 1 : 0;
    if (s->p.dtype == ZJ_F64)
        launch_stage<double>(s, (const double*)Q, (const double*)nullptr, (double*)dQ,
                             (const double*)nullptr, (double*)nullptr, 2, 0, 1, slopes,
                             (int*)nullptr, (unsigned long long*)nullptr, st, force_o1);
    else
        launch_stage<float>(s, (const float*)Q, (const float*)nullptr, (float*)dQ,
                            (const double*)nullptr, (float*)nullptr, 2, 0, 1, slopes,
                            (int*)nullptr, (unsigned long long*)nullptr, st, force_o1);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // extern "C"

// ─────────────────────────────────────────────────────────────────────────
// Batched ensemble: independent instances concatenated into one mesh (cells
// of instance i are off[i] .. off[i+1]-1, no face joins two instances).  A
// step is the same two stage launches over the whole union's patch list --
// every patch is one work item, so uneven instances balance by themselves --
// with each cell reading its instance's dt, and the per-wave dt partials
// folded per instance.  The steps can be captured in a HIP graph.
// ─────────────────────────────────────────────────────────────────────────
namespace {

// dt of each instance's state: per cell min over faces, two partials per
// wave (waves of 64 consecutive cells: the block of 256 cells is one
// "patch" of the fold below).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_dt_ens(DevMesh m, Opts o, const T* __restrict__ Q,
                                                  const double2* __restrict__ sn,
                                                  const double* __restrict__ sr,
                                                  const int* __restrict__ inst,
                                                  double* __restrict__ part2) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    double loc = INFINITY;
    int my = -1;
    if (c < m.N) {
        const int N = m.N;
        T q[4];
        load_q(Q, N, c, q);
        const T hn = hinv_of(q[1], T(o.eps));
        ZJ_FOR_FACES(k) {
            const double2 n = sn[(size_t)k * N + c];
            const double a = (double)max_abs_eig(q, hn, T(n.x), T(n.y), T(o.g));
            double l = o.cfl * 2.0 * sr[(size_t)k * N + c] / ((double)m.dim * a);
            if (l != l) l = 0.0;
            loc = fmin(loc, l);
        }
        my = inst[c];
    }
    const int i0 = __shfl(my, 0, 64);
    double m0 = (my >= 0 && my == i0) ? loc : INFINITY;
    double m1 = (my >= 0 && my != i0) ? loc : INFINITY;
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmin(m0, __shfl_xor(m0, off, 64));
        m1 = fmin(m1, __shfl_xor(m1, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        const size_t w = (size_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
        part2[2 * w] = m0;
        part2[2 * w + 1] = m1;
    }
}

// Per-instance fold of the wave partials (one block per instance): the waves
// of patch p (P cells each) cover cells p P + 64 j .. + 63, j < 4; a wave's
// first partial belongs to the instance of its first cell, its second to an
// instance that starts inside it.  Deterministic (a minimum).
__global__ __launch_bounds__(BLOCK) void k_reduce_inst(const int* __restrict__ off,
                                                       const int* __restrict__ inst,
                                                       const double* __restrict__ part2,
                                                       int P, int N, double* __restrict__ red) {
    __shared__ double lds[BLOCK / 64];
    const int i = blockIdx.x;
    const int c_lo = off[i], c_hi = off[i + 1];  // [c_lo, c_hi)
    double v = INFINITY;
    if (c_hi > c_lo) {
        const int p0 = c_lo / P, p1 = (c_hi - 1) / P;
        const int nw = (p1 - p0 + 1) * (SBLOCK / 64);
        for (int t = threadIdx.x; t < nw; t += BLOCK) {
            const int p = p0 + t / (SBLOCK / 64), j = t % (SBLOCK / 64);
            const int fc = p * P + 64 * j;      // first cell of the wave
            if (64 * j >= P || fc >= N) continue;
            const int lc = min(fc + 63, min(p * P + P - 1, N - 1));  // its last cell
            if (lc < c_lo || fc >= c_hi) continue;
            const size_t w = (size_t)p * (SBLOCK / 64) + j;
            v = fmin(v, part2[2 * w + (inst[fc] == i ? 0 : 1)]);
        }
    }
    v = block_reduce(v, [](double a, double b) { return fmin(a, b); }, lds);
    if (threadIdx.x == 0) red[i] = v;
}

// Per instance: t += the previous dt (advance), then dt = min(fold, dt_max,
// t_end - t).
__global__ void k_ens_head(int n, const double* __restrict__ red, double* __restrict__ t,
                           double* __restrict__ dt, double t_end, double dt_max, int advance) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (advance) t[i] += dt[i];
    double d = red[i];
    if (dt_max > 0.0) d = fmin(d, dt_max);
    if (t_end >= 0.0) d = fmin(d, t_end - t[i]);
    dt[i] = d;
}

__global__ void k_ens_advance(int n, double* __restrict__ t, const double* __restrict__ dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] += dt[i];
}

// Compensated FP64 mass of each instance: every instance is cut into chunks
// of MCHUNK cells (chunk b holds cells [lo[b], hi[b]) of one instance), one
// block per chunk writes its compensated partial, and k_mass_inst_fold adds
// an instance's chunk partials in order -- a fixed decomposition and fold
// order, so the result is deterministic.
constexpr int MCHUNK = 8192;

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_mass_chunk(DevMesh m, const T* __restrict__ Q,
                                                      const int* __restrict__ lo,
                                                      const int* __restrict__ hi,
                                                      double2* __restrict__ part) {
    __shared__ double2 lds[BLOCK];
    const int b = blockIdx.x;
    double s = 0.0, comp = 0.0;
    for (int c = lo[b] + threadIdx.x; c < hi[b]; c += BLOCK) {
        double2 v[MAXF];
#pragma unroll
        for (int k = 0; k < MAXF; ++k)
            v[k] = k < m.mf ? m.verts[m.cv[k * m.N + c]] : make_double2(0.0, 0.0);
        neumaier(s, comp, (double)Q[m.N + c] * cell_volume(m, v));
    }
    lds[threadIdx.x] = make_double2(s, comp);
    fold_pairs<BLOCK>(lds);
    if (threadIdx.x == 0) part[b] = lds[0];
}

__global__ __launch_bounds__(BLOCK) void k_mass_inst_fold(int n, const int* __restrict__ chunk0,
                                                          const double2* __restrict__ part,
                                                          double* __restrict__ out) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s = 0.0, comp = 0.0;
    for (int b = chunk0[i]; b < chunk0[i + 1]; ++b) {
        const double2 p = part[b];
        neumaier(s, comp, p.x);
        comp += p.y;
    }
    out[i] = s + comp;
}

}  // namespace

struct zj_ensemble {
    zj_solver* s = nullptr;
    int n = 0;
    int* off = nullptr;        // (n + 1) cell offsets, device
    int* inst = nullptr;       // (N) instance of each cell, device
    int nchunk = 0;            // mass chunks (MCHUNK cells of one instance)
    int* chunk_lo = nullptr;   // (nchunk) first cell, device
    int* chunk_hi = nullptr;   // (nchunk) one past the last cell, device
    int* chunk0 = nullptr;     // (n + 1) first chunk of each instance, device
    double2* mpart = nullptr;  // (nchunk) compensated partials
    double* part2 = nullptr;   // two dt partials per wave
    double* red = nullptr;     // (n) per-instance dt fold
    hipStream_t cap = nullptr; // stream the graph is captured on
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <typename T>
void ens_dt(zj_ensemble* e, const T* Q, hipStream_t st) {
    const zj_mesh* M = e->s->m;
    const int nb = nblocks(M->d.N);
    hipLaunchKernelGGL(k_dt_ens<T>, dim3(nb), dim3(BLOCK), 0, st, M->d, e->s->o, Q,
                       (const double2*)M->slot_n, (const double*)M->slot_r,
                       (const int*)e->inst, e->part2);
    hipLaunchKernelGGL(k_reduce_inst, dim3(e->n), dim3(BLOCK), 0, st, (const int*)e->off,
                       (const int*)e->inst, (const double*)e->part2, BLOCK, M->d.N, e->red);
}

// n_steps SSP-RK2 steps of every instance, Q in place; the first head uses
// the dt fold of Q already in e->red.  Kernel launches only (capturable).
template <typename T>
void ens_steps(zj_ensemble* e, T* Q, int n_steps, double t_end, double* t, double* dt,
               int advance_first, hipStream_t st) {
    zj_solver* s = e->s;
    const zj_mesh* M = s->m;
    T* Q1 = (T*)s->Q1;
    const EnsArgs stage1{e->inst, dt, nullptr};
    const EnsArgs stage2{e->inst, dt, e->part2};
    const int hb = (e->n + 63) / 64;
    for (int k = 0; k < n_steps; ++k) {
        hipLaunchKernelGGL(k_ens_head, dim3(hb), dim3(64), 0, st, e->n, (const double*)e->red, t,
                           dt, t_end, s->p.dt_max, (k > 0 || advance_first) ? 1 : 0);
        launch_stage<T>(s, (const T*)Q, (const T*)nullptr, Q1, (const double*)nullptr,
                        (T*)nullptr, 0, 0, 0, 1, (int*)nullptr, (unsigned long long*)nullptr, st,
                        (const uint8_t*)nullptr, (uint8_t*)nullptr, (const int*)nullptr,
                        ReconOut<T>{}, &stage1);
        launch_stage<T>(s, (const T*)Q1, (const T*)Q, Q, (const double*)nullptr, (T*)nullptr, 1,
                        1, 0, 1, (int*)nullptr, (unsigned long long*)nullptr, st,
                        (const uint8_t*)nullptr, (uint8_t*)nullptr, (const int*)nullptr,
                        ReconOut<T>{}, &stage2);
        hipLaunchKernelGGL(k_reduce_inst, dim3(e->n), dim3(BLOCK), 0, st, (const int*)e->off,
                           (const int*)e->inst, (const double*)e->part2, M->pt.P, M->d.N, e->red);
    }
}

}  // namespace

extern "C" {

int zj_ensemble_create(zj_solver* s, int32_t n_inst, const int32_t* cell_offsets,
                       zj_ensemble** out) {
    if (!s || n_inst <= 0 || !cell_offsets || !out) return ZJ_EINVAL;
    const int N = s->m->d.N;
    // the scheme the ensemble kernels run: order-2 conservative MUSCL, no
    // MOOD, the step-start aux not frozen (in-place stage 2 reads only its
    // own cells of Q)
    if (s->o.order != 2 || s->mood || s->o.stage_aux || s->o.rs == RS_RUSANOV_NC)
        return ZJ_EUNSUP;
    // the per-wave dt partials need patches of consecutive cells
    if (s->m->pt.expl) return ZJ_EUNSUP;
    if (cell_offsets[0] != 0 || cell_offsets[n_inst] != N) return ZJ_EINVAL;
    std::vector<int32_t> inst(N);
    for (int i = 0; i < n_inst; ++i) {
        // every instance holds more than one wave of cells, so a wave of 64
        // consecutive cells meets at most two instances
        if (cell_offsets[i + 1] - cell_offsets[i] < 64) return ZJ_EINVAL;
        for (int c = cell_offsets[i]; c < cell_offsets[i + 1]; ++c) inst[c] = i;
    }
    // no face may join two instances
    {
        const int mf = s->m->d.mf;
        std::vector<int32_t> nbr((size_t)mf * N);
        HIPCHK(hipMemcpy(nbr.data(), s->m->d.nbr, sizeof(int32_t) * nbr.size(),
                         hipMemcpyDeviceToHost));
        for (int k = 0; k < mf; ++k)
            for (int c = 0; c < N; ++c) {
                const int nb = nbr[(size_t)k * N + c];
                if (nb >= 0 && inst[nb] != inst[c]) return ZJ_EINVAL;
            }
    }
    zj_ensemble* e = new (std::nothrow) zj_ensemble();
    if (!e) return ZJ_ENOMEM;
    e->s = s;
    e->n = n_inst;
    const size_t nwaves = (size_t)std::max(s->m->pt.npatch, nblocks(N)) * (SBLOCK / 64);
    std::vector<int32_t> clo, chi, c0(n_inst + 1);
    for (int i = 0; i < n_inst; ++i) {
        c0[i] = (int32_t)clo.size();
        for (int c = cell_offsets[i]; c < cell_offsets[i + 1]; c += MCHUNK) {
            clo.push_back(c);
            chi.push_back(std::min(c + MCHUNK, cell_offsets[i + 1]));
        }
    }
    c0[n_inst] = (int32_t)clo.size();
    e->nchunk = (int)clo.size();
    if (hipMalloc((void**)&e->off, sizeof(int) * (n_inst + 1)) != hipSuccess ||
        hipMalloc((void**)&e->inst, sizeof(int) * N) != hipSuccess ||
        hipMalloc((void**)&e->part2, sizeof(double) * 2 * nwaves) != hipSuccess ||
        hipMalloc((void**)&e->red, sizeof(double) * n_inst) != hipSuccess ||
        hipMalloc((void**)&e->chunk_lo, sizeof(int) * e->nchunk) != hipSuccess ||
        hipMalloc((void**)&e->chunk_hi, sizeof(int) * e->nchunk) != hipSuccess ||
        hipMalloc((void**)&e->chunk0, sizeof(int) * (n_inst + 1)) != hipSuccess ||
        hipMalloc((void**)&e->mpart, sizeof(double2) * e->nchunk) != hipSuccess ||
        hipStreamCreate(&e->cap) != hipSuccess) {
        zj_ensemble_destroy(e);
        return ZJ_ENOMEM;
    }
    if (hipMemcpy(e->off, cell_offsets, sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(e->inst, inst.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->chunk_lo, clo.data(), sizeof(int) * e->nchunk, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(e->chunk_hi, chi.data(), sizeof(int) * e->nchunk, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(e->chunk0, c0.data(), sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice) !=
            hipSuccess) {
        zj_ensemble_destroy(e);
        return ZJ_EHIP;
    }
    *out = e;
    return ZJ_OK;
}

int zj_ensemble_destroy(zj_ensemble* e) {
    if (!e) return ZJ_OK;
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->cap) (void)hipStreamDestroy(e->cap);
    for (void* p : {(void*)e->off, (void*)e->inst, (void*)e->part2, (void*)e->red,
                    (void*)e->chunk_lo, (void*)e->chunk_hi, (void*)e->chunk0, (void*)e->mpart})
        if (p) (void)hipFree(p);
    delete e;
    return ZJ_OK;
}

int zj_ensemble_dt(zj_ensemble* e, const void* Q, double* dt_out, void* stream) {
    if (!e || !Q || !dt_out) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (e->s->p.dtype == ZJ_F64) ens_dt(e, (const double*)Q, st);
    else ens_dt(e, (const float*)Q, st);
    HIPCHK(hipMemcpyAsync(dt_out, e->red, sizeof(double) * e->n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_ensemble_march(zj_ensemble* e, void* Q, int32_t n_steps, double t_end, double* t_dev,
                      double* dt_dev, void* stream) {
    if (!e || !Q || !t_dev || !dt_dev || n_steps < 0) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    zj_solver* s = e->s;
    const size_t esz = s->p.dtype == ZJ_F64 ? 8 : 4;
    // the stage buffer's bed row (stage 1 does not write it)
    HIPCHK(hipMemcpyAsync(s->Q1, Q, esz * s->m->d.N, hipMemcpyDeviceToDevice, st));
    if (s->p.dtype == ZJ_F64) {
        ens_dt(e, (const double*)Q, st);
        ens_steps(e, (double*)Q, n_steps, t_end, t_dev, dt_dev, 0, st);
    } else {
        ens_dt(e, (const float*)Q, st);
        ens_steps(e, (float*)Q, n_steps, t_end, t_dev, dt_dev, 0, st);
    }
    if (n_steps > 0)
        hipLaunchKernelGGL(k_ens_advance, dim3((e->n + 63) / 64), dim3(64), 0, st, e->n, t_dev,
                           (const double*)dt_dev);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_ensemble_capture(zj_ensemble* e, void* Q, int32_t n_steps, double t_end, double* t_dev,
                        double* dt_dev, void* stream) {
    if (!e || !Q || !t_dev || !dt_dev || n_steps <= 0) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    zj_solver* s = e->s;
    const size_t esz = s->p.dtype == ZJ_F64 ? 8 : 4;
    if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    if (e->graph) {
        (void)hipGraphDestroy(e->graph);
        e->graph = nullptr;
    }
    // outside the graph: the stage buffer's bed row and the dt of Q; the
    // graph holds kernel launches only (no memset or copy nodes)
    HIPCHK(hipMemcpyAsync(s->Q1, Q, esz * s->m->d.N, hipMemcpyDeviceToDevice, st));
    if (s->p.dtype == ZJ_F64) ens_dt(e, (const double*)Q, st);
    else ens_dt(e, (const float*)Q, st);
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
    // every replay continues the march: its first head advances t by the
    // last step's dt except on the first replay, which starts from the
    // state's own dt in e->red -- the graph starts with an advance-free
    // head and ends with the advance
    if (s->p.dtype == ZJ_F64) ens_steps(e, (double*)Q, n_steps, t_end, t_dev, dt_dev, 0, e->cap);
    else ens_steps(e, (float*)Q, n_steps, t_end, t_dev, dt_dev, 0, e->cap);
    hipLaunchKernelGGL(k_ens_advance, dim3((e->n + 63) / 64), dim3(64), 0, e->cap, e->n, t_dev,
                       (const double*)dt_dev);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->cap, &g);
    if (ce != hipSuccess || !g) {
        fprintf(stderr, "zoomy_hip: graph capture failed: %s\n", hipGetErrorString(ce));
        return ZJ_EHIP;
    }
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return ZJ_OK;
}

int zj_ensemble_replay(zj_ensemble* e, void* stream) {
    if (!e || !e->exec) return ZJ_EINVAL;
    HIPCHK(hipGraphLaunch(e->exec, (hipStream_t)stream));
    return ZJ_OK;
}

int zj_ensemble_mass(zj_ensemble* e, const void* Q, double* mass_out, void* stream) {
    if (!e || !Q || !mass_out) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const DevMesh& m = e->s->m->d;
    if (e->s->p.dtype == ZJ_F64)
        hipLaunchKernelGGL(k_mass_chunk<double>, dim3(e->nchunk), dim3(BLOCK), 0, st, m,
                           (const double*)Q, (const int*)e->chunk_lo, (const int*)e->chunk_hi,
                           e->mpart);
    else
        hipLaunchKernelGGL(k_mass_chunk<float>, dim3(e->nchunk), dim3(BLOCK), 0, st, m,
                           (const float*)Q, (const int*)e->chunk_lo, (const int*)e->chunk_hi,
                           e->mpart);
    hipLaunchKernelGGL(k_mass_inst_fold, dim3((e->n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st,
                       e->n, (const int*)e->chunk0, (const double2*)e->mpart, mass_out);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // extern "C"
