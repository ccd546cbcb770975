/* FP64 CPU oracle for the finite-volume time step -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, as the checker.  See zj_oracle.h for the reference
 * files it restates.  The arithmetic follows the reference's own data layout
 * (faces with owner/neighbour, scatter-style accumulation) rather than the
 * cell-gather layout the HIP kernels use, so that the two implementations
 * are independent.
 */
#include "zj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef ZJO_OMP
#include <omp.h>
#endif

/* The same source builds the serial checker (liboracle.so) and, with
 * -fopenmp -DZJO_OMP, the multi-threaded CPU baseline (liboracle_omp.so).
 * Only loops whose result does not depend on the thread count run in
 * parallel: per-cell and per-face loops writing their own entries, min/max
 * scatters (exact in any order) and integer counts; every floating-point
 * sum keeps its serial order, so both builds give identical bits. */
#ifdef ZJO_OMP
#define ZJO_PRAGMA(x) _Pragma(#x)
#define ZJO_PAR_FOR ZJO_PRAGMA(omp parallel for schedule(static))
#define ZJO_PAR_FOR_MAXMIN(a, b, n) \
    ZJO_PRAGMA(omp parallel for schedule(static) reduction(max : a[:n]) reduction(min : b[:n]))
#define ZJO_PAR_FOR_MIN(a, n) ZJO_PRAGMA(omp parallel for schedule(static) reduction(min : a[:n]))
#define ZJO_PAR_FOR_MIN_SCALAR(x) ZJO_PRAGMA(omp parallel for schedule(static) reduction(min : x))
#define ZJO_PAR_FOR_SUM_INT(x) ZJO_PRAGMA(omp parallel for schedule(static) reduction(+ : x))
#else
#define ZJO_PAR_FOR
#define ZJO_PAR_FOR_MAXMIN(a, b, n)
#define ZJO_PAR_FOR_MIN(a, n)
#define ZJO_PAR_FOR_MIN_SCALAR(x)
#define ZJO_PAR_FOR_SUM_INT(x)
#endif

int zjo_set_threads(int n)
{
#ifdef ZJO_OMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

#define NV 4 /* b, h, hu, hv */
#define IB 0
#define IH 1
#define IU 2
#define IVv 3

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
static double clip01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* KP-desingularised 1/h: sqrt(2) h / sqrt(h^4 + max(h, eps)^4)
 * (zoomy_jax/transformation/jax_runtime.py:446-448; solver_jax.py:291). */
double zjo_hinv(double h, double eps)
{
    double m = dmax(h, eps);
    double m2 = m * m, h2 = h * h;
    return sqrt(2.0) * h / sqrt(h2 * h2 + m2 * m2);
}

/* Boundary ghost state at the face (reference BC kernels from zoomy_core,
 * semantics stated at zoomy_jax/fvm/reconstruction_jax.py:494-510): Wall
 * mirrors the normal momentum, Extrapolation copies the inner state. */
void zjo_bc_ghost(int kind, const double *q, const double *n, double *out)
{
    for (int v = 0; v < NV; ++v) out[v] = q[v];
    if (kind == ZJO_BC_WALL) {
        double qn = q[IU] * n[0] + q[IVv] * n[1];
        out[IU] = q[IU] - 2.0 * qn * n[0];
        out[IVv] = q[IVv] - 2.0 * qn * n[1];
    }
}

/* Ghost state of boundary face b from the state q at the face: the kind's
 * ghost, then the prescribed rows of a Dirichlet face, whose value is the
 * ghost itself (tests/refs/bc_inflow holds to round-off with that; the
 * ghost-centre lift 2 v - q misses it by 0.28). */
static void ghost_at(const zjo_mesh *m, int b, const double *q,
                     const double *n, double *out)
{
    zjo_bc_ghost(m->bface_kind[b], q, n, out);
    if (m->bface_kind[b] != ZJO_BC_DIRICHLET || !m->bface_mask) return;
    for (int v = 0; v < NV; ++v)
        if (m->bface_mask[b] >> v & 1) {
            double val = m->bface_value[(size_t)v * m->n_bf + b];
            out[v] = val;
        }
}

/* LSQ gradient (zoomy_jax/mesh/mesh.py:74-120, reconstruction_jax.py:
 * 433-522): grad = sum_j w_j (u_j - u_c) + sum_g w_g (u_bf,g - u_c).
 * ubf == NULL is the extrapolation convention (ghost delta zero). */
void zjo_lsq_grad(const zjo_mesh *m, const double *u, const double *ubf,
                  double *grad)
{
    const int N = m->n_cells, S = m->max_st;
    ZJO_PAR_FOR
    for (int c = 0; c < N; ++c) {
        double gx = 0.0, gy = 0.0, uc = u[c];
        for (int j = 0; j < S; ++j) {
            int s = m->stencil[(size_t)c * S + j];
            if (s < 0) continue;
            double du = u[s] - uc;
            gx += m->lsq_w[((size_t)c * S + j) * 2 + 0] * du;
            gy += m->lsq_w[((size_t)c * S + j) * 2 + 1] * du;
        }
        for (int g = 0; g < 2; ++g) {
            int bf = m->gslot[(size_t)c * 2 + g];
            if (bf < 0) continue;
            double du = ubf ? (ubf[bf] - uc) : 0.0;
            gx += m->lsq_wg[((size_t)c * 2 + g) * 2 + 0] * du;
            gy += m->lsq_wg[((size_t)c * 2 + g) * 2 + 1] * du;
        }
        grad[c] = gx;
        grad[N + c] = gy;
    }
}

/* Per-face limiter candidate (reconstruction_jax.py:603-691).  minmod is the
 * same clipped ratio as Barth-Jespersen in the reference; the VK negative
 * branch divides by min(den, -eps) and therefore never limits; both are
 * reproduced as written.  Van Leer is not in the reference: its candidate
 * y(4-y)/4 (y = admissible increment / face increment, 1 for y >= 2) is the
 * function that reproduces 1-D van Leer MUSCL exactly on uniform meshes with
 * central LSQ gradients. */
static double lim_cand(int lim, double d, double dp, double dn, double ev2)
{
    const double eps = 1e-30;
    switch (lim) {
    case ZJO_LIM_NONE:
        return 1.0;
    case ZJO_LIM_BJ:
    case ZJO_LIM_MINMOD:
        if (d > eps) return clip01(dp / dmax(d, eps));
        if (d < -eps) return clip01(dn / dmin(d, -eps));
        return 1.0;
    case ZJO_LIM_VK: {
        if (d > eps) {
            double num = dp * dp + ev2 + 2.0 * d * dp;
            double den = dp * dp + 2.0 * d * d + d * dp + ev2;
            return clip01(num / dmax(den, eps));
        }
        if (d < -eps) {
            double num = dn * dn + ev2 + 2.0 * d * dn;
            double den = dn * dn + 2.0 * d * d + d * dn + ev2;
            return clip01(-num / dmin(den, -eps));
        }
        return 1.0;
    }
    case ZJO_LIM_VANLEER: {
        double y;
        if (d > eps) y = dp / d;
        else if (d < -eps) y = dn / d;
        else return 1.0;
        if (y <= 0.0) return 0.0;
        return y >= 2.0 ? 1.0 : y * (4.0 - y) * 0.25;
    }
    }
    return 1.0;
}

/* Ghost state per boundary face evaluated at the CELL state (flux operator
 * step 1, solver_jax.py:727-746); periodic faces take the partner cell. */
static void boundary_values(const zjo_mesh *m, const double *Q, double *bfv,
                            int *bf_bound_cell)
{
    const int N = m->n_cells, nbf = m->n_bf;
    for (int b = 0; b < nbf; ++b) {
        int f = m->bface_face[b];
        int c = m->bface_cell[b];
        double n[2] = {m->face_normals[f], m->face_normals[m->n_faces + f]};
        double q[NV], o[NV];
        int kind = m->bface_kind[b];
        int cc = c;
        if (kind == ZJO_BC_PERIODIC && m->bface_partner &&
            m->bface_partner[b] >= 0)
            cc = m->bface_cell[m->bface_partner[b]];
        for (int v = 0; v < NV; ++v) q[v] = Q[(size_t)v * N + cc];
        if (kind == ZJO_BC_PERIODIC) {
            for (int v = 0; v < NV; ++v) o[v] = q[v];
        } else {
            ghost_at(m, b, q, n, o);
        }
        for (int v = 0; v < NV; ++v) bfv[(size_t)v * nbf + b] = o[v];
        if (bf_bound_cell) bf_bound_cell[b] = cc;
    }
}

/* Limited slopes of each row of W (4, N) with ghost values bW (4, n_bf):
 * phi (4, N), grads (4, 2, N).  Bounds per reconstruction_jax.py:535-579
 * (interior face neighbours + boundary ghosts 2*bf - u). */
static void limited_slopes(const zjo_mesh *m, const zjo_params *p,
                           const double *W, const double *bW,
                           double *grads, double *phi)
{
    const int N = m->n_cells, F = m->n_faces, nbf = m->n_bf;
    double *umax = malloc(sizeof(double) * N);
    double *umin = malloc(sizeof(double) * N);
    for (int v = 0; v < NV; ++v) {
        const double *u = W + (size_t)v * N;
        double *gr = grads + (size_t)v * 2 * N;
        double *ph = phi + (size_t)v * N;
        zjo_lsq_grad(m, u, bW + (size_t)v * nbf, gr);
        ZJO_PAR_FOR
        for (int c = 0; c < N; ++c) { umax[c] = u[c]; umin[c] = u[c]; }
#ifdef ZJO_OMP
        if (m->cell_faces) {
            /* the same maxima and minima gathered per cell */
            ZJO_PAR_FOR
            for (int c = 0; c < N; ++c)
                for (int k = 0; k < m->max_cf; ++k) {
                    int f = m->cell_faces[(size_t)c * m->max_cf + k];
                    if (f < 0) continue;
                    int A = m->face_cells[f], B = m->face_cells[F + f];
                    if (B < 0) continue;
                    int o = A == c ? B : A;
                    umax[c] = dmax(umax[c], u[o]); umin[c] = dmin(umin[c], u[o]);
                }
        } else
#endif
        ZJO_PAR_FOR_MAXMIN(umax, umin, N)
        for (int f = 0; f < F; ++f) {
            int A = m->face_cells[f], B = m->face_cells[F + f];
            if (B < 0) continue;
            umax[A] = dmax(umax[A], u[B]); umin[A] = dmin(umin[A], u[B]);
            umax[B] = dmax(umax[B], u[A]); umin[B] = dmin(umin[B], u[A]);
        }
        for (int b = 0; b < nbf; ++b) {
            /* the ghost sample sits one cell spacing away: 2 bf - u; a
             * periodic ghost is the partner cell itself */
            int c = m->bface_cell[b];
            double gv = m->bface_kind[b] == ZJO_BC_PERIODIC
                            ? bW[(size_t)v * nbf + b]
                            : 2.0 * bW[(size_t)v * nbf + b] - u[c];
            umax[c] = dmax(umax[c], gv); umin[c] = dmin(umin[c], gv);
        }
        ZJO_PAR_FOR
        for (int c = 0; c < N; ++c) ph[c] = 1.0;
        if (p->limiter == ZJO_LIM_NONE) continue;
#ifdef ZJO_OMP
        if (m->cell_faces) {
            /* the same face candidates, minimised per cell */
            ZJO_PAR_FOR
            for (int c = 0; c < N; ++c)
                for (int k = 0; k < m->max_cf; ++k) {
                    int f = m->cell_faces[(size_t)c * m->max_cf + k];
                    if (f < 0) continue;
                    double rx = m->face_centers[f] - m->cell_centers[c];
                    double ry = m->face_centers[F + f] - m->cell_centers[N + c];
                    double d = gr[c] * rx + gr[N + c] * ry;
                    double ev2 = p->limiter == ZJO_LIM_VK
                                     ? pow(m->cell_volumes[c], 2.0 / (m->dim > 1 ? m->dim : 1))
                                     : 0.0; /* read by VK only */
                    double cand = lim_cand(p->limiter, d, umax[c] - u[c],
                                           umin[c] - u[c], ev2);
                    if (cand < ph[c]) ph[c] = cand;
                }
        } else
#endif
        ZJO_PAR_FOR_MIN(ph, N)
        for (int f = 0; f < F; ++f) {
            int cells[2] = {m->face_cells[f], m->face_cells[F + f]};
            for (int side = 0; side < 2; ++side) {
                int c = cells[side];
                if (c < 0) continue;
                double rx = m->face_centers[f] - m->cell_centers[c];
                double ry = m->face_centers[F + f] - m->cell_centers[N + c];
                double d = gr[c] * rx + gr[N + c] * ry;
                double ev2 = p->limiter == ZJO_LIM_VK
                                     ? pow(m->cell_volumes[c], 2.0 / (m->dim > 1 ? m->dim : 1))
                                     : 0.0; /* read by VK only */
                double cand = lim_cand(p->limiter, d, umax[c] - u[c],
                                       umin[c] - u[c], ev2);
                if (cand < ph[c]) ph[c] = cand;
            }
        }
        ZJO_PAR_FOR
        for (int c = 0; c < N; ++c) ph[c] = clip01(ph[c]);
    }
    free(umax);
    free(umin);
}

/* Per-cell first-order demotion (MOOD force_o1, reconstruction_jax.py
 * :1211-1225): phi := 0 for the flagged cells, applied after the limiter
 * ties and before reconstruction.  NULL: none. */
static const unsigned char *g_force_o1 = NULL;

static void demote(int N, double *phi)
{
    if (!g_force_o1) return;
    for (int c = 0; c < N; ++c)
        if (g_force_o1[c])
            for (int v = 0; v < NV; ++v) phi[(size_t)v * N + c] = 0.0;
}

/* Xing-Zhang-Shu cell-mean positivity factor (reconstruction_jax.py
 * :840-850): with the face-midpoint depths h_f = hbar_c + s_c . (x_f - x_c)
 * over every face of cell c, theta_c = clip(hbar_c / max(hbar_c - min_f h_f,
 * eps), 0, 1) where min_f h_f < 0, else 1.  s: depth slope (2, N). */
static void pp_theta(const zjo_mesh *m, const double *hbar, const double *s,
                     double eps, double *theta)
{
    const int N = m->n_cells, F = m->n_faces;
    double *hmin = malloc(sizeof(double) * N);
    for (int c = 0; c < N; ++c) hmin[c] = INFINITY;
    ZJO_PAR_FOR_MIN(hmin, N)
    for (int f = 0; f < F; ++f) {
        double fx = m->face_centers[f], fy = m->face_centers[F + f];
        for (int side = 0; side < 2; ++side) {
            int c = m->face_cells[(size_t)side * F + f];
            if (c < 0) continue;
            double d = s[c] * (fx - m->cell_centers[c]) +
                       s[N + c] * (fy - m->cell_centers[N + c]);
            hmin[c] = dmin(hmin[c], hbar[c] + d);
        }
    }
    for (int c = 0; c < N; ++c) {
        double den = dmax(hbar[c] - hmin[c], eps);
        theta[c] = hmin[c] < 0.0 ? clip01(hbar[c] / den) : 1.0;
    }
    free(hmin);
}

static void face_values(const zjo_mesh *m, const double *W,
                        const double *grads, const double *phi, double *WL,
                        double *WR, const double *bW)
{
    const int N = m->n_cells, F = m->n_faces, nbf = m->n_bf;
    ZJO_PAR_FOR
    for (int f = 0; f < F; ++f) {
        int A = m->face_cells[f], B = m->face_cells[F + f];
        double fx = m->face_centers[f], fy = m->face_centers[F + f];
        for (int v = 0; v < NV; ++v) {
            const double *u = W + (size_t)v * N;
            const double *g = grads + (size_t)v * 2 * N;
            const double *ph = phi + (size_t)v * N;
            double rx = fx - m->cell_centers[A], ry = fy - m->cell_centers[N + A];
            WL[(size_t)v * F + f] = u[A] + ph[A] * (g[A] * rx + g[N + A] * ry);
            if (B >= 0) {
                double sx = fx - m->cell_centers[B], sy = fy - m->cell_centers[N + B];
                WR[(size_t)v * F + f] = u[B] + ph[B] * (g[B] * sx + g[N + B] * sy);
            }
        }
    }
    for (int b = 0; b < nbf; ++b) {
        int f = m->bface_face[b];
        for (int v = 0; v < NV; ++v)
            WR[(size_t)v * F + f] = bW[(size_t)v * nbf + b];
    }
}

/* Face reconstruction (LSQMUSCLReconstructionJAX.__call__ :398-429 and the
 * eta/wet-dry core EmittedWLSQMUSCLJAX._eta_core :1134-1292).  Boundary Q_R
 * is left at the ghost of the cell state; the flux operator replaces it by
 * BC(Q_L) as the reference does (solver_jax.py:834-858). */
void zjo_reconstruct(const zjo_mesh *m, const zjo_params *p, const double *Q,
                     double *QL, double *QR, double *lim_grad)
{
    const int N = m->n_cells, F = m->n_faces, nbf = m->n_bf;
    double *bfv = malloc(sizeof(double) * NV * (nbf > 0 ? nbf : 1));
    int *bcell = malloc(sizeof(int) * (nbf > 0 ? nbf : 1));
    boundary_values(m, Q, bfv, bcell);

    if (p->order < 2) {
        for (int f = 0; f < F; ++f) {
            int A = m->face_cells[f], B = m->face_cells[F + f];
            for (int v = 0; v < NV; ++v) {
                QL[(size_t)v * F + f] = Q[(size_t)v * N + A];
                if (B >= 0) QR[(size_t)v * F + f] = Q[(size_t)v * N + B];
            }
        }
        for (int b = 0; b < nbf; ++b)
            for (int v = 0; v < NV; ++v)
                QR[(size_t)v * F + m->bface_face[b]] = bfv[(size_t)v * nbf + b];
        if (lim_grad) memset(lim_grad, 0, sizeof(double) * NV * 2 * N);
        free(bfv);
        free(bcell);
        return;
    }

    double *grads = malloc(sizeof(double) * NV * 2 * N);
    double *phi = malloc(sizeof(double) * NV * N);
    if (p->recon != ZJO_RECON_ETA && p->recon != ZJO_RECON_ETA_HU) {
        limited_slopes(m, p, Q, bfv, grads, phi);
        if (p->recon == ZJO_RECON_XZ) {
            /* PositivityPreservingLSQMUSCLJAX._xing_zhang_scale_phi
             * (:924-937): theta of the limited depth slope scales the depth
             * and momentum limiters, never the bed's */
            double *s = malloc(sizeof(double) * 2 * N);
            double *th = malloc(sizeof(double) * N);
            for (int d = 0; d < 2; ++d)
                for (int c = 0; c < N; ++c)
                    s[(size_t)d * N + c] = grads[((size_t)1 * 2 + d) * N + c] * phi[N + c];
            pp_theta(m, Q + N, s, 1e-14, th);
            for (int c = 0; c < N; ++c)
                for (int v = 1; v < NV; ++v) phi[(size_t)v * N + c] *= th[c];
            free(s);
            free(th);
        }
        demote(N, phi);
        face_values(m, Q, grads, phi, QL, QR, bfv);
    } else {
        /* W = (b, eta = h + b, u = hinv q, v = hinv q_y); ghost W through the
         * same map with the ghost's own hinv.  ETA_HU keeps the conservative
         * momentum rows (b, eta, hu, hv) with the 1e-3 depth threshold of the
         * earlier EtaWellBalancedLSQMUSCLJAX (its replacement's docstring,
         * reconstruction_jax.py:995-1008): cells below it are demoted to
         * first order and face momentum below it is zeroed. */
        const int vel = p->recon == ZJO_RECON_ETA;
        const double eps_hu = 1e-3;
        double *W = malloc(sizeof(double) * NV * N);
        double *bW = malloc(sizeof(double) * NV * (nbf > 0 ? nbf : 1));
        for (int c = 0; c < N; ++c) {
            double b = Q[c], h = Q[N + c];
            double hi = zjo_hinv(h, p->eps_wet);
            W[c] = b;
            W[N + c] = h + b;
            W[2 * (size_t)N + c] = Q[2 * (size_t)N + c] * (vel ? hi : 1.0);
            W[3 * (size_t)N + c] = Q[3 * (size_t)N + c] * (vel ? hi : 1.0);
        }
        for (int k = 0; k < nbf; ++k) {
            double b = bfv[k], h = bfv[nbf + k];
            double hi = zjo_hinv(h, p->eps_wet);
            bW[k] = b;
            bW[nbf + k] = h + b;
            bW[2 * (size_t)nbf + k] = bfv[2 * (size_t)nbf + k] * (vel ? hi : 1.0);
            bW[3 * (size_t)nbf + k] = bfv[3 * (size_t)nbf + k] * (vel ? hi : 1.0);
        }
        limited_slopes(m, p, W, bW, grads, phi);
        /* tie phi_b and the velocity limiters to phi_eta (:1206-1209) */
        for (int c = 0; c < N; ++c) {
            double pe = phi[N + c];
            phi[c] = dmin(phi[c], pe);
            phi[2 * (size_t)N + c] = dmin(phi[2 * (size_t)N + c], pe);
            phi[3 * (size_t)N + c] = dmin(phi[3 * (size_t)N + c], pe);
        }
        /* front pre-detector and Zhang-Shu cap (:1211-1244), both from the
         * tied depth slope sh = phi_eta grad eta - phi_b grad b */
        unsigned char *front = NULL;
        if (p->front_tol > 0.0 || p->positivity == ZJO_POS_ZHANG_SHU) {
            double *sh = malloc(sizeof(double) * 2 * N);
            double *th = malloc(sizeof(double) * N);
            for (int d = 0; d < 2; ++d)
                for (int c = 0; c < N; ++c)
                    sh[(size_t)d * N + c] =
                        phi[N + c] * grads[((size_t)1 * 2 + d) * N + c] -
                        phi[c] * grads[(size_t)d * N + c];
            pp_theta(m, Q + N, sh, 1e-14, th);
            if (p->front_tol > 0.0) {
                front = calloc((size_t)N, 1);
                for (int c = 0; c < N; ++c) front[c] = th[c] < p->front_tol;
            }
            if (p->positivity == ZJO_POS_ZHANG_SHU) {
                /* eta slope := phi_b grad b + theta sh with phi_eta = 1;
                 * the velocity limiters are scaled by theta */
                for (int c = 0; c < N; ++c) {
                    for (int d = 0; d < 2; ++d)
                        grads[((size_t)1 * 2 + d) * N + c] =
                            phi[c] * grads[(size_t)d * N + c] + th[c] * sh[(size_t)d * N + c];
                    phi[N + c] = 1.0;
                    phi[2 * (size_t)N + c] *= th[c];
                    phi[3 * (size_t)N + c] *= th[c];
                }
            }
            free(sh);
            free(th);
        }
        if (front)
            for (int c = 0; c < N; ++c)
                if (front[c])
                    for (int v = 0; v < NV; ++v) phi[(size_t)v * N + c] = 0.0;
        free(front);
        if (!vel)
            for (int c = 0; c < N; ++c)
                if (Q[N + c] < eps_hu)
                    for (int v = 0; v < NV; ++v) phi[(size_t)v * N + c] = 0.0;
        demote(N, phi);
        double *WL = malloc(sizeof(double) * NV * F);
        double *WR = malloc(sizeof(double) * NV * F);
        face_values(m, W, grads, phi, WL, WR, bW);
        for (int f = 0; f < F; ++f) {
            int B = m->face_cells[F + f];
            for (int side = 0; side < 2; ++side) {
                double *Wx = side ? WR : WL;
                double *Qx = side ? QR : QL;
                if (side == 1 && B < 0) { /* boundary: ghost of cell state */
                    continue;
                }
                double b = Wx[f];
                double eta = dmax(Wx[(size_t)F + f], b); /* :1269-1272 */
                double h = eta - b;
                Qx[f] = b;
                Qx[(size_t)F + f] = h;
                for (int v = 2; v < NV; ++v) {
                    double w = Wx[(size_t)v * F + f];
                    Qx[(size_t)v * F + f] = vel ? w * h : (h < eps_hu ? 0.0 : w);
                }
            }
        }
        for (int b = 0; b < nbf; ++b)
            for (int v = 0; v < NV; ++v)
                QR[(size_t)v * F + m->bface_face[b]] = bfv[(size_t)v * nbf + b];
        if (lim_grad) {
            /* reconstruct_with_grad (:1313-1366): the limited W slopes
             * pushed to conservative rows through the inverse map at the
             * cell, dh = deta - db, d(u h) = h du + u dh (eta); the
             * conservative momentum rows of eta_hu stay as they are */
            for (int d = 0; d < 2; ++d)
                for (int c = 0; c < N; ++c) {
                    double s[NV];
                    for (int v = 0; v < NV; ++v)
                        s[v] = phi[(size_t)v * N + c] * grads[((size_t)v * 2 + d) * N + c];
                    double dh = s[1] - s[0];
                    if (vel) {
                        double hc = W[N + c] - W[c];
                        for (int v = 2; v < NV; ++v) s[v] = hc * s[v] + W[(size_t)v * N + c] * dh;
                    }
                    s[1] = dh;
                    for (int v = 0; v < NV; ++v) lim_grad[((size_t)v * 2 + d) * N + c] = s[v];
                }
            free(grads);
            free(phi);
            grads = phi = NULL;
        }
        free(WL);
        free(WR);
        free(W);
        free(bW);
    }
    if (lim_grad && grads) {
        for (int v = 0; v < NV; ++v)
            for (int d = 0; d < 2; ++d)
                for (int c = 0; c < N; ++c)
                    lim_grad[((size_t)v * 2 + d) * N + c] =
                        phi[(size_t)v * N + c] * grads[((size_t)v * 2 + d) * N + c];
    }
    free(grads);
    free(phi);
    free(bfv);
    free(bcell);
}

static double max_abs_eig(const zjo_params *p, const double *q, double hinv,
                          const double *n)
{
    double un = (q[IU] * n[0] + q[IVv] * n[1]) * hinv;
    return fabs(un) + sqrt(p->g * dmax(q[IH], 0.0));
}

static void phys_flux(const zjo_params *p, const double *q, double hinv,
                      const double *n, double *F)
{
    double un = (q[IU] * n[0] + q[IVv] * n[1]) * hinv;
    double pr = 0.5 * p->g * q[IH] * q[IH];
    F[IB] = 0.0;
    F[IH] = q[IU] * n[0] + q[IVv] * n[1];
    F[IU] = q[IU] * un + pr * n[0];
    F[IVv] = q[IVv] * un + pr * n[1];
}

/* Face numerics in the reference's (flux, fluctuation) split
 * (jax_runtime.py:780-806; update rule solver_jax.py:806-809):
 *   Rusanov: F = centred flux; Dm/Dp = (B_int dQ -/+ s_M I' dQ) / 2 with the
 *   path-conservative integral of B[q_d, b] = g h n_d over the straight path
 *   (nonconservative_flux.py:61-186); s_M takes 1/h of each face state
 *   itself (the KP formula), not the cell aux that the centred flux uses --
 *   at a periodic seam the ghost is the partner cell while the aux passed is
 *   the inner cell's (solver_jax.py:858-861), and tests/refs/bc_periodic
 *   holds to round-off only with the state's own 1/h.  The NC variant dissipates every row
 *   (I' = I); the positive/well-balanced variant dissipates h by the jump of
 *   eta = h + b and leaves b untouched (identified from tests/refs/wb_lake).
 *   HLL: hydrostatic reconstruction (Audusse et al. 2004) of both face states,
 *   HLL flux with Davis/Einfeldt speeds, and the pressure corrections
 *   g/2 (h^2 - h*^2) n as the fluctuations. */
void zjo_face_flux(const zjo_params *p, const double *qL, const double *qR,
                   double hinvA, double hinvB, const double *n, double *F,
                   double *Dp, double *Dm)
{
    const double g = p->g;
    if (p->riemann == ZJO_RIEMANN_HLL) {
        double bs = dmax(qL[IB], qR[IB]);
        double hLs = dmax(0.0, qL[IH] + qL[IB] - bs);
        double hRs = dmax(0.0, qR[IH] + qR[IB] - bs);
        double uL = qL[IU] * hinvA, vL = qL[IVv] * hinvA;
        double uR = qR[IU] * hinvB, vR = qR[IVv] * hinvB;
        double unL = uL * n[0] + vL * n[1], unR = uR * n[0] + vR * n[1];
        double cL = sqrt(g * hLs), cR = sqrt(g * hRs);
        double sL = dmin(unL - cL, unR - cR);
        double sR = dmax(unL + cL, unR + cR);
        double UL[3] = {hLs, hLs * uL, hLs * vL};
        double UR[3] = {hRs, hRs * uR, hRs * vR};
        double FL[3] = {hLs * unL, hLs * uL * unL + 0.5 * g * hLs * hLs * n[0],
                        hLs * vL * unL + 0.5 * g * hLs * hLs * n[1]};
        double FR[3] = {hRs * unR, hRs * uR * unR + 0.5 * g * hRs * hRs * n[0],
                        hRs * vR * unR + 0.5 * g * hRs * hRs * n[1]};
        double Fh[3];
        if (sL >= 0.0) {
            for (int k = 0; k < 3; ++k) Fh[k] = FL[k];
        } else if (sR <= 0.0) {
            for (int k = 0; k < 3; ++k) Fh[k] = FR[k];
        } else {
            double inv = 1.0 / (sR - sL);
            for (int k = 0; k < 3; ++k)
                Fh[k] = (sR * FL[k] - sL * FR[k] + sL * sR * (UR[k] - UL[k])) * inv;
        }
        F[IB] = 0.0;
        F[IH] = Fh[0];
        F[IU] = Fh[1];
        F[IVv] = Fh[2];
        double cA = 0.5 * g * (qL[IH] * qL[IH] - hLs * hLs);
        double cB = 0.5 * g * (qR[IH] * qR[IH] - hRs * hRs);
        Dm[IB] = Dm[IH] = 0.0;
        Dm[IU] = cA * n[0];
        Dm[IVv] = cA * n[1];
        Dp[IB] = Dp[IH] = 0.0;
        Dp[IU] = -cB * n[0];
        Dp[IVv] = -cB * n[1];
        return;
    }
    double FLx[NV], FRx[NV];
    phys_flux(p, qL, hinvA, n, FLx);
    phys_flux(p, qR, hinvB, n, FRx);
    double s = dmax(max_abs_eig(p, qL, zjo_hinv(qL[IH], p->eps_wet), n),
                    max_abs_eig(p, qR, zjo_hinv(qR[IH], p->eps_wet), n));
    double dq[NV];
    for (int v = 0; v < NV; ++v) {
        F[v] = 0.5 * (FLx[v] + FRx[v]);
        dq[v] = qR[v] - qL[v];
    }
    double Bd[NV] = {0.0, 0.0, 0.0, 0.0};
    double hb = 0.5 * (qL[IH] + qR[IH]);
    Bd[IU] = g * hb * dq[IB] * n[0];
    Bd[IVv] = g * hb * dq[IB] * n[1];
    double diss[NV];
    for (int v = 0; v < NV; ++v) diss[v] = s * dq[v];
    if (p->riemann == ZJO_RIEMANN_RUSANOV_POS) {
        diss[IB] = 0.0;
        diss[IH] = s * (dq[IH] + dq[IB]);
    }
    for (int v = 0; v < NV; ++v) {
        Dm[v] = 0.5 * (Bd[v] - diss[v]);
        Dp[v] = 0.5 * (Bd[v] + diss[v]);
    }
}

/* Cell-edge path integral of B between the cell mean and its own edge value
 * (the emitted interior nonconservative term, solver_jax.py:811-825 and
 * :866-883): q_d row = g (h_c + h_e)/2 (b_e - b_c) n_d. */
static void cell_edge_source(const zjo_params *p, const double *qc,
                             const double *qe, const double *n, double *out)
{
    double hb = 0.5 * (qc[IH] + qe[IH]);
    double db = qe[IB] - qc[IB];
    out[IB] = out[IH] = 0.0;
    out[IU] = p->g * hb * db * n[0];
    out[IVv] = p->g * hb * db * n[1];
}

/* The flux operator with the cell aux 1/h taken from the state H (Q itself,
 * or the step-start state for a frozen second RK stage). */
static void flux_operator_h(const zjo_mesh *m, const zjo_params *p,
                            const double *Q, const double *H, double *dQ);

void zjo_flux_operator(const zjo_mesh *m, const zjo_params *p, const double *Q,
                       double *dQ)
{
    flux_operator_h(m, p, Q, Q, dQ);
}

void zjo_flux_operator_o1(const zjo_mesh *m, const zjo_params *p, const double *Q,
                          const unsigned char *force_o1, double *dQ)
{
    g_force_o1 = force_o1;
    flux_operator_h(m, p, Q, Q, dQ);
    g_force_o1 = NULL;
}

static void flux_operator_h(const zjo_mesh *m, const zjo_params *p,
                            const double *Q, const double *H, double *dQ)
{
    const int N = m->n_cells, F = m->n_faces, nbf = m->n_bf;
    double *QL = malloc(sizeof(double) * NV * F);
    double *QR = malloc(sizeof(double) * NV * F);
    double *hinv = malloc(sizeof(double) * N);
    zjo_reconstruct(m, p, Q, QL, QR, NULL);
    ZJO_PAR_FOR
    for (int c = 0; c < N; ++c) hinv[c] = zjo_hinv(H[N + c], p->eps_wet);
    memset(dQ, 0, sizeof(double) * NV * N);
    /* boundary Q_R = BC(Q_L reconstructed) or the periodic partner state */
    for (int b = 0; b < nbf; ++b) {
        int f = m->bface_face[b];
        double n[2] = {m->face_normals[f], m->face_normals[F + f]};
        double ql[NV], o[NV];
        for (int v = 0; v < NV; ++v) ql[v] = QL[(size_t)v * F + f];
        if (m->bface_kind[b] == ZJO_BC_PERIODIC && m->bface_partner &&
            m->bface_partner[b] >= 0) {
            int pc = m->bface_cell[m->bface_partner[b]];
            for (int v = 0; v < NV; ++v) o[v] = Q[(size_t)v * N + pc];
        } else {
            ghost_at(m, b, ql, n, o);
        }
        for (int v = 0; v < NV; ++v) QR[(size_t)v * F + f] = o[v];
    }
    int use_ces = p->order >= 2;
    /* per-face terms (parallel), then the scatter in face order (serial):
     * interior faces A side, then B side (solver_jax.py:806-809), each face's
     * flux before its cell-edge terms; boundary faces to the owner only
     * (solver_jax.py:827-883) */
    double *tA = malloc(sizeof(double) * NV * F);   /* (F + Dm) |f| / |A|   */
    double *tB = malloc(sizeof(double) * NV * F);   /* (-F + Dp) |f| / |B|  */
    double *eA = use_ces ? malloc(sizeof(double) * NV * F) : NULL;
    double *eB = use_ces ? malloc(sizeof(double) * NV * F) : NULL;
    ZJO_PAR_FOR
    for (int f = 0; f < F; ++f) {
        int A = m->face_cells[f], B = m->face_cells[F + f];
        double n[2] = {m->face_normals[f], m->face_normals[F + f]};
        double ql[NV], qr[NV], Fx[NV], Dp[NV], Dm[NV];
        for (int v = 0; v < NV; ++v) {
            ql[v] = QL[(size_t)v * F + f];
            qr[v] = QR[(size_t)v * F + f];
        }
        zjo_face_flux(p, ql, qr, hinv[A], B >= 0 ? hinv[B] : hinv[A], n, Fx, Dp, Dm);
        double fa = m->face_volumes[f] / m->cell_volumes[A];
        double fb = B >= 0 ? m->face_volumes[f] / m->cell_volumes[B] : 0.0;
        for (int v = 0; v < NV; ++v) {
            tA[(size_t)v * F + f] = (Fx[v] + Dm[v]) * fa;
            tB[(size_t)v * F + f] = (-Fx[v] + Dp[v]) * fb;
        }
        if (use_ces) {
            double qa[NV], qb[NV], sa[NV], sb[NV], nm[2] = {-n[0], -n[1]};
            for (int v = 0; v < NV; ++v) qa[v] = Q[(size_t)v * N + A];
            cell_edge_source(p, qa, ql, n, sa);
            for (int v = 0; v < NV; ++v) eA[(size_t)v * F + f] = sa[v] * fa;
            if (B >= 0) {
                for (int v = 0; v < NV; ++v) qb[v] = Q[(size_t)v * N + B];
                cell_edge_source(p, qb, qr, nm, sb);
                for (int v = 0; v < NV; ++v) eB[(size_t)v * F + f] = sb[v] * fb;
            }
        }
    }
#ifdef ZJO_OMP
    if (m->cell_faces) {
        /* per cell over its faces in increasing id (interior faces precede
         * boundary faces): the serial scatter's order of operations */
        ZJO_PAR_FOR
        for (int c = 0; c < N; ++c)
            for (int k = 0; k < m->max_cf; ++k) {
                int f = m->cell_faces[(size_t)c * m->max_cf + k];
                if (f < 0) continue;
                const double *t = m->face_cells[f] == c ? tA : tB;
                const double *e = m->face_cells[f] == c ? eA : eB;
                for (int v = 0; v < NV; ++v) dQ[(size_t)v * N + c] -= t[(size_t)v * F + f];
                if (use_ces)
                    for (int v = 0; v < NV; ++v) dQ[(size_t)v * N + c] -= e[(size_t)v * F + f];
            }
        goto scattered;
    }
#endif
    for (int f = 0; f < F; ++f) {
        int A = m->face_cells[f], B = m->face_cells[F + f];
        if (B < 0) continue;
        for (int v = 0; v < NV; ++v) {
            dQ[(size_t)v * N + A] -= tA[(size_t)v * F + f];
            dQ[(size_t)v * N + B] -= tB[(size_t)v * F + f];
        }
        if (use_ces)
            for (int v = 0; v < NV; ++v) {
                dQ[(size_t)v * N + A] -= eA[(size_t)v * F + f];
                dQ[(size_t)v * N + B] -= eB[(size_t)v * F + f];
            }
    }
    for (int b = 0; b < nbf; ++b) {
        int f = m->bface_face[b];
        int A = m->face_cells[f];
        for (int v = 0; v < NV; ++v) dQ[(size_t)v * N + A] -= tA[(size_t)v * F + f];
        if (use_ces)
            for (int v = 0; v < NV; ++v) dQ[(size_t)v * N + A] -= eA[(size_t)v * F + f];
    }
#ifdef ZJO_OMP
scattered:
#endif
    free(tA);
    free(tB);
    free(eA);
    free(eB);
    free(QL);
    free(QR);
    free(hinv);
}

/* dt law (blocks.py:36-95, solver_jax.py:434-457, :971-973): per face the max
 * |lambda| of the two cell states along the face normal (a boundary face uses
 * its owner on both sides, context.py:111-115), dt = min_f CFL 2 r_f /
 * (d a_f), r_f = min(r_A, r_B); a wave-free face drops out; capped by dt_max
 * and the remaining time. */
double zjo_compute_dt(const zjo_mesh *m, const zjo_params *p, const double *Q,
                      double t_remaining)
{
    const int N = m->n_cells, F = m->n_faces;
    double dt = INFINITY;
    ZJO_PAR_FOR_MIN_SCALAR(dt)
    for (int f = 0; f < F; ++f) {
        int A = m->face_cells[f], B = m->face_cells[F + f];
        if (B < 0) B = A;
        double n[2] = {m->face_normals[f], m->face_normals[F + f]};
        double qa[NV], qb[NV];
        for (int v = 0; v < NV; ++v) {
            qa[v] = Q[(size_t)v * N + A];
            qb[v] = Q[(size_t)v * N + B];
        }
        double a = dmax(max_abs_eig(p, qa, zjo_hinv(qa[IH], p->eps_wet), n),
                        max_abs_eig(p, qb, zjo_hinv(qb[IH], p->eps_wet), n));
        double r = dmin(m->cell_inradius[A], m->cell_inradius[B]);
        double loc = p->cfl * 2.0 * r / ((double)m->dim * a);
        dt = dmin(loc, dt);
    }
    if (p->dt_max > 0.0) dt = dmin(dt, p->dt_max);
    if (t_remaining >= 0.0) dt = dmin(dt, t_remaining);
    return dt;
}

void zjo_manning_source(const zjo_params *p, const double *q, double hinv,
                        double *S)
{
    S[IB] = S[IH] = S[IU] = S[IVv] = 0.0;
    if (p->manning_n <= 0.0) return;
    double u = q[IU] * hinv, v = q[IVv] * hinv;
    double sp = sqrt(u * u + v * v);
    double k = p->g * p->manning_n * p->manning_n * cbrt(hinv);
    S[IU] = -k * u * sp;
    S[IVv] = -k * v * sp;
}

static int count_troubled(int N, const double *Q)
{
    int bad = 0;
    ZJO_PAR_FOR_SUM_INT(bad)
    for (int c = 0; c < N; ++c) {
        int t = Q[(size_t)N + c] < 0.0;
        for (int v = 0; v < NV; ++v) t |= !isfinite(Q[(size_t)v * N + c]);
        bad += t;
    }
    return bad;
}

/* _explicit_hyperbolic_step (solver_jax.py:990-1074) + explicit source
 * (ode.RK1, :1129).  The aux (hinv) is a function of the stage state, so the
 * stage-2 refresh of :1041-1042 is implicit here. */
int zjo_step(const zjo_mesh *m, const zjo_params *p, double dt,
             const double *Q, double *Qout)
{
    const int N = m->n_cells;
    size_t n = (size_t)NV * N;
    double *dQ = malloc(sizeof(double) * n);
    zjo_flux_operator(m, p, Q, dQ);
    ZJO_PAR_FOR
    for (size_t i = 0; i < n; ++i) Qout[i] = Q[i] + dt * dQ[i];
    if (p->order >= 2) {
        double *Q1 = malloc(sizeof(double) * n);
        memcpy(Q1, Qout, sizeof(double) * n);
        flux_operator_h(m, p, Q1,
                        p->stage_aux == ZJO_STAGE_AUX_FROZEN ? Q : Q1, dQ);
        ZJO_PAR_FOR
        for (size_t i = 0; i < n; ++i) {
            double q2 = Q1[i] + dt * dQ[i];
            Qout[i] = 0.5 * (Q[i] + q2);
        }
        free(Q1);
    }
    if (p->manning_n > 0.0) {
        ZJO_PAR_FOR
        for (int c = 0; c < N; ++c) {
            double q[NV], S[NV];
            for (int v = 0; v < NV; ++v) q[v] = Qout[(size_t)v * N + c];
            zjo_manning_source(p, q, zjo_hinv(Q[(size_t)N + c], p->eps_wet), S);
            for (int v = 0; v < NV; ++v) Qout[(size_t)v * N + c] += dt * S[v];
        }
    }
    free(dQ);
    return count_troubled(N, Qout);
}

/* A-posteriori MOOD (solver_jax.py:1055-1071): the order-2 candidate; if a
 * cell has h < 0 or a non-finite component, the whole step is redone with
 * those cells reconstructed at first order (shared face fluxes: still
 * conservative).  Returns the troubled count of the accepted state;
 * *n_demoted receives the number of demoted cells (0: candidate accepted). */
int zjo_step_mood(const zjo_mesh *m, const zjo_params *p, double dt,
                  const double *Q, double *Qout, int *n_demoted)
{
    const int N = m->n_cells;
    int bad = zjo_step(m, p, dt, Q, Qout);
    *n_demoted = 0;
    if (!bad || p->order < 2) return bad;
    unsigned char *mask = calloc((size_t)N, 1);
    for (int c = 0; c < N; ++c) {
        int t = Qout[(size_t)N + c] < 0.0;
        for (int v = 0; v < NV; ++v) t |= !isfinite(Qout[(size_t)v * N + c]);
        mask[c] = (unsigned char)t;
        *n_demoted += t;
    }
    g_force_o1 = mask;
    bad = zjo_step(m, p, dt, Q, Qout);
    g_force_o1 = NULL;
    free(mask);
    return bad;
}

void zjo_update_aux(const zjo_mesh *m, const zjo_params *p, const double *Q,
                    double *Qaux)
{
    const int N = m->n_cells, D = m->dim;
    double *gr = malloc(sizeof(double) * 2 * N);
    int row = 0;
    for (int d = 0; d < D; ++d) {
        zjo_lsq_grad(m, Q + (size_t)(IU + d) * N, NULL, gr);
        memcpy(Qaux + (size_t)row++ * N, gr + (size_t)d * N, sizeof(double) * N);
        zjo_lsq_grad(m, Q + (size_t)IH * N, NULL, gr);
        memcpy(Qaux + (size_t)row++ * N, gr + (size_t)d * N, sizeof(double) * N);
    }
    zjo_lsq_grad(m, Q + (size_t)IB * N, NULL, gr);
    for (int d = 0; d < D; ++d)
        memcpy(Qaux + (size_t)row++ * N, gr + (size_t)d * N, sizeof(double) * N);
    ZJO_PAR_FOR
    for (int c = 0; c < N; ++c)
        Qaux[(size_t)row * N + c] = zjo_hinv(Q[(size_t)N + c], p->eps_wet);
    free(gr);
}

double zjo_mass(const zjo_mesh *m, const double *Q)
{
    const int N = m->n_cells;
    double s = 0.0, comp = 0.0;
    for (int c = 0; c < N; ++c) {
        double x = Q[(size_t)N + c] * m->cell_volumes[c];
        double t = s + x;
        if (fabs(s) >= fabs(x)) comp += (s - t) + x;
        else comp += (x - t) + s;
        s = t;
    }
    return s + comp;
}

/* _build_implicit_local (solver_imex_jax.py:268-342): Newton from Qexp,
 * A = I - dt dS/dQ, d = A^{-1}(-(Q - Qexp - dt S(Q))).  h and b have no source,
 * so only the momentum block is solved (the h column of the chain term
 * dS/dhinv dhinv/dh multiplies a zero h increment). */
void zjo_imex_local(const zjo_mesh *m, const zjo_params *p, double dt,
                    int maxiter, const double *Qexp, double *Qout)
{
    const int N = m->n_cells;
    memcpy(Qout, Qexp, sizeof(double) * NV * N);
    if (p->manning_n <= 0.0) return;
    for (int c = 0; c < N; ++c) {
        double hinv = zjo_hinv(Qexp[(size_t)N + c], p->eps_wet);
        double k = p->g * p->manning_n * p->manning_n * cbrt(hinv);
        double qx = Qexp[(size_t)IU * N + c], qy = Qexp[(size_t)IVv * N + c];
        double ex = qx, ey = qy;
        for (int it = 0; it < maxiter; ++it) {
            double u = qx * hinv, v = qy * hinv;
            double sp = sqrt(u * u + v * v);
            double Sx = -k * u * sp, Sy = -k * v * sp;
            double jxx = 0, jxy = 0, jyx = 0, jyy = 0;
            if (sp > 0.0) {
                jxx = -k * hinv * (sp + u * u / sp);
                jxy = -k * hinv * (u * v / sp);
                jyx = jxy;
                jyy = -k * hinv * (sp + v * v / sp);
            }
            double a11 = 1.0 - dt * jxx, a12 = -dt * jxy;
            double a21 = -dt * jyx, a22 = 1.0 - dt * jyy;
            double rx = qx - ex - dt * Sx, ry = qy - ey - dt * Sy;
            double det = a11 * a22 - a12 * a21;
            double dx = (-rx * a22 + ry * a12) / det;
            double dy = (-ry * a11 + rx * a21) / det;
            qx += dx;
            qy += dy;
        }
        Qout[(size_t)IU * N + c] = qx;
        Qout[(size_t)IVv * N + c] = qy;
    }
}

int zjo_march(const zjo_mesh *m, const zjo_params *p, double *Q, double *t_io,
              double t_end, int n_steps)
{
    const int N = m->n_cells;
    double *Qn = malloc(sizeof(double) * NV * N);
    double t = *t_io;
    int k = 0;
    for (;;) {
        if (n_steps >= 0) {
            if (k >= n_steps) break;
        } else if (!(t < t_end)) {
            break;
        }
        double dt = zjo_compute_dt(m, p, Q, n_steps >= 0 ? -1.0 : t_end - t);
        if (!(dt > 0.0) || !isfinite(dt)) break;
        zjo_step(m, p, dt, Q, Qn);
        memcpy(Q, Qn, sizeof(double) * NV * N);
        t += dt;
        ++k;
    }
    *t_io = t;
    free(Qn);
    return k;
}
