/* zoomy_hip -- C ABI of the MI355X finite-volume time step.
 *
 * Every entry point takes plain pointers and sizes, returns 0 or a ZJ_E*
 * code, and is asynchronous on the given HIP stream (hipStream_t passed as
 * void*; NULL = default stream).  Device arrays are laid out as in the
 * reference: state Q is (4, N) row-major SoA with rows (b, h, hu, hv); a 1-D
 * mesh keeps hv == 0.  Scalars that the step produces (dt, mass, troubled
 * count) stay in device memory; the host reads them only where the reference
 * synchronises (the dt guard, zoomy_jax/fvm/solver2/blocks.py:98-124).
 *
 * The Python host (zoomy_hip.fvm) binds these with ctypes; a jax.ffi custom
 * call would bind the same symbols.
 */
#ifndef ZOOMY_HIP_H
#define ZOOMY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define ZJ_OK 0
#define ZJ_EINVAL 1     /* bad argument / shape                        */
#define ZJ_EHIP 2       /* a HIP runtime call failed                   */
#define ZJ_ENOMEM 3     /* device allocation failed                    */
#define ZJ_EUNSUP 4     /* unsupported option combination              */
#define ZJ_ESOLVE 5     /* pressure solve missed its accept gate       */

/* boundary kinds, limiters, Riemann solvers, reconstruction variables */
enum { ZJ_BC_EXTRAPOLATION = 0, ZJ_BC_WALL = 1, ZJ_BC_PERIODIC = 2,
       ZJ_BC_DIRICHLET = 3 };
enum { ZJ_LIM_NONE = 0, ZJ_LIM_BARTH_JESPERSEN = 1, ZJ_LIM_MINMOD = 2,
       ZJ_LIM_VENKATAKRISHNAN = 3, ZJ_LIM_VANLEER = 4 };
enum { ZJ_RIEMANN_RUSANOV_NC = 0, ZJ_RIEMANN_RUSANOV = 1, ZJ_RIEMANN_HLL = 2 };
/* ZJ_RECON_XZ: conservative variables with the Xing-Zhang cell-mean
 * positivity scaling (PositivityPreservingLSQMUSCLJAX,
 * reconstruction_jax.py:853-984).  ZJ_RECON_ETA: (b, eta, u, v) with the
 * face clip (EmittedWLSQMUSCLJAX, :987-1292).  ZJ_RECON_ETA_HU: (b, eta,
 * hu, hv), cells with h < 1e-3 at first order and face momentum zeroed
 * where h_f < 1e-3 -- the class EmittedWLSQMUSCLJAX replaced (:995-1008),
 * with which the reference's ritter_dry_o2_small golden was recorded. */
enum { ZJ_RECON_CONSERVATIVE = 0, ZJ_RECON_ETA = 1, ZJ_RECON_XZ = 2,
       ZJ_RECON_ETA_HU = 3 };
enum { ZJ_F64 = 0, ZJ_F32 = 1 };

/* Host description of a mesh (topology + vertex coordinates).  All arrays
 * are host pointers; zj_mesh_create copies them to the device.
 *   cell_vertices (max_faces, n_cells)  slot k = face from vertex k to k+1
 *   cell_nbr      (max_faces, n_cells)  cell across slot k, or -1-bface
 *   cell_extra    (max_extra, n_cells)  other vertex neighbours, -1 padded
 *   verts         (n_verts, 2)          float64
 *   cell_center   (2, n_cells)          float64
 *   cell_inradius (n_cells)             float64
 *   bface_cell, bface_slot, bface_kind, bface_partner (n_bf)
 */
typedef struct {
    int32_t dim, n_cells, n_verts, max_faces, max_extra, n_bf;
    const int32_t *cell_vertices, *cell_nbr, *cell_extra;
    const double *verts, *cell_center, *cell_inradius;
    const int32_t *bface_cell, *bface_slot, *bface_kind, *bface_partner;
    /* ZJ_BC_DIRICHLET faces (may be NULL when there are none): bit v of
     * bface_mask[b] prescribes state row v (b, h, hu, hv) of face b at
     * bface_value[v * n_bf + b]; the other rows extrapolate */
    const int32_t *bface_mask;
    const double *bface_value;
    /* patch decomposition of the stage kernel (0, 0: automatic):
     * patch_cells = cells per patch (a power of two, 32..256; halved while
     * the FP64 working set exceeds the LDS budget); patch_order = 1:
     * consecutive ids, 2: clusters of the Morton order of the cell centres
     * (2-D), 0: consecutive ids unless their ring cells exceed 3x the patch
     * and clusters do better */
    int32_t patch_cells, patch_order;
} zj_mesh_desc;

typedef struct zj_mesh zj_mesh;       /* device-resident mesh   */
typedef struct zj_solver zj_solver;   /* mesh + params + scratch */

typedef struct {
    double g, cfl, eps_wet, dt_max, manning_n;
    int32_t order, limiter, riemann, recon, dtype;
} zj_params;

/* Replaces convert_mesh_to_jax (zoomy_jax/mesh/mesh.py:186-264) and
 * prepare_mesh (zoomy_jax/fvm/solver2/context.py:99-163).  Meshes of more
 * than 268,435,455 cells (8 rows of 32-bit offsets) are refused with
 * ZJ_EUNSUP. */
int zj_mesh_create(const zj_mesh_desc *desc, zj_mesh **out);
int zj_mesh_destroy(zj_mesh *m);

/* Reverse Cuthill-McKee permutation of a symmetric CSR graph (host).  The
 * reference has no RCM (only a Morton order, gnn_blueprint/
 * mesh_spectral_geometry.py:34-46); this follows the algorithm of
 * scipy.sparse.csgraph.reverse_cuthill_mckee with a stable degree sort. */
int zj_rcm(int32_t n, const int32_t *indptr, const int32_t *indices,
           int32_t *perm_out);
/* The same ordering with the search of start's component begun at start
 * (a caller's pseudo-peripheral node, or the start node another RCM chose
 * among equal-degree candidates); the other components as in zj_rcm. */
int zj_rcm_from(int32_t n, const int32_t *indptr, const int32_t *indices,
                int32_t start, int32_t *perm_out);

/* Host-only check of the LDS patch decomposition zj_mesh_create builds:
 * patches of P consecutive cells with their face ring and stencil ring,
 * local tables shared between patches of identical structure.  stats_out[6]
 * = {P, npatch, ntemplates, max_load, max_slope, max_vert}.  No counterpart
 * in the reference (it keeps no device mesh layout). */
int zj_patch_check(const zj_mesh_desc *desc, int32_t P, int32_t *stats_out);
/* Host-only: the decomposition zj_mesh_create picks for desc.  stats_out[7]
 * = {P, npatch, ntemplates, max_load, max_slope, max_vert, clustered}.  No
 * counterpart in the reference. */
int zj_patch_plan(const zj_mesh_desc *desc, int32_t *stats_out);

/* Solver context: owns the per-stage scratch (limited slopes, stage state,
 * reduction scalars).  Replaces HyperbolicSolver.setup_simulation
 * (zoomy_jax/fvm/solver_jax.py:901-988). */
int zj_solver_create(zj_mesh *m, const zj_params *p, zj_solver **out);
int zj_solver_destroy(zj_solver *s);

/* Spatial residual dQ = L(Q) (flux_operator, solver_jax.py:722-895). */
int zj_flux_operator(zj_solver *s, const void *Q, void *dQ, void *stream);

/* dt of Q (compute_timestep, solver_jax.py:1222-1236; blocks.py:36-95):
 * writes the device scalar *dt_out.  t_remaining < 0 means no time cap. */
int zj_compute_dt(zj_solver *s, const void *Q, double t_remaining,
                  double *dt_out, void *stream);

/* zj_compute_dt without device allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Same kernels and result as zj_compute_dt,
 * with its block partials and reduction word in a static device array of
 * the library (512 KiB per GPU, kept until the process ends) instead of the
 * solver's scratch: meshes of more than 16,777,216 cells are refused
 * (ZJ_EINVAL).  Calls that overlap on one GPU share that array and race. */
int zj_compute_dt_noalloc(zj_solver *s, const void *Q, double t_remaining,
                          double *dt_out, void *stream);

/* One step with device dt (step, solver_jax.py:1095-1131): Qout = step(Q).
 * troubled_out (device int32, may be NULL) receives the number of cells with
 * h < 0 or a non-finite component (blocks.py:290-301). */
int zj_step(zj_solver *s, const double *dt, const void *Q, void *Qout,
            int32_t *troubled_out, void *stream);

/* A-posteriori MOOD (solver_jax.py:1055-1071, HyperbolicSolver.mood): every
 * order-2 step (zj_step, zj_march) first computes a candidate; if a cell has
 * h < 0 or a non-finite component, the step is redone from the same input
 * with those cells reconstructed at first order (phi := 0, shared face
 * fluxes: conservative).  The redo is decided on the device (no host
 * synchronisation).  Off by default. */
int zj_solver_set_mood(zj_solver *s, int32_t on);
/* A-priori wet/dry guards of the eta reconstruction (HyperbolicSolver
 * positivity_method / front_theta_tol, solver_jax.py:205-217; the
 * reconstruction at reconstruction_jax.py:1211-1244): method
 * ZJ_POSITIVITY_ZHANG_SHU caps the depth deviation of every cell by the
 * Xing-Zhang-Shu theta (the bed slope kept, the velocity limiters scaled by
 * theta); front_tol in (0, 1] reconstructs at first order every cell whose
 * theta is below it (<= 0: off).  Both are ignored by the other
 * reconstructions, as in the reference. */
enum { ZJ_POSITIVITY_NONE = 0, ZJ_POSITIVITY_ZHANG_SHU = 1 };
int zj_solver_set_positivity(zj_solver *s, int32_t method, double front_tol);
/* Which 1/h the fluxes of the second SSP-RK2 stage read: REFRESH (the
 * default) that stage's own state, as solver_jax.py:1026-1041 does;
 * FROZEN the step-start aux, which the reference's recorded order-2 goldens
 * (stoker_wet_o2_small, ritter_dry_o2_small, boundary_o2_small) were made
 * with.  The reconstruction's own 1/h (eta velocity) is the stage state's
 * either way. */
enum { ZJ_STAGE_AUX_REFRESH = 0, ZJ_STAGE_AUX_FROZEN = 1 };
int zj_solver_set_stage_aux(zj_solver *s, int32_t mode);
/* {steps redone, cells demoted} since MOOD was enabled (host-synchronous). */
int zj_solver_mood_stats(zj_solver *s, int32_t *out2);
/* post_step aux refresh (solver_jax.py:1133-1172): Qaux rows per direction
 * d: [d q_d/dx_d, d h/dx_d], then d b/dx_d, then hinv. */
int zj_update_aux(zj_solver *s, const void *Q, void *Qaux, void *stream);

/* Device march of n_steps adaptive steps (run_simulation's while loop,
 * solver_jax.py:1238-1352): Q is updated in place, *t_dev advanced, dt of
 * every step capped by t_end - t (t_end < 0: no cap).  Qaux (may be NULL)
 * receives the post_step aux refresh after every step. */
int zj_march(zj_solver *s, void *Q, void *Qaux, int32_t n_steps,
             double t_end, double *t_dev, double *dt_dev,
             int32_t *troubled_out, void *stream);

/* flux_operator with the per-cell force_o1 demotion mask
 * (solver_jax.py:724-785): cells with force_o1[c] != 0 reconstruct at
 * first order; force_o1 (device uint8, N) may be NULL. */
int zj_flux_operator_o1(zj_solver *s, const void *Q, const uint8_t *force_o1,
                        void *dQ, void *stream);

/* ───────────── the block contract (zoomy_jax/fvm/solver2/blocks.py) ─────────────
 * Separate kernels on stored per-face arrays, in the face order of the
 * reference's mesh (face_cells owner / neighbour, -1 on boundary faces).
 * Arrays are device arrays of the solver's dtype unless noted. */
/* The reference-layout face list: face f joins cell face_cells[f] (its slot
 * face_slot[f]) and face_cells[n_faces + f] (slot face_slot[n_faces + f]),
 * -1 for a boundary face; every cell slot must be named exactly once
 * (host arrays, checked; ZJ_EINVAL otherwise).  Needed by the calls below
 * (MeshJAX.face_cells, zoomy_jax/mesh/mesh.py:123-169). */
int zj_mesh_set_faces(zj_mesh *m, int32_t n_faces, const int32_t *face_cells,
                      const int32_t *face_slot);
/* dt_pass (blocks.py:36-61) and get_compute_max_abs_eigenvalue
 * (solver_jax.py:434-457): per face the lower and upper bounds lo, hi (F) of
 * the eigenvalues {0, u.n - c, u.n, u.n + c} of both cell states (the owner
 * on both sides of a boundary face); max |lambda| = max(|lo|, |hi|).
 * hinv_aux (N, may be NULL: from Q) is the aux 1/h row. */
int zj_face_speeds(zj_solver *s, const void *Q, const void *hinv_aux, void *lo,
                   void *hi, void *stream);
/* reduce_dt (blocks.py:64-95): *dt_out = min_f CFL 2 r_f / (d (2k+1)
 * max(|lo_f|, |hi_f|)), wave-free faces dropping out, capped by dt_max
 * (> 0) and clamp (>= 0: the remaining time or the coupling window);
 * inradius_f (F, FP64) = min(r_A, r_B). */
int zj_reduce_dt(zj_solver *s, const void *lo, const void *hi,
                 const double *inradius_f, double cfl, int32_t dimension,
                 int32_t degree, double dt_max, double clamp, double *dt_out,
                 void *stream);
/* halo_bc (blocks.py:129-157): bf_values (4, n_bf) = the BC ghost of each
 * boundary face's cell state. */
int zj_bc_values(zj_solver *s, const void *Q, void *bf_values, void *stream);
/* The reconstruction object (reconstruction_jax.py:398-429; eta:
 * :1307-1366): face states QL, QR (4, F); Q_R of a boundary face is the
 * ghost of the cell state (the halo_bc values).  force_o1 (uint8, N, may be
 * NULL) demotes cells to first order; order1 != 0 gives the constant
 * reconstruction (MOOD redo).  lim_grad (4, dim, N, may be NULL) receives
 * the limited cell gradient in conservative variables (reconstruct_with_grad;
 * the eta reconstruction's W slopes pushed through the inverse map). */
int zj_reconstruct(zj_solver *s, const void *Q, const uint8_t *force_o1,
                   int32_t order1, void *QL, void *QR, void *lim_grad,
                   void *stream);
/* flux_pass (blocks.py:185-238): F, Dp, Dm (4, F) per face; a boundary
 * face's Q_R is recomputed as BC(Q_L) (the partner cell at a periodic seam);
 * hinv_aux (N, may be NULL: from Q) is the aux 1/h the fluxes read. */
int zj_flux_pass(zj_solver *s, const void *Q, const void *hinv_aux,
                 const void *QL, const void *QR, void *Fface, void *Dp,
                 void *Dm, void *stream);
/* The runtime slots numerical_flux / numerical_fluctuations
 * (transformation/jax_runtime.py:789-806): the solver's Riemann solver on n
 * given face states qL, qR (4, n) with the aux of the faces' cells qauxL,
 * qauxR (3 dim + 1, n; the hinv row is read; NULL: the hinv of the face
 * state) and FP64 unit normals (dim, n).  Writes any of F (4, n) and the
 * fluctuations Dp, Dm (4, n) (both or neither). */
int zj_face_numerics(zj_solver *s, int32_t n, const void *qL, const void *qR,
                     const void *qauxL, const void *qauxR, const double *normals,
                     void *Fface, void *Dp, void *Dm, void *stream);
/* gather_update (blocks.py:241-301): Qcand = Q0 + a dt (-div(F + D) +
 * cell term + source) gathered per cell over its faces; the cell term is
 * cell_term (4, N) or, when that is NULL and lim_grad (4, dim, N) is given,
 * -B(Q0) . lim_grad (the interior nonconservative integral, march.py
 * :144-149); source != 0 adds the explicit Manning friction of Q0.
 * troubled (uint8, N, may be NULL) flags h < 0 or a non-finite component;
 * *count (device int32, may be NULL) is incremented by their number.
 * dt: device scalar. */
int zj_gather_update(zj_solver *s, const void *Q0, const void *Fface,
                     const void *Dp, const void *Dm, const double *dt,
                     double a_stage, const void *cell_term,
                     const void *lim_grad, int32_t source,
                     const void *hinv_aux, void *Qcand, uint8_t *troubled,
                     int32_t *count, void *stream);

/* Profiling: average duration (ms) of each kernel of a step over `reps`
 * launches, timed with HIP events on `stream`: ms_out[5] = {stage 1, stage 2
 * (+ RK average, source, troubled flag, dt of the new state), standalone dt,
 * standalone aux refresh, stage 1 with the fused aux refresh}.  No
 * counterpart in the reference (it has no profiler integration); used by
 * bench.py for the per-kernel roofline. */
int zj_time_kernels(zj_solver *s, const void *Q, const double *dt,
                    int32_t reps, float *ms_out, void *stream);

/* ───────────────────────── batched ensemble ─────────────────────────
 * Independent instances stored as one mesh: the cells of instance i are
 * cell_offsets[i] .. cell_offsets[i+1]-1 (each more than 64 cells) and no
 * face joins two instances.  Every step advances every instance with its
 * own dt (CFL law per instance) in the same two stage launches over the
 * whole mesh; per-instance t and dt are device arrays (n_inst doubles).  The
 * solver must be order 2 without MOOD or the frozen stage aux.  No
 * counterpart in the reference (it marches one mesh at a time); the
 * numerics per instance are zj_march's. */
typedef struct zj_ensemble zj_ensemble;
int zj_ensemble_create(zj_solver *s, int32_t n_inst, const int32_t *cell_offsets,
                       zj_ensemble **out);
int zj_ensemble_destroy(zj_ensemble *e);
/* dt of every instance's state into dt_out (n_inst, device). */
int zj_ensemble_dt(zj_ensemble *e, const void *Q, double *dt_out, void *stream);
/* n_steps steps of every instance, Q in place, t_dev advanced; dt of each
 * step capped by t_end - t (t_end < 0: no cap). */
int zj_ensemble_march(zj_ensemble *e, void *Q, int32_t n_steps, double t_end,
                      double *t_dev, double *dt_dev, void *stream);
/* Captures n_steps steps on Q, t_dev, dt_dev into a HIP graph (kernel
 * nodes only); zj_ensemble_replay launches it on `stream`, each replay
 * continuing the march.  Capturing computes the dt of the current Q and
 * synchronises the stream. */
int zj_ensemble_capture(zj_ensemble *e, void *Q, int32_t n_steps, double t_end,
                        double *t_dev, double *dt_dev, void *stream);
int zj_ensemble_replay(zj_ensemble *e, void *stream);
/* Compensated FP64 mass of every instance into mass_out (n_inst, device). */
int zj_ensemble_mass(zj_ensemble *e, const void *Q, double *mass_out, void *stream);

/* Compensated FP64 mass sum(h |cell|) into the device scalar *mass_out. */
int zj_mass(zj_solver *s, const void *Q, double *mass_out, void *stream);

/* Local implicit source stage (solver_imex_jax.py:268-342): maxiter Newton
 * iterations per cell on Q - Qexp - dt S(Q) = 0. */
int zj_imex_local(zj_solver *s, const double *dt, int32_t maxiter,
                  const void *Qexp, void *Qout, void *stream);

/* Global implicit source stage (_build_implicit_global,
 * solver_imex_jax.py:344-396): Newton on R(Q) = Q - Qexp - dt S(Q, aux(Q))
 * for a source that reads the derivative aux rows -- rows q_d: the model's
 * Manning friction and -alpha g h dh/dx_d, dh/dx_d the LSQ aux row of the
 * post-step refresh -- so the stage couples every cell to its stencil.
 * Each Newton step solves J delta = -R by GMRES(restart) with the exact
 * Jacobian-vector product (the linearisation, through aux(v)), relative
 * tolerance gmres_tol, at most gmres_maxiter restart cycles; a GMRES that
 * misses its tolerance takes delta = -R (the reference's fallback).  The
 * Newton loop stops at maxiter steps or ||R||_2 <= tol.  FP64 solvers only;
 * host-synchronous: the Krylov scalars stay on the device, the host reads
 * ||R|| once per Newton step and the GMRES convergence flags once per chunk
 * of Arnoldi iterations. */
typedef struct zj_imex_global zj_imex_global;
int zj_imex_global_create(zj_solver *s, double alpha, int32_t restart,
                          zj_imex_global **out);
int zj_imex_global_destroy(zj_imex_global *g);
int zj_imex_global_solve(zj_imex_global *g, double dt, const double *Qexp,
                         double *Qout, int32_t maxiter, double tol,
                         double gmres_tol, int32_t gmres_maxiter, void *stream);
/* {Newton steps, GMRES iterations, GMRES fallbacks} and the final ||R||_2
 * of the last solve. */
int zj_imex_global_stats(zj_imex_global *g, int32_t *out3, double *rnorm);

/* ─────────────────── implicit diffusion (Crank-Nicolson) ───────────────────
 * The scalar nu-Laplacian of DiffusionOperatorJAX (zoomy_jax/fvm/
 * reconstruction_jax.py:1423-1520), which the IMEX solver applies per state
 * row between its explicit hyperbolic stage and its implicit source
 * (solver_imex_jax.py:396-460, 617-619): two-point fluxes
 * nu |f| (u_b - u_a) / |x_b - x_a| / max(|n.e|, 0.1) on interior faces,
 * none on boundary faces. */
typedef struct zj_diffusion zj_diffusion;
/* Face coefficients and volumes from the host mesh (explicit() of the
 * reference's setup, :1437-1479). */
int zj_diffusion_create(const zj_mesh_desc *desc, zj_diffusion **out);
int zj_diffusion_destroy(zj_diffusion *d);
/* Lu = L u, FP64 device arrays of n_cells (DiffusionOperatorJAX.explicit,
 * :1481-1483). */
int zj_diffusion_explicit(zj_diffusion *d, double nu, const double *u, double *Lu,
                          void *stream);
/* u_out solves (I - dt/2 L) u_out = (I + dt/2 L) u_star (implicit_solve,
 * :1485-1520; the reference solves densely): Jacobi-preconditioned CG on the
 * volume-weighted system from u_star, to a relative residual <= tol within
 * maxiter iterations, else ZJ_ESOLVE.  The host reads the convergence flag
 * of each chunk of 8 iterations while the next chunk runs, and returns with
 * the stream idle; u_out must not alias u_star. */
int zj_diffusion_cn(zj_diffusion *d, double nu, double dt, const double *u_star,
                    double *u_out, double tol, int32_t maxiter, void *stream);
/* CG iterations and relative residual of the last zj_diffusion_cn. */
int zj_diffusion_stats(zj_diffusion *d, int32_t *iters, double *relres);

/* ───────────────────────── Chorin projection ─────────────────────────────
 * Non-hydrostatic projection on an nx x ny wall-bounded quad grid whose
 * cells are numbered i*ny + j (Mesh.quad order).  State (5, N) rows
 * (b, h, hu, hv, hw).  The FP64 pressure operator S is symmetric positive
 * definite (5-point; a_c = h_c + b_c - b_e, c_f = |f|/d_f hinv(h_f),
 * S_cc = sum_f c_f a_c^2 + 4|A_c| hinv(h_c), S_ce = -c_f a_c a_e).
 * Pressure solves are host-synchronous once per chunk of CG iterations.
 * Replaces ChorinSplitVAMSolverJax (zoomy_jax/fvm/solver_chorin_vam_jax.py). */
/* Preconditioners of the FP64 CG: the FP64 V-cycle; the FP32 V-cycle
 * (mixed precision, flexible CG); the FP32 V-cycle plus one BF16
 * message-passing correction (learned, flexible CG).  The low-precision
 * forms are checked by a fresh FP64 residual and redone plainly when they
 * miss tol.  ZJ_PRECOND_LEARNED_GUESS applies the learned V-cycle once, to
 * the right-hand side, as the initial guess of the FP64 V-cycle CG, and
 * retries from zero when that solve misses tol (the learned initial guess
 * of gnn_blueprint/imex_child_solver.py:253-283). */
enum { ZJ_PRECOND_MG = 0, ZJ_PRECOND_LEARNED = 1, ZJ_PRECOND_MIXED = 2,
       ZJ_PRECOND_LEARNED_GUESS = 3 };
typedef struct zj_chorin zj_chorin;
typedef struct zj_mpnn zj_mpnn;     /* learned preconditioner (zj_mpnn_*) */
typedef struct {
    double tol;          /* relative residual target (1e-8)                  */
    double eps_wet;      /* hinv desingularisation (1e-8)                    */
    int32_t maxiter;     /* CG iteration cap per solve                       */
    int32_t nu;          /* smoothing sweeps per level and direction         */
    int32_t coarse_sweeps; /* symmetric sweep pairs on the coarsest level    */
    int32_t min_coarse;  /* stop coarsening before a side drops below this   */
    int32_t precond;     /* ZJ_PRECOND_*                                     */
    int32_t time_order;  /* 1: one stage; 2: two stages + average            */
} zj_chorin_params;
/* Counters of the last zj_chorin_solve / _stage / _step call.  lp_*: the
 * solves with a low-precision preconditioner (mixed or learned), their CG
 * iterations and worst fresh FP64 residual; fallbacks: those redone
 * plainly.  iterations / relres: of the accepted solves. */
typedef struct {
    int32_t solves, iterations, fallbacks, lp_iterations, lp_solves;
    double relres, lp_relres; /* worst relative residual of the call */
} zj_chorin_stats;
/* fv: an FP64 zj_solver of the same nx x ny mesh (hyperbolic predictor). */
int zj_chorin_create(zj_solver *fv, int32_t nx, int32_t ny, double dx, double dy,
                     const zj_chorin_params *prm, zj_chorin **out);
int zj_chorin_destroy(zj_chorin *c);
/* Cell order of the FV solver fv (host array, a permutation): FV cell k is
 * grid cell order[k] (e.g. a Morton order, so that the predictor's patch
 * kernel gathers from compact tiles).  NULL: the grid order.  No counterpart
 * in the reference (its Chorin solver has no device mesh layout). */
int zj_chorin_set_fv_order(zj_chorin *c, const int32_t *order);
/* Mixed-precision predictor: the FV fluxes of the hyperbolic stage on fv32,
 * an FP32 zj_solver of the same cells in the same order as the FP64 one
 * (NULL: back to FP64).  The state, the pressure residual and the
 * pressure solve stay FP64. */
int zj_chorin_set_fv32(zj_chorin *c, zj_solver *fv32);
/* Fused multigrid kernels on/off (default on; used when nu == 1).  Both
 * give bit-identical V-cycles.  No counterpart in the reference. */
int zj_chorin_set_mg_fused(zj_chorin *c, int32_t on);
/* Select the preconditioner; net must be non-NULL for ZJ_PRECOND_LEARNED
 * and ZJ_PRECOND_LEARNED_GUESS. */
int zj_chorin_set_precond(zj_chorin *c, int32_t precond, zj_mpnn *net);
/* chorin_cycle (solver_chorin_vam_jax.py:574-615): one step of time_order
 * stages; returns ZJ_ESOLVE when a pressure solve misses the accept gate
 * max(1e-6, 10 tol) (:72-91, EllipticSolveError). */
int zj_chorin_step(zj_chorin *c, double dt, const double *Q, double *Qout, void *stream);
/* chorin_stage (:524-572): hyperbolic Euler -> pressure -> corrector; with
 * Q0 != NULL the result is averaged with Q0 (second SSP-RK2 stage). */
int zj_chorin_stage(zj_chorin *c, double dt, const double *Q, double *Qout,
                    const double *Q0, void *stream);
/* _hyperbolic_stage (:496-522): Qs = Q + dt L(Q), hw advected. */
int zj_chorin_predict(zj_chorin *c, double dt, const double *Q, double *Qs, void *stream);
/* Right-hand side of the pressure equation of a predicted state:
 * rhs = -D(Qs)/dt (_step_pressure_pure's residual, :728-743). */
int zj_chorin_rhs(zj_chorin *c, double dt, const double *Qs, double *rhs, void *stream);
/* _step_corrector_pure (:885-907). */
int zj_chorin_correct(zj_chorin *c, double dt, const double *Qs, const double *p,
                      const double *Q0, double *Qout, void *stream);
/* Build the operator hierarchy of S(h, b) (device rows b, h). */
int zj_chorin_setup_operator(zj_chorin *c, const double *b, const double *h, void *stream);
/* y = S x on the current operator (the JVP matvec, :759-761). */
int zj_chorin_apply(zj_chorin *c, const double *x, double *y, void *stream);
/* z = V-cycle(r) on the current operator, FP64 or FP32 hierarchy. */
int zj_chorin_vcycle(zj_chorin *c, const void *r, void *z, int32_t dtype, void *stream);
/* Solve S p = rhs with the selected preconditioner (replaces the GMRES of
 * :800-805): FP64 CG, flexible for the mixed and learned preconditioners,
 * whose results are accepted only when a fresh FP64 residual meets tol and
 * are redone with the FP64 V-cycle otherwise. */
int zj_chorin_solve(zj_chorin *c, const double *rhs, double *p, void *stream);
/* Number of multigrid levels and their (nx, ny). */
int zj_chorin_levels(zj_chorin *c, int32_t *n_levels, int32_t *dims_out, int32_t max_levels);
/* Statistics of the last step / stage / solve call. */
int zj_chorin_get_stats(zj_chorin *c, zj_chorin_stats *out);
/* The learned V-cycle z = M r (FP32, level 0 of the current operator; needs
 * the network set by zj_chorin_set_precond):
 *   x = V r (the FP32 V-cycle);  z = x + N(r - S x).
 * The network is a learned smoother, the role of smooth_step in the
 * reference's learned V-cycle prior
 * (zoomy_jax/gnn_blueprint/vcycle_structured_gnn.py:74-122, :145-200).
 * train != 0 keeps the activations for the backward below. */
int zj_chorin_learned_vcycle(zj_chorin *c, const float *r, float *z, int32_t train,
                             void *stream);
/* Accumulates into the network's gradient the parameter gradient of
 * sum(g * z) for the last training zj_chorin_learned_vcycle. */
int zj_chorin_learned_vcycle_backward(zj_chorin *c, const float *g, void *stream);

/* ───────────────── learned message-passing preconditioner ─────────────────
 * Network on 2x2 cell blocks (zoomy_jax/gnn_blueprint/graph_mp_multibranch.py
 * :9-16 made linear in the residual): h0 = (r/diag) W_in; three layers
 * h' = sigmoid(phi U_l) * ([h | sum_nbr h] W_l), hidden 64, BF16 MFMA;
 * z = h3 W_out.  FP32 master parameters (zj_mpnn_param_count() floats in
 * the order W_in (4x64) | W_l (128x64), U_l (16x64), l = 0..2 | W_out
 * (64x4)); Adam as in the reference's training
 * (train_vcycle_structured_poisson.py:266-312). */
int64_t zj_mpnn_param_count(void);
int zj_mpnn_create(zj_mpnn **out);
int zj_mpnn_destroy(zj_mpnn *n);
int zj_mpnn_set_params(zj_mpnn *n, const float *host, int64_t count);
int zj_mpnn_get_params(zj_mpnn *n, float *host, int64_t count);
int zj_mpnn_get_grad(zj_mpnn *n, float *host, int64_t count);
int zj_mpnn_zero_grad(zj_mpnn *n, void *stream);
int zj_mpnn_adam(zj_mpnn *n, float lr, float beta1, float beta2, float eps, void *stream);
/* z (+)= N(r) on B stacked nx x ny grids with FP32 stencils (diag, cE, cN);
 * slot 0/1 keep activations for zj_mpnn_backward, slot 2 does not.  The
 * operator features are recomputed when diag or the grid differ from the
 * last call, or after zj_mpnn_invalidate_features (call it when the stencil
 * arrays were rewritten in place). */
int zj_mpnn_forward(zj_mpnn *n, int32_t B, int32_t nx, int32_t ny, const float *r,
                    const float *diag, const float *cE, const float *cN, float *z,
                    int32_t add, int32_t slot, void *stream);
int zj_mpnn_invalidate_features(zj_mpnn *n);
/* Profiling: average duration (ms) of `reps` back-to-back inference
 * launches on one nx x ny grid, timed with HIP events on `stream`.  No
 * counterpart in the reference; used by bench.py for the MFMA rate. */
int zj_mpnn_time_forward(zj_mpnn *n, int32_t nx, int32_t ny, const float *r, const float *diag,
                         const float *cE, const float *cN, float *z, int32_t reps,
                         float *ms_out, void *stream);
/* Inference (slot 2) in one fused launch (default) or layer by layer; both
 * give the same result.  No counterpart in the reference. */
int zj_mpnn_set_fused(zj_mpnn *n, int32_t on);
/* Accumulates the gradient of sum(dz * z) for the forward of `slot`; dr
 * (may be NULL) receives the input gradient. */
int zj_mpnn_backward(zj_mpnn *n, int32_t B, int32_t nx, int32_t ny, const float *dz,
                     const float *diag, float *dr, int32_t slot, void *stream);

/* Library identification string. */
const char *zj_version(void);

#ifdef __cplusplus
}
#endif
#endif
