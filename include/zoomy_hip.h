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

/* boundary kinds, limiters, Riemann solvers, reconstruction variables */
enum { ZJ_BC_EXTRAPOLATION = 0, ZJ_BC_WALL = 1, ZJ_BC_PERIODIC = 2 };
enum { ZJ_LIM_NONE = 0, ZJ_LIM_BARTH_JESPERSEN = 1, ZJ_LIM_MINMOD = 2,
       ZJ_LIM_VENKATAKRISHNAN = 3, ZJ_LIM_VANLEER = 4 };
enum { ZJ_RIEMANN_RUSANOV_NC = 0, ZJ_RIEMANN_RUSANOV = 1, ZJ_RIEMANN_HLL = 2 };
enum { ZJ_RECON_CONSERVATIVE = 0, ZJ_RECON_ETA = 1 };
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
} zj_mesh_desc;

typedef struct zj_mesh zj_mesh;       /* device-resident mesh   */
typedef struct zj_solver zj_solver;   /* mesh + params + scratch */

typedef struct {
    double g, cfl, eps_wet, dt_max, manning_n;
    int32_t order, limiter, riemann, recon, dtype;
} zj_params;

/* Replaces convert_mesh_to_jax (zoomy_jax/mesh/mesh.py:186-264) and
 * prepare_mesh (zoomy_jax/fvm/solver2/context.py:99-163). */
int zj_mesh_create(const zj_mesh_desc *desc, zj_mesh **out);
int zj_mesh_destroy(zj_mesh *m);

/* Reverse Cuthill-McKee permutation of a symmetric CSR graph (host).  The
 * reference has no RCM (only a Morton order, gnn_blueprint/
 * mesh_spectral_geometry.py:34-46); this follows the algorithm of
 * scipy.sparse.csgraph.reverse_cuthill_mckee with a stable degree sort. */
int zj_rcm(int32_t n, const int32_t *indptr, const int32_t *indices,
           int32_t *perm_out);

/* Host-only check of the LDS patch decomposition zj_mesh_create builds:
 * patches of P consecutive cells with their face ring and stencil ring,
 * local tables shared between patches of identical structure.  stats_out[6]
 * = {P, npatch, ntemplates, max_load, max_slope, max_vert}.  No counterpart
 * in the reference (it keeps no device mesh layout). */
int zj_patch_check(const zj_mesh_desc *desc, int32_t P, int32_t *stats_out);

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

/* One step with device dt (step, solver_jax.py:1095-1131): Qout = step(Q).
 * troubled_out (device int32, may be NULL) receives the number of cells with
 * h < 0 or a non-finite component (blocks.py:290-301). */
int zj_step(zj_solver *s, const double *dt, const void *Q, void *Qout,
            int32_t *troubled_out, void *stream);

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

/* Profiling: average duration (ms) of each kernel of a step over `reps`
 * launches, timed with HIP events on `stream`: ms_out[5] = {stage 1, stage 2
 * (+ RK average, source, troubled flag, dt of the new state), standalone dt,
 * standalone aux refresh, stage 1 with the fused aux refresh}.  No
 * counterpart in the reference (it has no profiler integration); used by
 * bench.py for the per-kernel roofline. */
int zj_time_kernels(zj_solver *s, const void *Q, const double *dt,
                    int32_t reps, float *ms_out, void *stream);

/* Compensated FP64 mass sum(h |cell|) into the device scalar *mass_out. */
int zj_mass(zj_solver *s, const void *Q, double *mass_out, void *stream);

/* Local implicit source stage (solver_imex_jax.py:268-342): maxiter Newton
 * iterations per cell on Q - Qexp - dt S(Q) = 0. */
int zj_imex_local(zj_solver *s, const double *dt, int32_t maxiter,
                  const void *Qexp, void *Qout, void *stream);

/* Library identification string. */
const char *zj_version(void);

#ifdef __cplusplus
}
#endif
#endif
