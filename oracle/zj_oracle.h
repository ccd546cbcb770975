/* FP64 CPU oracle for the finite-volume time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU restatement).  The product path never calls it.
 *
 * It restates, in plain C, the arithmetic of the reference's hot path on the
 * reference's own data layout (state Q is (n_var, N) row-major SoA, faces carry
 * owner A = face_cells[0] and neighbour B = face_cells[1], boundary faces have
 * the inner cell as owner and an outward normal):
 *
 *   zoomy_jax/fvm/solver_jax.py:563-897      flux operator (BC, reconstruct,
 *                                            face flux, gather update)
 *   zoomy_jax/fvm/reconstruction_jax.py      LSQ-MUSCL, limiters, eta recon
 *   zoomy_jax/mesh/mesh.py:74-120            LSQ derivative kernel
 *   zoomy_jax/fvm/solver2/blocks.py:36-301   dt law, gather update, troubled
 *   zoomy_jax/fvm/ode.py:14-30               RK1 / SSP-RK2
 *   zoomy_jax/fvm/nonconservative_flux.py    path-conservative Rusanov
 *   zoomy_jax/fvm/solver_imex_jax.py:268-342 local Newton for the source
 *
 * State rows are always (b, h, hu, hv); a 1-D mesh has n_y = 0 everywhere so
 * the hv row stays identically zero.
 */
#ifndef ZJ_ORACLE_H
#define ZJ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ZJO_BC_EXTRAP = 0, ZJO_BC_WALL = 1, ZJO_BC_PERIODIC = 2,
       ZJO_BC_DIRICHLET = 3 };
enum { ZJO_LIM_NONE = 0, ZJO_LIM_BJ = 1, ZJO_LIM_MINMOD = 2, ZJO_LIM_VK = 3,
       ZJO_LIM_VANLEER = 4 };
enum { ZJO_RIEMANN_RUSANOV_NC = 0, ZJO_RIEMANN_RUSANOV_POS = 1,
       ZJO_RIEMANN_HLL = 2 };
enum { ZJO_RECON_CONSERVATIVE = 0, ZJO_RECON_ETA = 1, ZJO_RECON_XZ = 2,
       ZJO_RECON_ETA_HU = 3 };
/* which 1/h the second SSP-RK2 stage's fluxes read: the stage state's own
 * (solver_jax.py:1026-1041) or the step-start aux (the reference's recorded
 * order-2 goldens were made with the latter) */
enum { ZJO_STAGE_AUX_REFRESH = 0, ZJO_STAGE_AUX_FROZEN = 1 };
enum { ZJO_POS_NONE = 0, ZJO_POS_ZHANG_SHU = 1 };

typedef struct {
    int dim, n_cells, n_faces, n_bf, max_st;
    const int *face_cells;       /* (2, F), B = -1 on boundary faces      */
    const double *face_normals;  /* (2, F), unit, A -> B / outward         */
    const double *face_volumes;  /* (F)                                    */
    const double *face_centers;  /* (2, F)                                 */
    const double *cell_centers;  /* (2, N)                                 */
    const double *cell_volumes;  /* (N)                                    */
    const double *cell_inradius; /* (N)                                    */
    const int *bface_face;       /* (n_bf) face id                         */
    const int *bface_cell;       /* (n_bf) inner cell                      */
    const int *bface_kind;       /* (n_bf) ZJO_BC_*                        */
    const int *bface_partner;    /* (n_bf) partner bface (periodic) or -1  */
    const int *stencil;          /* (N, max_st) LSQ cells, -1 padded       */
    const double *lsq_w;         /* (N, max_st, 2) weights per component  */
    const int *gslot;            /* (N, 2) boundary faces in the stencil   */
    const double *lsq_wg;        /* (N, 2, 2) ghost weights [slot][comp]   */
    const int *bface_mask;       /* (n_bf) Dirichlet rows (bit v), or NULL */
    const double *bface_value;   /* (4, n_bf) Dirichlet values, or NULL    */
    /* (N, max_cf) faces of each cell, -1 padded, or NULL: the OpenMP build
     * turns the face scatters of min/max into per-cell gathers with it */
    int max_cf;
    const int *cell_faces;
} zjo_mesh;

typedef struct {
    double g, cfl, eps_wet, dt_max, manning_n;
    int order, limiter, riemann, recon;
    /* eta reconstruction only (solver_jax.py:205-217): Zhang-Shu
     * deviation cap, and the front pre-detector tolerance (<= 0: off) */
    int positivity;
    double front_tol;
    int stage_aux;               /* ZJO_STAGE_AUX_* */
} zjo_params;

double zjo_hinv(double h, double eps);
void zjo_bc_ghost(int kind, const double *q, const double *n, double *out);

/* LSQ gradient of one field: grad (2, N).  ubf = ghost values per boundary
 * face (NULL: extrapolation, i.e. zero ghost deltas). */
void zjo_lsq_grad(const zjo_mesh *m, const double *u, const double *ubf,
                  double *grad);

/* Face states (4, F) each; lim_grad (4, 2, N) or NULL. */
void zjo_reconstruct(const zjo_mesh *m, const zjo_params *p, const double *Q,
                     double *QL, double *QR, double *lim_grad);

/* One face: conservative flux F and fluctuations Dp, Dm (4 each). */
void zjo_face_flux(const zjo_params *p, const double *qL, const double *qR,
                   double hinvA, double hinvB, const double *n, double *F,
                   double *Dp, double *Dm);

/* Spatial residual dQ (4, N) of state Q (4, N).  Returns the troubled count
 * of nothing (flux only). */
void zjo_flux_operator(const zjo_mesh *m, const zjo_params *p, const double *Q,
                       double *dQ);

/* The same with a per-cell first-order mask (the flux operator's
 * force_o1, solver_jax.py:724-785; honoured by the eta reconstructions). */
void zjo_flux_operator_o1(const zjo_mesh *m, const zjo_params *p, const double *Q,
                          const unsigned char *force_o1, double *dQ);

/* dt law; t_remaining < 0 means "no cap". */
double zjo_compute_dt(const zjo_mesh *m, const zjo_params *p, const double *Q,
                      double t_remaining);

/* One full step (hyperbolic RK1 / SSP-RK2, then explicit source RK1).
 * Returns number of troubled cells (h < 0 or non-finite) in the result. */
int zjo_step(const zjo_mesh *m, const zjo_params *p, double dt,
             const double *Q, double *Qout);

/* Aux refresh (post_step): rows per horizontal direction d:
 * [d q_d/dx_d, d h/dx_d], then [d b/dx_d] per d, then hinv. */
int zjo_step_mood(const zjo_mesh *m, const zjo_params *p, double dt,
                  const double *Q, double *Qout, int *n_demoted);
void zjo_update_aux(const zjo_mesh *m, const zjo_params *p, const double *Q,
                    double *Qaux);

/* Neumaier-compensated sum of h * |cell|. */
double zjo_mass(const zjo_mesh *m, const double *Q);

/* Manning source S (4) at state q with aux hinv. */
void zjo_manning_source(const zjo_params *p, const double *q, double hinv,
                        double *S);

/* Local implicit source stage (solver_imex_jax.py:268-342): per cell
 * `maxiter` Newton iterations on Q - Qexp - dt S(Q) = 0. */
void zjo_imex_local(const zjo_mesh *m, const zjo_params *p, double dt,
                    int maxiter, const double *Qexp, double *Qout);

/* March n_steps adaptive steps (or to t_end when n_steps < 0).  Returns the
 * number of steps taken; *t_io is advanced. */
/* Threads of the OpenMP build (n <= 0: leave as is); returns the count in
 * effect (1 for the serial build). */
int zjo_set_threads(int n);

int zjo_march(const zjo_mesh *m, const zjo_params *p, double *Q, double *t_io,
              double t_end, int n_steps);

#ifdef __cplusplus
}
#endif
#endif
