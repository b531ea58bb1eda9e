/*
 * poromec_hip.h -- C ABI of libpmhip, the MI355X (gfx950) implementation of
 * poroMec's mef Biot u-p solve path: hexa20/8 element assembly, the sparse
 * matrix-vector product, the Krylov solvers with their dot / axpy / Jacobi
 * steps, and the backward-Euler time step around them.
 *
 * Plain pointers and sizes only.  Host arrays use the reference's layouts
 * (Fortran column-major, 1-based node and equation numbers):
 *   ix    (21, numel)  int32  element nodes 1..20 + material id (column 21)
 *   x     (3, nnodev)  double vertex coordinates
 *   id    (4, nnode)   int32  restraint flags in, equation numbers out
 *   u     (4, nnode)   double u, v, w, p per node
 *   e     (15, numat)  double material constants of elmt17_pm
 *   ia    (neq+1)      int64  1-based row pointers of the strict lower CSRC
 *   ja    (nad)        int32  1-based columns
 *
 * Every call returns PM_OK (0) or a negative PM_ERR_* code; pm_error_string
 * names it.  A context is owned by one host thread at a time.
 */
#ifndef POROMEC_HIP_H
#define POROMEC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_ctx pm_ctx;

/* error codes: these replace the reference's stop_mef() exits */
#define PM_OK 0
#define PM_ERR_HIP (-1)          /* HIP runtime error (no device, out of memory, ...) */
#define PM_ERR_ARG (-2)          /* bad argument or size */
#define PM_ERR_ZERO_DIAG (-3)    /* zero diagonal coefficient (Iters.f:401-406) */
#define PM_ERR_MAXIT (-4)        /* no convergence in maxit (Iters.f:499-501) */
#define PM_ERR_BREAKDOWN (-5)    /* Krylov breakdown (Iters.f:2955-2991, 6260-6307) */
#define PM_ERR_JACOBIAN (-6)     /* element Jacobian det <= 1e-14 (Elmlib.f:2363-2366) */
#define PM_ERR_NLIT (-7)         /* nonlinear loop reached maxnlit (Main.f:1046-1053) */
#define PM_ERR_STATE (-8)        /* call out of order (no mesh, no matrix, ...) */
#define PM_ERR_UNSUPPORTED (-9)  /* feature outside this path (element type, load type) */

/* solver ids (Solv.f dispatch) and preconditioner ids (Precond.f) */
#define PM_SOLVER_PCG 1
#define PM_SOLVER_GMRES 2
#define PM_SOLVER_BICGSTAB 4
#define PM_SOLVER_MINRES 7
#define PM_SOLVER_SQMR 11
#define PM_PRECOND_DIAG 2
#define PM_PRECOND_ABSDIAG 5

/* arithmetic of the Krylov solve */
#define PM_PREC_FP64 0   /* everything in FP64 (the oracle form) */
#define PM_PREC_FP32IR 1 /* FP32 matrix values and Jacobi inside FP64 iterative refinement */

typedef struct {
  int solver;      /* PM_SOLVER_* (reference default 11) */
  int precond;     /* PM_PRECOND_* (reference default 2) */
  int maxit;       /* Krylov iteration limit (reference default 50000) */
  int nkrylov;     /* GMRES restart length ngram (reference default 150) */
  double tol;      /* relative tolerance of the Krylov stop test */
  int precision;   /* PM_PREC_* */
  double ir_inner_tol; /* FP32IR: relative tolerance of each inner solve */
  int ir_maxouter; /* FP32IR: refinement sweeps */
  int maxnlit;     /* nonlinear iterations per step (nl macro, default 500) */
  double nltol;    /* nonlinear tolerance (nl macro, default 1e-4) */
  int stop_crit;   /* cal_residuo_pm criterion: 1 or 2 */
  int maxnlit_accept; /* 0: reaching maxnlit is PM_ERR_NLIT, as stop_mef at
                         Main.f:1046-1053; nonzero: the step ends there and is
                         kept (the goto 420 that code keeps commented out, the
                         behaviour the decks shipped with the reference's
                         outputs were run with) */
} pm_solver_cfg;

typedef struct {
  int iters;        /* Krylov iterations (FP32IR: inner iterations, all sweeps) */
  int outer;        /* FP32IR refinement sweeps (0 for FP64) */
  int used_solver;  /* 2 when SQMR fell back to GMRES */
  double resid;     /* final recurrence residual of the stop test */
  double true_rel;  /* ||b - A x|| / ||b|| recomputed in FP64 */
  double ms;        /* device time of the solve, milliseconds */
  int spmv64, spmv32; /* SpMV launches by precision */
} pm_solve_info;

typedef struct {
  int nlit;             /* nonlinear iterations */
  int nsolve;           /* linear solves */
  int iters[16];        /* Krylov iterations of each solve */
  int used[16];         /* solver that finished each solve */
  double true_rel[16];  /* FP64 relative residual of each solve */
  double ms_total;      /* device time of the step */
  double ms_solve;      /* device time inside the solves */
  double ms_element;    /* device time inside element loops */
} pm_step_info;

const char *pm_error_string(int code);
/* Text of the last HIP runtime error seen by this thread (PM_ERR_HIP). */
const char *pm_last_error(void);

/* Creates a context on HIP device `device` (one stream, default priority). */
int pm_ctx_create(pm_ctx **ctx, int device);
void pm_ctx_destroy(pm_ctx *ctx);

/* ---- set-up: the mesh macro (Main.f:481-870) ----
 * hexa8 input (ix columns 9..20 zero) is promoted to hexa20/8 as the deck
 * reader does (Rdata.f:479-521, mk_elconn_quad_v1 Adjacency.f:583-693), the
 * mid-edge restraints follow mk_bound_quad (Rdata.f:2337-2405), nodes are
 * RCM-ordered (reord/genrcm_v1, Reord.f) when reorder != 0, equations are
 * numbered by numeqpmec2 (Numeq.f:86-151), and the matrix pattern is built
 * on the device.  Replaces rdat_pm's element block + reord + numeqpmec2 +
 * datastruct_pm (Main.f:496-771). */
int pm_mesh_setup(pm_ctx *ctx, int numel, int nnodev, const int *ix, const double *x,
                  const int *idv, int reorder, int *nnode, int *neq, int64_t *nad);

/* The mesh macro for either element family, by the mesh's maxno (nen):
 * nen = 20 is pm_mesh_setup (ix (21, numel), hexa8 input promoted to the
 * hexa20/8 element); nen = 8 keeps a hexa8 mesh as read (ix (9, numel): 8
 * vertices + material), as the reader sizes a maxno-8 mesh (nst = nenv *
 * ndf = 32, Rdata.f:650-660), for the equal-order Q1/Q1 Biot element --
 * element type 7, the ELMT07_PM the reference reserves without implementing
 * it (Elmlib03.f:22-23).  Every node is then a vertex with u, v, w, p; the
 * element is elmt17_pm's weak form with trilinear displacements and the
 * 2 x 2 x 2 Gauss rule.  On such a mesh pm_set_materials takes element-7
 * materials: e(1..8) as for elmt17_pm, and e(9) = beta, the element's
 * pressure stabilization (0: none; Kpp gains -tau int grad N . grad N, tau =
 * beta h^2 / (lambda + 2 mu), h the cube root of the element's volume: the
 * flow-equation perturbation of Aguilar et al., IJNME 75 (2008) 1282-1300,
 * whose 1-D value beta = 1/4 removes the equal-order pressure oscillation of
 * a near-undrained step; pm_adjoint refuses stabilized materials);
 * pm_set_vprop accepts only ie = 7 with vprop off,
 * and the reference summation order (pm_set_ordered) is refused by the
 * element loop (PM_ERR_UNSUPPORTED); pm_get_numbering returns its ix in the
 * (21, numel) layout with columns 9..20 zero. */
int pm_mesh_setup_nen(pm_ctx *ctx, int nen, int numel, int nnodev, const int *ix,
                      const double *x, const int *idv, int reorder, int *nnode, int *neq,
                      int64_t *nad);

/* The same promotion, restraints, ordering and numbering on the host only
 * (no device needed).  ix20 (21 x numel), id (4 x nnode_max) and perm
 * (nnode_max) receive the results; nnode_max >= nnodev + 12 * numel always
 * suffices.  nad is the size of the reference's strict-lower CSRC. */
int pm_mesh_numbering(int numel, int nnodev, const int *ix, const int *idv, int reorder,
                      int nnode_max, int *ix20, int *id, int *perm, int *nnode, int *neq,
                      int64_t *nad);

/* pm_mesh_numbering by the mesh's nen, as pm_mesh_setup_nen reads it (nen =
 * 8: ix (9, numel), no promotion, nnode = nnodev; ix20 still (21, numel)). */
int pm_mesh_numbering_nen(int nen, int numel, int nnodev, const int *ix, const int *idv,
                          int reorder, int nnode_max, int *ix20, int *id, int *perm, int *nnode,
                          int *neq, int64_t *nad);

/* The promoted connectivity (21 x numel), equation numbers (4 x nnode) and
 * the RCM permutation (nnode, perm[i] = node at position i+1), any may be NULL. */
int pm_get_numbering(pm_ctx *ctx, int *ix20, int *id, int *perm);

/* Materials, face loads, load table and gravity (materials / elmtloads /
 * loads / gravity macros, Rdata.f:1132-1180, 1487, 3543).  Element type 17
 * (elmt17_pm) only.  load: load(3,200), fload: fload(1000,5,200) of load.fi.
 * Like the mesh macro (initial_porosity, Main.f:820-828) it also sets the
 * nodal porosity to the vertex average of e(8). */
int pm_set_materials(pm_ctx *ctx, int numat, const double *e, const int *eload,
                     const int *eloadp, const int *load, const double *fload,
                     const double *gravity, double gravity_mod);

/* Variable properties at the integration points (elmt18_pm,
 * Elmlib03.f:2640-3640): ie (numat) gives each material's element type, 17
 * or 18 (elmt17_pm / elmt18_pm), and vprop[4] the conseq macro's flags
 * (read_constitutive_equation, Rdata.f:2545-2640): vprop, kozeny_carman,
 * density, hashin_shtrikman.  Like check_element (Rdata.f:3703-3760) it
 * requires an elmt18 material exactly when vprop is on.  It then runs
 * initial_prop (Pform.f:1434-1540), which turns e(1), e(2) into the
 * solid's K and mu (e(6) into its density) and sets the 7 properties of
 * every element's 64 points; every step then runs update_prop
 * (Pform.f:1176-1330) after the nonlinear loop, and the nodal porosity
 * becomes the points' porosity extrapolated to the vertices.  Call after
 * pm_set_materials (which resets it); PM_ERR_ARG where ks_and_mus has no
 * solution, PM_ERR_UNSUPPORTED in the batching form. */
int pm_set_vprop(pm_ctx *ctx, const int *ie, const int *vprop);
/* The points' properties (7, 64, numel) in the reference's vpropel layout. */
int pm_get_vprop(pm_ctx *ctx, double *vpropel);

/* Nodal loads and prescribed values: nload (4, nnode) time-function ids and
 * f (4, nnode) values, as pload_pm reads them (Pload.f:1-92).  Optional. */
int pm_set_nodal_loads(pm_ctx *ctx, const int *nload, const double *f);

/* ---- element loop + assembly: pform_pm (Pform.f:1-429) with elmt17_pm
 * (Elmlib03.f:1822-2636) and assbly_pm/csr (Assbly.f:89-304) ----
 * isw 2: stiffness (lhs) and internal force b -= K_e u_e (rhs) from the
 *        device state u; isw 4: b0 -= volume, face and previous-step terms.
 * target 0 assembles the right-hand side into b0, 1 into b. */
int pm_pform(pm_ctx *ctx, int isw, int lhs, int rhs, double dt, double t, int target);

/* ---- linear solve: solv_pm (Solv.f:27-780) on the device matrix ----
 * b: host right-hand side (neq) or NULL to use the device b; x: host
 * solution (neq) or NULL to keep it on the device. */
int pm_solve(pm_ctx *ctx, const pm_solver_cfg *cfg, const double *b, double *x,
             pm_solve_info *info);

/* ---- pm_solve without allocating (no counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops) ----
 * The same solve as pm_solve, with its scratch (Krylov work vectors, the
 * Jacobi vector, reduction partials, device copies of b and x) in a static
 * device array of the library, (19 * PM_NOALLOC_MAX_NEQ + 32768) doubles
 * (80 MB), that each GPU holds from the library's first use on it until
 * the process ends.  It calls no hipMalloc or hipFree.  The matrix must be
 * set up before (pm_mesh_setup + pm_pform, or pm_struct_upload +
 * pm_values_upload).  Limits: neq <= PM_NOALLOC_MAX_NEQ (PM_ERR_ARG beyond),
 * not GMRES and no SQMR -> GMRES fallback (SQMR at maxit returns
 * PM_ERR_MAXIT), not the batching form (PM_ERR_UNSUPPORTED), FP32IR only
 * once the FP32 values exist (PM_ERR_STATE otherwise; a pm_solve with
 * FP32IR or pm_tune_spmv(use32 = 1) makes them).  b / x as in pm_solve; NULL
 * needs the context's device vectors to exist (PM_ERR_STATE otherwise).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU race. */
#define PM_NOALLOC_MAX_NEQ 524288
int pm_solve_noalloc(pm_ctx *ctx, const pm_solver_cfg *cfg, const double *b, double *x,
                     pm_solve_info *info);

/* ---- drop-in solv_pm with the caller's CSRC matrix ----
 * Uploads the strict-lower CSRC pattern (ia, ja: Datastruct.f:124-150) and
 * values (ad, al), then solves A x = b with A = diag(ad) + L + L^T.
 * pm_struct_upload keeps the pattern; pm_values_upload replaces values. */
int pm_struct_upload(pm_ctx *ctx, int neq, int64_t nad, const int64_t *ia, const int *ja);
int pm_values_upload(pm_ctx *ctx, const double *ad, const double *al);

/* y = A x (matvec_csrc_sym_pm, Matvec.f:1504-1595), host vectors of neq. */
int pm_spmv(pm_ctx *ctx, const double *x, double *y);

/* The device matrix as the reference's CSRC: ia (neq+1), ja (nad), ad (neq),
 * al (nad); any may be NULL. */
int pm_export_csrc(pm_ctx *ctx, int64_t *ia, int *ja, double *ad, double *al);

/* ---- time march: the solv macro (Main.f:876-1123) ----
 * State arrays: u, u0 (4, nnode), dp, porosity (nnode), time t. */
int pm_set_state(pm_ctx *ctx, const double *u, const double *u0, const double *dp,
                 const double *porosity, double t);
int pm_get_state(pm_ctx *ctx, double *u, double *u0, double *dp, double *porosity,
                 double *t);
/* Device vectors b0, b, x0 (neq), any may be NULL. */
int pm_get_vectors(pm_ctx *ctx, double *b0, double *b, double *x0);
int pm_set_vectors(pm_ctx *ctx, const double *b0, const double *b, const double *x0);
/* One backward-Euler step of size dt: pload_pm, pform isw=4, predictor,
 * nonlinear loop (pform isw=2, cal_residuo_pm, solv_pm, delta_update_pm),
 * porosity_form, update_res_v2. */
int pm_step(pm_ctx *ctx, double dt, const pm_solver_cfg *cfg, pm_step_info *info);

/* Vertex stresses and Darcy flux of the current state, what the pnsf macro
 * prints: tform_pm (Pform.f:433-700) with elmt17_pm isw 3
 * (Elmlib03.f:2206-2256, darcy_flux Elmlib.f:3178-3200) and
 * effective_stress (Pform.f:2487-2503).  tx, txe, txb: (6, nnode) total
 * stress D eps - alpha dp, Terzaghi's tx + p and Biot's tx + alpha p,
 * components xx yy zz xy yz xz; flux: (3, nnode) perm (rho_f g - grad p).
 * Each vertex gets the average over its elements, summed in element order;
 * other nodes get zero (txe: the pressure shift).  Any pointer may be NULL. */
int pm_tform(pm_ctx *ctx, double *tx, double *txe, double *txb, double *flux);

/* ---- reference summation order (no counterpart in the reference: it is the
 * reference's own arithmetic order) ----
 * on != 0: every dot product and norm of the solvers and of the residual
 * test is summed sequentially in equation order (dot_par, Matvec.f:1714-1743,
 * one rank), and the SpMV sums each row as matvec_csrc_sym_pm does (diagonal,
 * lower columns ascending, transposed terms ascending; Matvec.f:1504-1595).
 * With a matrix whose two triangles hold identical values (pm_values_upload)
 * the solvers then reproduce the reference's iterates bit for bit.  The
 * reductions run on one thread: for verification, not for speed.  Not used
 * by the batching form. */
int pm_set_ordered(pm_ctx *ctx, int on);

/* ---- batching form (no counterpart in the reference: the reference runs one
 * problem per process) ----
 * The mesh given to pm_mesh_setup holds nbox disconnected boxes, box b owning
 * elements [elem_off[b], elem_off[b+1]) and, after RCM, a contiguous range
 * of equations.  Afterwards pm_solve and pm_step treat every box as its own
 * system -- own Krylov scalars, own Krylov and nonlinear stop tests -- while
 * all boxes share each kernel launch.  Solvers: PCG and SQMR. */
int pm_batch_define(pm_ctx *ctx, int nbox, const int *elem_off);
/* Krylov iterations of every box in the last solve (nbox). */
int pm_batch_iters(pm_ctx *ctx, int *iters);
/* 1 (default): replay the iteration chunks of batched solves from a captured
 * hipGraph; 0: launch them one by one. */
int pm_batch_graphs(pm_ctx *ctx, int on);

/* ---- adjoint form (no counterpart in the reference: it has no derivatives) ----
 * Marches nsteps backward-Euler steps of size dt from the current state (as
 * nsteps calls of pm_step), measuring the misfit
 *   J = sum_n sum_i (u[sensor_dof[i]] after step n - target[n * nsens + i])^2
 * where sensor_dof[i] = 4 * node + dof names a free dof of the nodal state
 * (dof 3 = pressure), then solves the adjoint equations backwards in time
 * and returns dJ/dlog(e(1)) (Young's modulus) in grad_logE and
 * dJ/dlog(e(3)) (permeability) in grad_logk, one entry per material (numat;
 * either may be NULL).  Linear elastic materials, one solve per step
 * (tight solver tolerance); forward states are recomputed from checkpoints
 * taken every ckpt_every steps.  Not available in the batching form.  On an
 * error return the nodal state is that of some step of the march. */
int pm_adjoint(pm_ctx *ctx, int nsteps, double dt, const pm_solver_cfg *cfg, int nsens,
               const int *sensor_dof, const double *target, int ckpt_every, double *J,
               double *grad_logE, double *grad_logk);

/* ---- measurement hooks (no counterpart in the reference) ----
 * Sizes of the device matrix: equations, row blocks (free nodes), stored
 * values (full matrix) and shared column indices. */
int pm_matrix_info(pm_ctx *ctx, int *neq, int *nblocks, int64_t *nnz, int64_t *ncols);
/* Runs of <= 4 consecutive columns the run-per-lane SpMV shapes read in
 * place of the column indices (one 32-bit word each). */
int pm_matrix_runs(pm_ctx *ctx, int64_t *nruns);
/* Average device time of one SpMV launch (FP64 values, or FP32 values when
 * use32 != 0) over reps launches, timed with HIP events on the library's
 * stream. */
int pm_bench_spmv(pm_ctx *ctx, int use32, int reps, double *ms_per_launch);
/* SpMV shape tuner (no counterpart in the reference): times each of the
 * PM_SPMV_SHAPES shapes on the current matrix over `reps` launches: 0-5
 * (lanes per node block, columns in flight per lane), 6-8 the same with the
 * column indices loaded one block ahead, 9-10 one or two runs of <= 4
 * consecutive columns per lane, 11-12 the runs and 13 the (64, 5) prefetch
 * shape with the workgroups of each XCD sweeping its blocks together, 14
 * the sweeping runs with one run per lane in blocks of <= 64 runs and two in
 * wider blocks, 15 the sweeping runs with 32 lanes per block (two blocks per
 * wavefront; for the Q1/Q1 element's 27-run blocks; the batching form runs
 * shape 5's kernel for it).  It keeps the fastest for FP64
 * (use32 = 0) or FP32 values, and returns its index and, if ms_per_shape is
 * not NULL, every shape's average launch time.  The batching form tunes its
 * own kernel.  Shapes sum in different orders: results move at the rounding
 * level when the shape changes, and stay fixed until the next change. */
#define PM_SPMV_SHAPES 16
int pm_tune_spmv(pm_ctx *ctx, int use32, int reps, int *best, double *ms_per_shape);
/* Selects SpMV shape `shape` (0 .. PM_SPMV_SHAPES-1) without timing. */
int pm_set_spmv_shape(pm_ctx *ctx, int use32, int shape);
/* Live SpMV timing: with on = k > 0 every k-th SpMV the single-system
 * solvers launch is bracketed by HIP events on the library's stream (launches
 * skipped after convergence are not counted; on = 0 switches it off).  stats
 * (may be NULL) receives what was accumulated since the previous call --
 * {ms FP64, launches timed FP64, ms FP32, launches timed FP32} -- and the
 * accumulators restart from zero. */
int pm_spmv_timing(pm_ctx *ctx, int on, double *stats);
/* Average device time of a plain copy kernel over `bytes` bytes (reads
 * bytes, writes bytes): the reference point for HBM traffic counters. */
int pm_bench_copy(pm_ctx *ctx, int64_t bytes, int reps, double *ms_per_launch);

/* Average device time of a streaming kernel over `bytes` bytes (no
 * counterpart in the reference): what HBM delivers on this device, the
 * reference point of the SpMV's bandwidth figures.  kind 0: copy, 32 B per
 * lane, grid-stride (pm_bench_copy's kernel); 1: read-only, 4 x 16 B in
 * flight per lane; 2: the same with nontemporal loads; 3: copy, 4 x 16 B per
 * lane; 4: write-only, 4 x 16 B per lane; 5: read-only, 8 x 16 B per lane.
 * Kinds 1-5 launch one workgroup of 256 lanes per 16 KB (32 KB for kind 5).
 * A copy moves `bytes` in and `bytes` out; a read or a fill `bytes` once. */
int pm_bench_stream(pm_ctx *ctx, int kind, int64_t bytes, int reps, double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
