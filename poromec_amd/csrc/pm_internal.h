// pm_internal.h -- internal definitions of libpmhip (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/poromec_hip.h"

#define PM_NEN 20
#define PM_NENV 8
#define PM_NDF 4
#define PM_NST 68
#define PM_PROP 15
#define PM_NUMLOAD 200
#define PM_NTERMS 5
#define PM_NPAR 1000

#define PM_CHECK_HIP(expr)                                  \
  do {                                                      \
    hipError_t e__ = (expr);                                \
    if (e__ != hipSuccess) {                                \
      pm_set_last_error(hipGetErrorString(e__), __FILE__, __LINE__); \
      return PM_ERR_HIP;                                    \
    }                                                       \
  } while (0)

#define PM_TRY(expr)             \
  do {                           \
    int rc__ = (expr);           \
    if (rc__ != PM_OK) return rc__; \
  } while (0)

void pm_set_last_error(const char *msg, const char *file, int line);

// ---------------------------------------------------------------------------
// host structure (pm_structure.cpp)
// ---------------------------------------------------------------------------
struct PmHostMesh {
  int numel = 0, nnodev = 0, nnode = 0, neq = 0;
  int nen = 20;             // 20: hexa20/8 (promoted mesh); 8: Q1/Q1, nodes = vertices
  int64_t nad = 0;          // strict-lower count (reference CSRC size)
  std::vector<int> ix20;    // (21, numel) 1-based, reference layout (nen 8: columns 9..20 zero)
  std::vector<int> idflags; // (4, nnode) restraint flags
  std::vector<int> id;      // (4, nnode) reference equation numbers (1-based, 0 none)
  std::vector<int> perm;    // (nnode) RCM order, 1-based node ids
  std::vector<int> fnno;    // (nnode) 1 for vertices
  // node graph (graph_v1): 0-based CSR over nodes, neighbours sorted by id
  std::vector<int64_t> gptr;
  std::vector<int> gadj;
  // element colouring
  std::vector<int> colour_elems;  // elements grouped by colour (0-based)
  std::vector<int> colour_ptr;    // ncolour+1
};

// Node-shared-column matrix layout ("node blocks"): every free node k (in
// equation order) owns nr[k] consecutive rows starting at fr[k]; all its rows
// share the column list ncol[colptr[k] .. colptr[k+1]) and row q's values
// sit at val[valptr[k] + q * c_k + j], c_k = colptr[k+1] - colptr[k].
struct PmHostPattern {
  int nrn = 0;                 // free nodes (row blocks)
  int neq = 0;
  std::vector<int> fr, nr;     // per row block
  std::vector<int64_t> colptr; // nrn+1
  std::vector<int64_t> valptr; // nrn+1
  std::vector<int> ncol;       // 0-based equation ids
  std::vector<int> eq_block;   // per equation: its row block
};

int pm_host_promote(PmHostMesh &m, const int *ix8);
// a maxno-8 mesh as read: ix9 (9, numel) = 8 vertices + material, no
// promotion, every node a vertex (the Q1/Q1 element)
int pm_host_plain_hexa8(PmHostMesh &m, const int *ix9);
void pm_host_bound_quad(PmHostMesh &m);
void pm_host_graph(PmHostMesh &m);
void pm_host_rcm(PmHostMesh &m, int reorder);
void pm_host_numeq(PmHostMesh &m);
int pm_host_pattern(const PmHostMesh &m, PmHostPattern &p);

// element-order assembly maps (see pm_host_ordered_maps)
struct PmOrderedMaps {
  std::vector<int64_t> mptr, msrc;  // per stored value: e*68*68 + a*68 + b
  std::vector<int64_t> rptr, rsrc;  // per equation: e*68 + a
  std::vector<int64_t> pptr, psrc;  // per node: e*8 + vertex
};
int pm_host_ordered_maps(const PmHostMesh &m, const PmHostPattern &p, PmOrderedMaps &o);
int pm_host_pattern_from_csrc(int neq, int64_t nad, const int64_t *ia, const int *ja,
                              PmHostPattern &p);
void pm_host_colour(PmHostMesh &m);
int64_t pm_host_count_lower(const PmHostPattern &p);

// ---------------------------------------------------------------------------
// device context
// ---------------------------------------------------------------------------
struct PmScalars;  // device-side Krylov scalars (pm_device.h)
struct PmBatch;    // batching form (pm_batch.hip)
void pm_batch_free(PmBatch *b);

struct PmDevMatrix {
  int nrn = 0, neq = 0;
  int64_t nnz = 0, ncols = 0;
  int *fr = nullptr, *nr = nullptr, *ncol = nullptr, *eq_block = nullptr;
  int64_t *colptr = nullptr, *valptr = nullptr;
  // column runs of <= 4 consecutive columns per block: chunk[] holds
  // (first column << 2) | (length - 1), chptr[] each block's first run
  int64_t nchunks = 0;
  int64_t *chptr = nullptr;
  unsigned int *chunk = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;
  double *diag = nullptr;  // ad
  int64_t *diagpos = nullptr; // position of the diagonal entry of each row
  bool have32 = false;
  bool values_valid = false;
};

struct pm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cu_count = 256;

  PmHostMesh mesh;
  bool have_mesh = false;
  bool have_materials = false;
  PmDevMatrix A;
  bool have_matrix = false;

  // device mesh
  int *d_ix = nullptr;      // (numel, 20) 0-based nodes
  double *d_xv = nullptr;   // (nnodev, 3)
  int *d_eq = nullptr;      // (nnode, 4) 0-based equation or -1
  unsigned char *d_isp = nullptr; // (neq) 1 for pressure equations
  int *d_elem_err = nullptr; // element error flag
  std::vector<int64_t> csrc_ia; // caller's lower pattern (pm_struct_upload)
  std::vector<int> csrc_ja;
  int *d_mat = nullptr;     // (numel) 0-based material
  double *d_e = nullptr;    // (numat, 15)
  int numat = 0;
  int *d_eload = nullptr;   // (numel, 14): mechanical 7, hydraulic 7
  int has_face_loads = 0;
  int *d_load = nullptr;    // (200, 3)
  double *d_fload = nullptr;// (200, 5, 1000)
  std::vector<int> h_load;
  std::vector<double> h_fload;
  double gravity[3] = {0, 0, 0};
  double gravity_mod = 10.0;
  int *d_colour_elems = nullptr;
  std::vector<int> colour_ptr;
  int *d_ic = nullptr;      // (nnode) elements incident to each vertex
  // element-order assembly (reference summation order), built on first use
  int64_t *d_omptr = nullptr, *d_omsrc = nullptr, *d_orptr = nullptr, *d_orsrc = nullptr;
  int64_t *d_opptr = nullptr, *d_opsrc = nullptr;
  double *d_ke = nullptr, *d_pe = nullptr, *d_pv = nullptr;  // element arrays
  bool have_ordered_maps = false;
  // variable properties (elmt18_pm): the conseq flags, element type per
  // material, the 7 properties of every element's 64 integration points,
  // and the materials as pm_set_materials received them (initial_prop
  // rewrites the device copy)
  int vprop[4] = {0, 0, 0, 0};
  int *d_ie = nullptr;
  double *d_vp = nullptr;
  std::vector<double> h_e;
  // stress recovery: (element, corner) pairs per node and the outputs
  int64_t *d_tf_ptr = nullptr;
  int *d_tf_src = nullptr;
  double *d_tf_out = nullptr;  // tx, txe, txb (6 x nnode each), flux (3 x nnode)

  // nodal loads
  int *d_nload = nullptr;   // (nnode, 4)
  double *d_f = nullptr;    // (nnode, 4)
  bool have_nodal_loads = false;

  // state
  double *d_u = nullptr, *d_u0 = nullptr, *d_dp = nullptr, *d_poro = nullptr;
  double *d_b0 = nullptr, *d_b = nullptr, *d_x0 = nullptr, *d_m = nullptr;
  double *d_dporo = nullptr;
  double t = 0.0;

  // solver workspace
  std::vector<void *> work;   // freed at destroy
  double *d_wk[16] = {nullptr};
  int wk_n = 0;               // length of each work vector
  double *d_gm = nullptr;     // GMRES basis
  int gm_k = 0;
  double *d_partials = nullptr;
  unsigned int *d_counter = nullptr;
  PmScalars *d_scal = nullptr;
  PmScalars *h_scal = nullptr; // pinned mirror
  int red_blocks = 1024;
  // live SpMV timing (measurement hook): an event pair around every SpMV a
  // single-system solver launches, harvested after each host sync; pairs
  // tagged with an iteration past convergence (launched but skipped by the
  // kernels' done check) are dropped
  int time_spmv = 0;                // 0 off, k: every k-th SpMV launch
  long long spmv_seq = 0;
  bool tev_open = false;
  int cur_iter = 0;                 // iteration tag of the next launches
  std::vector<hipEvent_t> tev;      // 2 per pending pair
  std::vector<int> tev_kind, tev_iter;
  int tev_used = 0;
  double spmv_ms[2] = {0.0, 0.0};   // FP64, FP32 values
  long long spmv_n[2] = {0, 0};
  // SpMV shape (index into the tuner's table) per value precision
  int spmv_cfg[2] = {2, 2};
  int spmv_tuned[2] = {0, 0};
  // reference summation order: every reduction a sequential left-to-right
  // sum, the SpMV in the order of matvec_csrc_sym_pm (slow; for parity)
  int ordered = 0;
  // set for the duration of pm_solve_noalloc: the solve must not allocate,
  // so SQMR reaching maxit returns PM_ERR_MAXIT instead of falling back to
  // GMRES (whose basis would need (ngram + 1) more vectors)
  bool no_alloc = false;

  // batching form
  PmBatch *batch = nullptr;
  int *d_eq_box = nullptr;  // (neq) box of every equation

  ~pm_ctx();
};

int pm_upload_pattern(pm_ctx *c, const PmHostPattern &p);
int pm_dev_alloc(pm_ctx *c, void **ptr, size_t bytes);
int pm_ensure_workspace(pm_ctx *c, int n);

// kernels entry points (pm_blas.hip, pm_krylov.hip, pm_element.hip, pm_march.hip)
int pm_spmv_dev(pm_ctx *c, const double *x, double *y, int use32);
int pm_spmv_fused_dev(pm_ctx *c, const double *x, double *y, int use32);
int pm_solve_dev(pm_ctx *c, const pm_solver_cfg *cfg, const double *b, double *x,
                 pm_solve_info *info);
int pm_pform_dev(pm_ctx *c, int isw, int lhs, int rhs, double dt, double t, double *bvec);
int pm_porosity_dev(pm_ctx *c, double dt);
int pm_tform_dev(pm_ctx *c, double *tx, double *txe, double *txb, double *flux);
int pm_update_prop_dev(pm_ctx *c, double dt);
int pm_extract_diag(pm_ctx *c);
int pm_make_val32(pm_ctx *c);
int pm_step_dev(pm_ctx *c, double dt, const pm_solver_cfg *cfg, pm_step_info *info);
double pm_dot_host(pm_ctx *c, const double *a, const double *b, int n, int *rc);
int pm_batch_define_dev(pm_ctx *c, int nbox, const int *elem_off);
int pm_batch_nbox(pm_ctx *c);
int pm_batch_last_iters(pm_ctx *c, int *iters);
void pm_batch_set_graph(pm_ctx *c, int on);
int pm_batch_krylov(pm_ctx *c, int solver, int maxit, double tol, const double *b, double *x,
                    const double *m, int use32, pm_solve_info *info,
                    int cap_stragglers = 0);
int pm_batch_resnorms(pm_ctx *c, const double *b, std::vector<double> &ru2,
                      std::vector<double> &rp2);
int pm_batch_mask_rhs(pm_ctx *c, double *b, const std::vector<int> &active);
// boxes whose last batched solve ended unconverged (1) or converged (0)
int pm_batch_unconverged(pm_ctx *c, std::vector<int> &open);
int pm_batch_true_res(pm_ctx *c, const double *b, const double *x, double *res,
                      std::vector<double> &r2, std::vector<double> &b2);
int pm_adjoint_dev(pm_ctx *c, int nsteps, double dt, const pm_solver_cfg *cfg, int nsens,
                   const int *sensor_dof, const double *target, int ckpt_every, double *J,
                   double *grad_logE, double *grad_logk);
void pm_spmv_harvest(pm_ctx *c, int done, int iter);
int pm_batch_spmv_dev(pm_ctx *c, const double *x, double *y, int use32);
