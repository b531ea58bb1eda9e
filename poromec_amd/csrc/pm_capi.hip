// pm_capi.hip -- context, device data management and the C ABI of libpmhip
// (declared in include/poromec_hip.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "pm_device.h"
#include "pm_spmv.h"

static thread_local char g_last_error[512] = "";

void pm_set_last_error(const char *msg, const char *file, int line) {
  snprintf(g_last_error, sizeof(g_last_error), "%s (%s:%d)", msg, file, line);
}

extern "C" const char *pm_last_error(void) { return g_last_error; }

extern "C" const char *pm_error_string(int code) {
  switch (code) {
    case PM_OK: return "ok";
    case PM_ERR_HIP: return "HIP runtime error";
    case PM_ERR_ARG: return "invalid argument";
    case PM_ERR_ZERO_DIAG: return "zero diagonal coefficient";
    case PM_ERR_MAXIT: return "no convergence within maxit";
    case PM_ERR_BREAKDOWN: return "Krylov breakdown";
    case PM_ERR_JACOBIAN: return "element Jacobian determinant <= 0";
    case PM_ERR_NLIT: return "nonlinear loop reached maxnlit";
    case PM_ERR_STATE: return "call out of order";
    case PM_ERR_UNSUPPORTED: return "unsupported feature";
    default: return "unknown error";
  }
}

#define FREE(p)       \
  do {                \
    if (p) (void)hipFree(p); \
    p = nullptr;      \
  } while (0)

static void free_matrix(PmDevMatrix &A) {
  FREE(A.fr);
  FREE(A.nr);
  FREE(A.ncol);
  FREE(A.eq_block);
  FREE(A.colptr);
  FREE(A.valptr);
  FREE(A.chptr);
  FREE(A.chunk);
  FREE(A.val);
  FREE(A.val32);
  FREE(A.diag);
  FREE(A.diagpos);
  A = PmDevMatrix();
}

pm_ctx::~pm_ctx() {
  for (hipEvent_t e : tev) (void)hipEventDestroy(e);
  FREE(d_omptr);
  FREE(d_omsrc);
  FREE(d_orptr);
  FREE(d_orsrc);
  FREE(d_opptr);
  FREE(d_opsrc);
  FREE(d_ke);
  FREE(d_pe);
  FREE(d_pv);
  FREE(d_tf_ptr);
  FREE(d_tf_src);
  FREE(d_tf_out);
  FREE(d_ie);
  FREE(d_vp);
  if (batch) pm_batch_free(batch);
  FREE(d_eq_box);
  free_matrix(A);
  FREE(d_ix);
  FREE(d_xv);
  FREE(d_eq);
  FREE(d_isp);
  FREE(d_mat);
  FREE(d_e);
  FREE(d_eload);
  FREE(d_load);
  FREE(d_fload);
  FREE(d_colour_elems);
  FREE(d_ic);
  FREE(d_nload);
  FREE(d_f);
  FREE(d_u);
  FREE(d_u0);
  FREE(d_dp);
  FREE(d_poro);
  FREE(d_b0);
  FREE(d_b);
  FREE(d_x0);
  FREE(d_m);
  FREE(d_dporo);
  FREE(d_elem_err);
  for (int i = 0; i < 16; i++) FREE(d_wk[i]);
  FREE(d_gm);
  FREE(d_partials);
  FREE(d_counter);
  FREE(d_scal);
  if (h_scal) hipHostFree(h_scal);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (stream) hipStreamDestroy(stream);
}

template <class T>
static int dalloc(T **p, size_t n) {
  FREE(*p);
  if (n == 0) n = 1;
  PM_CHECK_HIP(hipMalloc((void **)p, n * sizeof(T)));
  return PM_OK;
}
template <class T>
static int dupload(pm_ctx *c, T **p, const T *h, size_t n) {
  PM_TRY(dalloc(p, n));
  if (n) PM_CHECK_HIP(hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return PM_OK;
}

extern "C" int pm_ctx_create(pm_ctx **out, int device) {
  if (!out) return PM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  PM_CHECK_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return PM_ERR_HIP;
  PM_CHECK_HIP(hipSetDevice(device));
  pm_ctx *c = new pm_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
  int rc = PM_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == PM_OK) {
      pm_set_last_error(hipGetErrorString(e), __FILE__, __LINE__);
      rc = PM_ERR_HIP;
    }
  };
  chk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  chk(hipEventCreate(&c->ev0));
  chk(hipEventCreate(&c->ev1));
  chk(hipMalloc(&c->d_partials, sizeof(double) * 8192 * 4));
  chk(hipMalloc(&c->d_counter, sizeof(unsigned int)));
  chk(hipMemset(c->d_counter, 0, sizeof(unsigned int)));
  chk(hipMalloc(&c->d_scal, sizeof(PmScalars)));
  chk(hipMemset(c->d_scal, 0, sizeof(PmScalars)));
  chk(hipHostMalloc((void **)&c->h_scal, sizeof(PmScalars), hipHostMallocDefault));
  chk(hipMalloc(&c->d_elem_err, sizeof(int)));
  if (rc != PM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return PM_OK;
}

extern "C" void pm_ctx_destroy(pm_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  delete c;
}

// ---------------------------------------------------------------------------
// matrix pattern on the device
// ---------------------------------------------------------------------------
int pm_upload_pattern(pm_ctx *c, const PmHostPattern &p) {
  // a batch definition (box ranges, workgroup assignment, captured graph)
  // refers to the matrix it was defined on: it goes with that matrix
  if (c->batch) {
    (void)hipStreamSynchronize(c->stream);
    pm_batch_free(c->batch);
    c->batch = nullptr;
  }
  FREE(c->d_eq_box);
  free_matrix(c->A);
  PmDevMatrix &A = c->A;
  A.nrn = p.nrn;
  A.neq = p.neq;
  A.ncols = p.colptr[p.nrn];
  A.nnz = p.valptr[p.nrn];
  PM_TRY(dupload(c, &A.fr, p.fr.data(), p.fr.size()));
  PM_TRY(dupload(c, &A.nr, p.nr.data(), p.nr.size()));
  PM_TRY(dupload(c, &A.colptr, p.colptr.data(), p.colptr.size()));
  PM_TRY(dupload(c, &A.valptr, p.valptr.data(), p.valptr.size()));
  PM_TRY(dupload(c, &A.ncol, p.ncol.data(), p.ncol.size()));
  PM_TRY(dupload(c, &A.eq_block, p.eq_block.data(), p.eq_block.size()));
  {
    // runs of consecutive columns, cut at 4, for the run-per-lane SpMV
    std::vector<int64_t> chptr((size_t)A.nrn + 1);
    std::vector<unsigned int> chunk;
    chunk.reserve((size_t)(A.ncols / 3 + A.nrn));
    for (int k = 0; k < A.nrn; k++) {
      chptr[k] = (int64_t)chunk.size();
      int64_t j = p.colptr[k];
      const int64_t e = p.colptr[k + 1];
      while (j < e) {
        const unsigned int s0 = (unsigned int)p.ncol[j];
        int len = 1;
        while (len < 4 && j + len < e && (unsigned int)p.ncol[j + len] == s0 + (unsigned int)len) len++;
        chunk.push_back((s0 << 2) | (unsigned int)(len - 1));
        j += len;
      }
    }
    chptr[A.nrn] = (int64_t)chunk.size();
    A.nchunks = (int64_t)chunk.size();
    PM_TRY(dupload(c, &A.chptr, chptr.data(), chptr.size()));
    PM_TRY(dupload(c, &A.chunk, chunk.data(), chunk.size()));
  }
  // 4 values of padding: the run-per-lane SpMV reads a whole 4-value run
  PM_TRY(dalloc(&A.val, (size_t)A.nnz + 4));
  PM_CHECK_HIP(hipMemsetAsync(A.val, 0, sizeof(double) * ((size_t)A.nnz + 4), c->stream));
  PM_TRY(dalloc(&A.diag, (size_t)A.neq));
  PM_TRY(dalloc(&A.diagpos, (size_t)A.neq));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->have_matrix = true;
  return PM_OK;
}

// position of the diagonal entry of every row (binary search in its block)
__global__ void k_diagpos(int neq, const int *eq_block, const int *fr, const int64_t *colptr,
                          const int64_t *valptr, const int *ncol, int64_t *diagpos) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  const int k = eq_block[r];
  const int64_t cp = colptr[k];
  const int cnt = (int)(colptr[k + 1] - cp);
  int lo = 0, hi = cnt;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ncol[cp + mid] < r) lo = mid + 1;
    else hi = mid;
  }
  diagpos[r] = (lo < cnt && ncol[cp + lo] == r) ? valptr[k] + (int64_t)(r - fr[k]) * cnt + lo : -1;
}

__global__ void k_diag(int neq, const int64_t *diagpos, const double *val, double *diag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  const int64_t p = diagpos[r];
  diag[r] = p >= 0 ? val[p] : 0.0;
}

int pm_extract_diag(pm_ctx *c) {
  PmDevMatrix &A = c->A;
  const int g = (A.neq + 255) / 256;
  hipLaunchKernelGGL(k_diagpos, dim3(g), dim3(256), 0, c->stream, A.neq, A.eq_block, A.fr,
                     A.colptr, A.valptr, A.ncol, A.diagpos);
  hipLaunchKernelGGL(k_diag, dim3(g), dim3(256), 0, c->stream, A.neq, A.diagpos, A.val, A.diag);
  PM_CHECK_HIP(hipGetLastError());
  return PM_OK;
}

__global__ void k_to32(int64_t n, const double *a, float *b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = i; k < n; k += stride) b[k] = (float)a[k];
}

int pm_make_val32(pm_ctx *c) {
  PmDevMatrix &A = c->A;
  if (!A.val32) {  // padded as val
    PM_TRY(dalloc(&A.val32, (size_t)A.nnz + 4));
    PM_CHECK_HIP(hipMemsetAsync(A.val32 + A.nnz, 0, 4 * sizeof(float), c->stream));
  }
  hipLaunchKernelGGL(k_to32, dim3(4096), dim3(256), 0, c->stream, A.nnz, A.val, A.val32);
  PM_CHECK_HIP(hipGetLastError());
  A.have32 = true;
  return PM_OK;
}

// CSRC values -> node-block values: every strict-lower entry (r, c) goes to
// (r, c) and (c, r), the diagonal to (r, r)
__global__ void k_csrc_values(int neq, const int64_t *ia, const int *ja, const double *ad,
                              const double *al, const int *eq_block, const int *fr,
                              const int64_t *colptr, const int64_t *valptr, const int *ncol,
                              double *val) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  auto pos = [&](int row, int col) -> int64_t {  // -1 when (row, col) is not stored
    const int k = eq_block[row];
    const int64_t cp = colptr[k];
    const int cnt = (int)(colptr[k + 1] - cp);
    int lo = 0, hi = cnt;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ncol[cp + mid] < col) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= cnt || ncol[cp + lo] != col) return -1;
    return valptr[k] + (int64_t)(row - fr[k]) * cnt + lo;
  };
  int64_t p = pos(r, r);
  if (p >= 0) val[p] = ad[r];
  for (int64_t k = ia[r] - 1; k < ia[r + 1] - 1; k++) {
    const int col = ja[k] - 1;
    const double v = al[k];
    p = pos(r, col);
    if (p >= 0) val[p] = v;
    p = pos(col, r);
    if (p >= 0) val[p] = v;
  }
}

// node-block matrix -> strict lower CSRC (counts, then fill)
__global__ void k_lower_count(int nrn, const int *fr, const int *nr, const int64_t *colptr,
                              const int *ncol, int64_t *cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrn) return;
  const int f = fr[k];
  for (int q = 0; q < nr[k]; q++) {
    int64_t c = 0;
    for (int64_t j = colptr[k]; j < colptr[k + 1]; j++)
      if (ncol[j] < f + q) c++;
    cnt[f + q] = c;
  }
}

__global__ void k_lower_fill(int nrn, const int *fr, const int *nr, const int64_t *colptr,
                             const int64_t *valptr, const int *ncol, const double *val,
                             const int64_t *ia, int *ja, double *al) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrn) return;
  const int f = fr[k];
  const int cnt = (int)(colptr[k + 1] - colptr[k]);
  for (int q = 0; q < nr[k]; q++) {
    int64_t w = ia[f + q] - 1;
    for (int j = 0; j < cnt; j++) {
      const int col = ncol[colptr[k] + j];
      if (col >= f + q) break;
      if (ja) ja[w] = col + 1;
      if (al) al[w] = val[valptr[k] + (int64_t)q * cnt + j];
      w++;
    }
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static int ensure_state(pm_ctx *c) {
  const int nn = c->mesh.nnode, neq = c->mesh.neq;
  PM_TRY(dalloc(&c->d_u, (size_t)nn * 4));
  PM_TRY(dalloc(&c->d_u0, (size_t)nn * 4));
  PM_TRY(dalloc(&c->d_dp, (size_t)nn));
  PM_TRY(dalloc(&c->d_poro, (size_t)nn));
  PM_TRY(dalloc(&c->d_dporo, (size_t)nn));
  PM_TRY(dalloc(&c->d_b0, (size_t)neq));
  PM_TRY(dalloc(&c->d_b, (size_t)neq));
  PM_TRY(dalloc(&c->d_x0, (size_t)neq));
  PM_TRY(dalloc(&c->d_m, (size_t)neq));
  PM_CHECK_HIP(hipMemsetAsync(c->d_u, 0, sizeof(double) * nn * 4, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_u0, 0, sizeof(double) * nn * 4, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_dp, 0, sizeof(double) * nn, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_poro, 0, sizeof(double) * nn, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_b0, 0, sizeof(double) * neq, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_b, 0, sizeof(double) * neq, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_x0, 0, sizeof(double) * neq, c->stream));
  c->t = 0.0;
  return PM_OK;
}

static int mesh_setup(pm_ctx *c, int nen, int numel, int nnodev, const int *ix, const double *x,
                      const int *idv, int reorder, int *nnode, int *neq, int64_t *nad) {
  if (!c || numel <= 0 || nnodev <= 0 || !ix || !x || !idv) return PM_ERR_ARG;
  if (nen != 8 && nen != 20) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  PmHostMesh &m = c->mesh;
  m = PmHostMesh();
  c->have_mesh = false;
  FREE(c->d_tf_ptr);  // per-mesh maps are rebuilt on first use
  FREE(c->d_tf_src);
  FREE(c->d_tf_out);
  c->have_ordered_maps = false;
  m.numel = numel;
  m.nnodev = nnodev;
  if (nen == 20) PM_TRY(pm_host_promote(m, ix));
  else PM_TRY(pm_host_plain_hexa8(m, ix));
  m.idflags.assign((size_t)m.nnode * 4, 0);
  for (size_t i = 0; i < (size_t)nnodev * 4; i++) m.idflags[i] = idv[i];
  if (nen == 20) pm_host_bound_quad(m);
  pm_host_graph(m);
  pm_host_rcm(m, reorder);
  pm_host_numeq(m);
  pm_host_colour(m);
  PmHostPattern p;
  PM_TRY(pm_host_pattern(m, p));
  m.nad = pm_host_count_lower(p);
  PM_TRY(pm_upload_pattern(c, p));
  // device mesh: 0-based nodes (nen per element), equations (-1 = none),
  // pressure flags, colours
  std::vector<int> ix0((size_t)numel * nen), eq0((size_t)m.nnode * 4, -1), mat0((size_t)numel);
  for (int e = 0; e < numel; e++) {
    for (int k = 0; k < nen; k++) ix0[(size_t)e * nen + k] = m.ix20[(size_t)e * 21 + k] - 1;
    mat0[e] = m.ix20[(size_t)e * 21 + 20] - 1;
  }
  std::vector<unsigned char> isp((size_t)m.neq, 0);
  for (int n = 0; n < m.nnode; n++) {
    const int nd = m.fnno[n] == 1 ? 4 : 3;
    for (int d = 0; d < nd; d++) {
      const int q = m.id[(size_t)n * 4 + d];
      if (q > 0) {
        eq0[(size_t)n * 4 + d] = q - 1;
        if (d == 3) isp[q - 1] = 1;
      }
    }
  }
  std::vector<int> ic((size_t)m.nnode, 0);
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < 8; k++) ic[m.ix20[(size_t)e * 21 + k] - 1]++;
  PM_TRY(dupload(c, &c->d_ix, ix0.data(), ix0.size()));
  PM_TRY(dupload(c, &c->d_xv, x, (size_t)nnodev * 3));
  PM_TRY(dupload(c, &c->d_eq, eq0.data(), eq0.size()));
  PM_TRY(dupload(c, &c->d_isp, isp.data(), isp.size()));
  PM_TRY(dupload(c, &c->d_mat, mat0.data(), mat0.size()));
  PM_TRY(dupload(c, &c->d_colour_elems, m.colour_elems.data(), m.colour_elems.size()));
  PM_TRY(dupload(c, &c->d_ic, ic.data(), ic.size()));
  c->colour_ptr = m.colour_ptr;
  c->mesh.neq = m.neq;
  PM_TRY(ensure_state(c));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->have_mesh = true;
  c->have_materials = false;
  c->have_nodal_loads = false;
  if (nnode) *nnode = m.nnode;
  if (neq) *neq = m.neq;
  if (nad) *nad = m.nad;
  return PM_OK;
}

extern "C" int pm_mesh_setup(pm_ctx *c, int numel, int nnodev, const int *ix, const double *x,
                             const int *idv, int reorder, int *nnode, int *neq, int64_t *nad) {
  return mesh_setup(c, 20, numel, nnodev, ix, x, idv, reorder, nnode, neq, nad);
}

extern "C" int pm_mesh_setup_nen(pm_ctx *c, int nen, int numel, int nnodev, const int *ix,
                                 const double *x, const int *idv, int reorder, int *nnode,
                                 int *neq, int64_t *nad) {
  return mesh_setup(c, nen, numel, nnodev, ix, x, idv, reorder, nnode, neq, nad);
}

static int mesh_numbering(int nen, int numel, int nnodev, const int *ix, const int *idv,
                          int reorder, int nnode_max, int *ix20, int *id, int *perm, int *nnode,
                          int *neq, int64_t *nad) {
  if (numel <= 0 || nnodev <= 0 || !ix || !idv || (nen != 8 && nen != 20)) return PM_ERR_ARG;
  PmHostMesh m;
  m.numel = numel;
  m.nnodev = nnodev;
  if (nen == 20) PM_TRY(pm_host_promote(m, ix));
  else PM_TRY(pm_host_plain_hexa8(m, ix));
  if (m.nnode > nnode_max) return PM_ERR_ARG;
  m.idflags.assign((size_t)m.nnode * 4, 0);
  for (size_t i = 0; i < (size_t)nnodev * 4; i++) m.idflags[i] = idv[i];
  if (nen == 20) pm_host_bound_quad(m);
  pm_host_graph(m);
  pm_host_rcm(m, reorder);
  pm_host_numeq(m);
  PmHostPattern p;
  PM_TRY(pm_host_pattern(m, p));
  if (ix20) memcpy(ix20, m.ix20.data(), sizeof(int) * m.ix20.size());
  if (id) memcpy(id, m.id.data(), sizeof(int) * m.id.size());
  if (perm) memcpy(perm, m.perm.data(), sizeof(int) * m.perm.size());
  if (nnode) *nnode = m.nnode;
  if (neq) *neq = m.neq;
  if (nad) *nad = pm_host_count_lower(p);
  return PM_OK;
}

extern "C" int pm_mesh_numbering(int numel, int nnodev, const int *ix, const int *idv,
                                 int reorder, int nnode_max, int *ix20, int *id, int *perm,
                                 int *nnode, int *neq, int64_t *nad) {
  return mesh_numbering(20, numel, nnodev, ix, idv, reorder, nnode_max, ix20, id, perm, nnode,
                        neq, nad);
}

extern "C" int pm_mesh_numbering_nen(int nen, int numel, int nnodev, const int *ix,
                                     const int *idv, int reorder, int nnode_max, int *ix20,
                                     int *id, int *perm, int *nnode, int *neq, int64_t *nad) {
  return mesh_numbering(nen, numel, nnodev, ix, idv, reorder, nnode_max, ix20, id, perm, nnode,
                        neq, nad);
}

extern "C" int pm_get_numbering(pm_ctx *c, int *ix20, int *id, int *perm) {
  if (!c || !c->have_mesh) return PM_ERR_STATE;
  const PmHostMesh &m = c->mesh;
  if (ix20) memcpy(ix20, m.ix20.data(), sizeof(int) * m.ix20.size());
  if (id) memcpy(id, m.id.data(), sizeof(int) * m.id.size());
  if (perm) memcpy(perm, m.perm.data(), sizeof(int) * m.perm.size());
  return PM_OK;
}

extern "C" int pm_set_materials(pm_ctx *c, int numat, const double *e, const int *eload,
                                const int *eloadp, const int *load, const double *fload,
                                const double *gravity, double gravity_mod) {
  if (!c || !c->have_mesh) return PM_ERR_STATE;
  if (numat <= 0 || !e) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const PmHostMesh &m = c->mesh;
  for (int el = 0; el < m.numel; el++) {
    const int ma = m.ix20[(size_t)el * 21 + 20];
    if (ma < 1 || ma > numat) return PM_ERR_ARG;
  }
  std::vector<int> el14((size_t)m.numel * 14, 0);
  int any = 0;
  for (int el = 0; el < m.numel; el++)
    for (int j = 0; j < 7; j++) {
      const int a = eload ? eload[(size_t)el * 7 + j] : 0;
      const int b = eloadp ? eloadp[(size_t)el * 7 + j] : 0;
      if (a < 0 || a > PM_NUMLOAD || b < 0 || b > PM_NUMLOAD) return PM_ERR_ARG;
      el14[(size_t)el * 14 + j] = a;
      el14[(size_t)el * 14 + 7 + j] = b;
      any |= a | b;
    }
  std::vector<int> ld((size_t)PM_NUMLOAD * 3, 0);
  std::vector<double> fl((size_t)PM_NUMLOAD * PM_NTERMS * PM_NPAR, 0.0);
  if (load) memcpy(ld.data(), load, sizeof(int) * ld.size());
  if (fload) memcpy(fl.data(), fload, sizeof(double) * fl.size());
  // every face load the elements use must be a table load (types 40, 41,
  // 43) with 1..1000 time points; type 42 is outside this path
  for (int el = 0; el < m.numel; el++)
    for (int j = 0; j < 14; j++) {
      const int nc = el14[(size_t)el * 14 + j];
      if (nc == 0) continue;
      const int type = ld[(size_t)3 * (nc - 1)], npt = ld[(size_t)3 * (nc - 1) + 2];
      if (type == 42) return PM_ERR_UNSUPPORTED;
      if (j < 7 && type != 40 && type != 41) return PM_ERR_ARG;
      if (j >= 7 && type != 43) return PM_ERR_ARG;
      if (npt < 1 || npt > PM_NPAR) return PM_ERR_ARG;
    }
  c->numat = numat;
  c->h_load = ld;
  c->h_e.assign(e, e + (size_t)numat * PM_PROP);
  c->vprop[0] = c->vprop[1] = c->vprop[2] = c->vprop[3] = 0;  // pm_set_vprop again
  PM_TRY(dupload(c, &c->d_e, e, (size_t)numat * PM_PROP));
  PM_TRY(dupload(c, &c->d_eload, el14.data(), el14.size()));
  PM_TRY(dupload(c, &c->d_load, ld.data(), ld.size()));
  PM_TRY(dupload(c, &c->d_fload, fl.data(), fl.size()));
  c->has_face_loads = any != 0;
  for (int d = 0; d < 3; d++) c->gravity[d] = gravity ? gravity[d] : 0.0;
  c->gravity_mod = gravity_mod;
  // initial_porosity (Pform.f:1331-1376, run by the mesh macro, Main.f:820-828):
  // every vertex starts at the average of e(8) over its elements, summed in
  // element order
  std::vector<double> poro((size_t)m.nnode, 0.0);
  std::vector<int> ic((size_t)m.nnode, 0);
  for (int el = 0; el < m.numel; el++) {
    const int ma = m.ix20[(size_t)el * 21 + 20];
    const double p = e[(size_t)(ma - 1) * PM_PROP + 7];
    for (int i = 0; i < 8; i++) {
      const int no = m.ix20[(size_t)el * 21 + i];
      if (no <= 0) continue;
      ic[no - 1]++;
      poro[no - 1] = poro[no - 1] + p;
    }
  }
  for (int n = 0; n < m.nnode; n++)
    if (m.fnno[n] == 1 && ic[n] > 0) poro[n] = poro[n] / ic[n];
  PM_CHECK_HIP(hipMemcpyAsync(c->d_poro, poro.data(), sizeof(double) * poro.size(),
                              hipMemcpyHostToDevice, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->have_materials = true;
  return PM_OK;
}

extern "C" int pm_set_nodal_loads(pm_ctx *c, const int *nload, const double *f) {
  if (!c || !c->have_mesh) return PM_ERR_STATE;
  const size_t n = (size_t)c->mesh.nnode * 4;
  if (nload) PM_TRY(dupload(c, &c->d_nload, nload, n));
  else {
    std::vector<int> z(n, 0);
    PM_TRY(dupload(c, &c->d_nload, z.data(), n));
  }
  if (nload) {  // time functions must index the load table within its bounds
    if (!c->have_materials) return PM_ERR_STATE;
    for (size_t i = 0; i < n; i++) {
      const int nc = nload[i];
      if (nc == 0) continue;
      if (nc < 0 || nc > PM_NUMLOAD) return PM_ERR_ARG;
      const int type = c->h_load[(size_t)3 * (nc - 1)];
      const int np2 = c->h_load[(size_t)3 * (nc - 1) + 1];
      const int np3 = c->h_load[(size_t)3 * (nc - 1) + 2];
      if (type >= 4 && type <= 6 && (np2 < 0 || np2 > PM_NTERMS)) return PM_ERR_ARG;
      if ((type == 7 || type == 8) && (np2 < 1 || np2 > PM_NPAR)) return PM_ERR_ARG;
      if (type >= 39 && type <= 43 && (np3 < 1 || np3 > PM_NPAR)) return PM_ERR_ARG;
    }
  }
  if (f) PM_TRY(dupload(c, &c->d_f, f, n));
  else {
    std::vector<double> z(n, 0.0);
    PM_TRY(dupload(c, &c->d_f, z.data(), n));
  }
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->have_nodal_loads = true;
  return PM_OK;
}

extern "C" int pm_pform(pm_ctx *c, int isw, int lhs, int rhs, double dt, double t, int target) {
  if (!c) return PM_ERR_ARG;
  if (!c->have_mesh) return PM_ERR_STATE;
  if (isw != 2 && isw != 4) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  int rc = pm_pform_dev(c, isw, lhs, rhs, dt, t, target ? c->d_b : c->d_b0);
  if (rc == PM_OK) PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int pm_solve(pm_ctx *c, const pm_solver_cfg *cfg, const double *b, double *x,
                        pm_solve_info *info) {
  if (!c || !cfg) return PM_ERR_ARG;
  if (!c->have_matrix || !c->A.values_valid) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const int n = c->A.neq;
  if (!c->d_b || !c->d_x0) {
    PM_TRY(dalloc(&c->d_b, (size_t)n));
    PM_TRY(dalloc(&c->d_x0, (size_t)n));
    PM_TRY(dalloc(&c->d_m, (size_t)n));
  }
  if (b)
    PM_CHECK_HIP(hipMemcpyAsync(c->d_b, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  pm_solve_info local;
  pm_solve_info *inf = info ? info : &local;
  int rc = pm_solve_dev(c, cfg, c->d_b, c->d_x0, inf);
  if (rc == PM_OK && x) {
    PM_CHECK_HIP(hipMemcpyAsync(x, c->d_x0, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  }
  return rc;
}

extern "C" int pm_struct_upload(pm_ctx *c, int neq, int64_t nad, const int64_t *ia,
                                const int *ja) {
  if (!c || !ia || (nad > 0 && !ja)) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  PmHostPattern p;
  PM_TRY(pm_host_pattern_from_csrc(neq, nad, ia, ja, p));
  PM_TRY(pm_upload_pattern(c, p));
  c->mesh.neq = neq;
  c->mesh.nad = nad;
  c->have_mesh = false;
  c->A.values_valid = false;
  FREE(c->d_b);
  FREE(c->d_x0);
  FREE(c->d_m);
  PM_TRY(dalloc(&c->d_b, (size_t)neq));
  PM_TRY(dalloc(&c->d_x0, (size_t)neq));
  PM_TRY(dalloc(&c->d_m, (size_t)neq));
  // keep the lower pattern for value uploads
  c->csrc_ia.assign(ia, ia + neq + 1);
  c->csrc_ja.assign(ja, ja + nad);
  return PM_OK;
}

extern "C" int pm_values_upload(pm_ctx *c, const double *ad, const double *al) {
  if (!c || !ad) return PM_ERR_ARG;
  if (!c->have_matrix || c->csrc_ia.empty()) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const int neq = c->A.neq;
  const int64_t nad = c->csrc_ia[neq] - 1;
  int64_t *d_ia = nullptr;
  int *d_ja = nullptr;
  double *d_ad = nullptr, *d_al = nullptr;
  PM_TRY(dupload(c, &d_ia, c->csrc_ia.data(), (size_t)neq + 1));
  PM_TRY(dupload(c, &d_ja, c->csrc_ja.data(), (size_t)nad));
  PM_TRY(dupload(c, &d_ad, ad, (size_t)neq));
  if (nad > 0) PM_TRY(dupload(c, &d_al, al, (size_t)nad));
  else PM_TRY(dalloc(&d_al, 1));
  PmDevMatrix &A = c->A;
  PM_CHECK_HIP(hipMemsetAsync(A.val, 0, sizeof(double) * (size_t)A.nnz, c->stream));
  hipLaunchKernelGGL(k_csrc_values, dim3((neq + 255) / 256), dim3(256), 0, c->stream, neq, d_ia,
                     d_ja, d_ad, d_al, A.eq_block, A.fr, A.colptr, A.valptr, A.ncol, A.val);
  PM_CHECK_HIP(hipGetLastError());
  PM_TRY(pm_extract_diag(c));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(d_ia);
  (void)hipFree(d_ja);
  (void)hipFree(d_ad);
  (void)hipFree(d_al);
  A.values_valid = true;
  A.have32 = false;
  return PM_OK;
}

extern "C" int pm_spmv(pm_ctx *c, const double *x, double *y) {
  if (!c || !x || !y) return PM_ERR_ARG;
  if (!c->have_matrix || !c->A.values_valid) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const int n = c->A.neq;
  double *dx = nullptr, *dy = nullptr;
  PM_CHECK_HIP(hipMalloc(&dx, sizeof(double) * n));
  PM_CHECK_HIP(hipMalloc(&dy, sizeof(double) * n));
  PM_CHECK_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  int rc = pm_spmv_dev(c, dx, dy, 0);
  if (rc == PM_OK) {
    PM_CHECK_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  }
  hipFree(dx);
  hipFree(dy);
  return rc;
}

extern "C" int pm_export_csrc(pm_ctx *c, int64_t *ia, int *ja, double *ad, double *al) {
  if (!c) return PM_ERR_ARG;
  if (!c->have_matrix) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  PmDevMatrix &A = c->A;
  const int neq = A.neq;
  int64_t *d_cnt = nullptr;
  PM_TRY(dalloc(&d_cnt, (size_t)neq + 1));
  hipLaunchKernelGGL(k_lower_count, dim3((A.nrn + 255) / 256), dim3(256), 0, c->stream, A.nrn,
                     A.fr, A.nr, A.colptr, A.ncol, d_cnt);
  std::vector<int64_t> cnt((size_t)neq);
  PM_CHECK_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int64_t) * neq, hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  std::vector<int64_t> hia((size_t)neq + 1);
  hia[0] = 1;
  for (int i = 0; i < neq; i++) hia[i + 1] = hia[i] + cnt[i];
  const int64_t nad = hia[neq] - 1;
  if (ia) memcpy(ia, hia.data(), sizeof(int64_t) * (neq + 1));
  if (ja || al) {
    int64_t *d_ia = nullptr;
    int *d_ja = nullptr;
    double *d_al = nullptr;
    PM_TRY(dupload(c, &d_ia, hia.data(), hia.size()));
    if (ja) PM_TRY(dalloc(&d_ja, (size_t)nad));
    if (al) PM_TRY(dalloc(&d_al, (size_t)nad));
    hipLaunchKernelGGL(k_lower_fill, dim3((A.nrn + 255) / 256), dim3(256), 0, c->stream, A.nrn,
                       A.fr, A.nr, A.colptr, A.valptr, A.ncol, A.val, d_ia, d_ja, d_al);
    if (ja)
      PM_CHECK_HIP(hipMemcpyAsync(ja, d_ja, sizeof(int) * nad, hipMemcpyDeviceToHost, c->stream));
    if (al)
      PM_CHECK_HIP(hipMemcpyAsync(al, d_al, sizeof(double) * nad, hipMemcpyDeviceToHost, c->stream));
    PM_CHECK_HIP(hipStreamSynchronize(c->stream));
    FREE(d_ia);
    FREE(d_ja);
    FREE(d_al);
  }
  if (ad)
    PM_CHECK_HIP(hipMemcpyAsync(ad, A.diag, sizeof(double) * neq, hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  FREE(d_cnt);
  return PM_OK;
}

extern "C" int pm_set_state(pm_ctx *c, const double *u, const double *u0, const double *dp,
                            const double *porosity, double t) {
  if (!c || !c->have_mesh) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const size_t nn = (size_t)c->mesh.nnode;
  if (u) PM_CHECK_HIP(hipMemcpyAsync(c->d_u, u, sizeof(double) * nn * 4, hipMemcpyHostToDevice, c->stream));
  if (u0) PM_CHECK_HIP(hipMemcpyAsync(c->d_u0, u0, sizeof(double) * nn * 4, hipMemcpyHostToDevice, c->stream));
  if (dp) PM_CHECK_HIP(hipMemcpyAsync(c->d_dp, dp, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  if (porosity)
    PM_CHECK_HIP(hipMemcpyAsync(c->d_poro, porosity, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->t = t;
  return PM_OK;
}

extern "C" int pm_get_state(pm_ctx *c, double *u, double *u0, double *dp, double *porosity,
                            double *t) {
  if (!c || !c->have_mesh) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const size_t nn = (size_t)c->mesh.nnode;
  if (u) PM_CHECK_HIP(hipMemcpyAsync(u, c->d_u, sizeof(double) * nn * 4, hipMemcpyDeviceToHost, c->stream));
  if (u0) PM_CHECK_HIP(hipMemcpyAsync(u0, c->d_u0, sizeof(double) * nn * 4, hipMemcpyDeviceToHost, c->stream));
  if (dp) PM_CHECK_HIP(hipMemcpyAsync(dp, c->d_dp, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  if (porosity)
    PM_CHECK_HIP(hipMemcpyAsync(porosity, c->d_poro, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (t) *t = c->t;
  return PM_OK;
}

extern "C" int pm_get_vectors(pm_ctx *c, double *b0, double *b, double *x0) {
  if (!c || !c->have_matrix) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)c->A.neq;
  if (b0 && c->d_b0) PM_CHECK_HIP(hipMemcpyAsync(b0, c->d_b0, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (b && c->d_b) PM_CHECK_HIP(hipMemcpyAsync(b, c->d_b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (x0 && c->d_x0) PM_CHECK_HIP(hipMemcpyAsync(x0, c->d_x0, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PM_OK;
}

extern "C" int pm_set_vectors(pm_ctx *c, const double *b0, const double *b, const double *x0) {
  if (!c || !c->have_matrix) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)c->A.neq;
  if (b0 && c->d_b0) PM_CHECK_HIP(hipMemcpyAsync(c->d_b0, b0, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (b && c->d_b) PM_CHECK_HIP(hipMemcpyAsync(c->d_b, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (x0 && c->d_x0) PM_CHECK_HIP(hipMemcpyAsync(c->d_x0, x0, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PM_OK;
}

extern "C" int pm_step(pm_ctx *c, double dt, const pm_solver_cfg *cfg, pm_step_info *info) {
  if (!c || !cfg) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  pm_step_info local;
  return pm_step_dev(c, dt, cfg, info ? info : &local);
}

// ks_and_mus (Poromec.f:48-111): the solid moduli whose Hashin-Shtrikman
// bound at poro is the drained (k, mu); false where the reference stops
static bool ks_and_mus(double poro, double &k, double &mu) {
  const double ha = 1.0 - poro;
  const double hb = 9.0 + 6.0 * poro;
  const double hc = 8.0 + 12.0 * poro;
  const double a = 32.0 * ha * ha;
  const double b = 4.0 * ha * (9.0 * k - mu * hc - 6.0 * poro * k);
  const double cc = 3.0 * mu * hc * poro * k - 4.0 * k * mu * hb;
  double delta = b * b - 4.0 * a * cc;
  if (!(delta > 0)) return false;
  delta = std::sqrt(delta);
  const double mus = 0.5 * (delta - b) / a;
  if (mus < 0) return false;
  const double ks = 4.0 * k * mus / (4.0 * mus * ha - 3.0 * poro * k);
  if (ks < 0) return false;
  k = ks;
  mu = mus;
  return true;
}

static void hashin_shtrikman_h(double poro, double ks, double mus, double *v) {
  const double ha = 1.0 - poro;
  const double hb = 9.0 + 6.0 * poro;
  const double hc = 8.0 + 12.0 * poro;
  v[0] = (4.0 * ks * mus * ha) / (3.0 * ks * poro + 4.0 * mus);
  v[1] = (mus * ha * (9.0 * ks + 8 * mus)) / (ks * hb + mus * hc);
  v[3] = 1.0 - v[0] / ks;
  v[2] = (v[3] - poro) / ks;
}

extern "C" int pm_set_vprop(pm_ctx *c, const int *ie, const int *vprop) {
  if (!c || !ie) return PM_ERR_ARG;
  if (!c->have_materials) return PM_ERR_STATE;
  if (c->batch) return PM_ERR_UNSUPPORTED;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const int numat = c->numat;
  int f[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4 && vprop; i++) f[i] = vprop[i] != 0;
  if (f[1] || f[2] || f[3]) f[0] = 1;  // read_constitutive_equation, Rdata.f:2630-2633
  if (c->mesh.nen == 8) {  // a maxno-8 mesh: every material is the Q1/Q1 element
    for (int ma = 0; ma < numat; ma++)
      if (ie[ma] != 7) return PM_ERR_UNSUPPORTED;
    return f[0] ? PM_ERR_UNSUPPORTED : PM_OK;
  }
  bool any18 = false;
  for (int ma = 0; ma < numat; ma++) {
    if (ie[ma] != 17 && ie[ma] != 18) return PM_ERR_UNSUPPORTED;
    any18 |= ie[ma] == 18;
  }
  if (f[0] != (int)any18) return PM_ERR_ARG;  // check_element, Rdata.f:3744-3758
  for (int i = 0; i < 4; i++) c->vprop[i] = 0;
  if (!f[0]) {
    PM_TRY(dupload(c, &c->d_e, c->h_e.data(), c->h_e.size()));
    return PM_OK;
  }
  // initial_prop (Pform.f:1434-1540), from the materials as given
  std::vector<double> e = c->h_e;
  for (int ma = 0; ma < numat; ma++) {
    double *em = e.data() + (size_t)ma * PM_PROP;
    const double ym = em[0], ps = em[1], ro = em[5], rof = em[6], poro = em[7];
    const double lam = ym * ps / ((1.0 + ps) * (1.0 - 2.0 * ps));
    double mu = (0.5 * ym) / (1.0 + ps);
    double k = lam + (2.0 / 3.0) * mu;
    if (f[2]) em[5] = (ro - poro * rof) / (1.0 - poro);
    if (f[3] && !ks_and_mus(poro, k, mu)) return PM_ERR_ARG;
    em[0] = k;
    em[1] = mu;
  }
  const PmHostMesh &m = c->mesh;
  std::vector<double> vp((size_t)m.numel * 64 * 7);
  for (int el = 0; el < m.numel; el++) {
    const int ma = m.ix20[(size_t)el * 21 + 20];
    const double *em = e.data() + (size_t)(ma - 1) * PM_PROP;
    const double k = em[0], mu = em[1], perm = em[2], ro = em[5], rof = em[6], poro = em[7];
    const double ibiot = 1 / em[3], cbiot = em[4];
    for (int i = 0; i < 64; i++) {
      double *v = vp.data() + ((size_t)el * 64 + i) * 7;
      v[0] = poro;
      v[1] = perm;
      v[2] = f[2] ? (1.0 - poro) * ro + poro * rof : ro;
      v[2] = v[2] * 1.0e-06;
      if (f[3]) {
        hashin_shtrikman_h(poro, k, mu, v + 3);
      } else {
        v[3] = k;
        v[4] = mu;
        v[5] = ibiot;
        v[6] = cbiot;
      }
    }
  }
  PM_TRY(dupload(c, &c->d_e, e.data(), e.size()));
  PM_TRY(dupload(c, &c->d_vp, vp.data(), vp.size()));
  PM_TRY(dupload(c, &c->d_ie, ie, (size_t)numat));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 4; i++) c->vprop[i] = f[i];
  return PM_OK;
}

extern "C" int pm_get_vprop(pm_ctx *c, double *vpropel) {
  if (!c || !vpropel) return PM_ERR_ARG;
  if (!c->vprop[0]) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  PM_CHECK_HIP(hipMemcpyAsync(vpropel, c->d_vp, sizeof(double) * (size_t)c->mesh.numel * 64 * 7,
                              hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PM_OK;
}

extern "C" int pm_tform(pm_ctx *c, double *tx, double *txe, double *txb, double *flux) {
  if (!c) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  return pm_tform_dev(c, tx, txe, txb, flux);
}

extern "C" int pm_set_ordered(pm_ctx *c, int on) {
  if (!c) return PM_ERR_ARG;
  c->ordered = on != 0;
  return PM_OK;
}

extern "C" int pm_adjoint(pm_ctx *c, int nsteps, double dt, const pm_solver_cfg *cfg, int nsens,
                          const int *sensor_dof, const double *target, int ckpt_every, double *J,
                          double *grad_logE, double *grad_logk) {
  if (!c || !cfg) return PM_ERR_ARG;
  if (c->vprop[0]) return PM_ERR_UNSUPPORTED;  // the adjoint is of elmt17_pm's march
  if (c->mesh.nen == 8)  // Q1/Q1: its pressure stabilization would make C depend on E
    for (int ma = 0; ma < c->numat; ma++)
      if (c->h_e[(size_t)ma * PM_PROP + 8] != 0.0) return PM_ERR_UNSUPPORTED;
  PM_CHECK_HIP(hipSetDevice(c->device));
  return pm_adjoint_dev(c, nsteps, dt, cfg, nsens, sensor_dof, target, ckpt_every, J, grad_logE,
                        grad_logk);
}

extern "C" int pm_spmv_timing(pm_ctx *c, int on, double *stats) {
  if (!c) return PM_ERR_ARG;
  if (stats) {
    stats[0] = c->spmv_ms[0];
    stats[1] = (double)c->spmv_n[0];
    stats[2] = c->spmv_ms[1];
    stats[3] = (double)c->spmv_n[1];
  }
  c->spmv_ms[0] = c->spmv_ms[1] = 0.0;
  c->spmv_n[0] = c->spmv_n[1] = 0;
  c->time_spmv = on > 0 ? on : 0;
  c->spmv_seq = 0;
  return PM_OK;
}

extern "C" int pm_matrix_info(pm_ctx *c, int *neq, int *nblocks, int64_t *nnz, int64_t *ncols) {
  if (!c || !c->have_matrix) return PM_ERR_STATE;
  if (neq) *neq = c->A.neq;
  if (nblocks) *nblocks = c->A.nrn;
  if (nnz) *nnz = c->A.nnz;
  if (ncols) *ncols = c->A.ncols;
  return PM_OK;
}

extern "C" int pm_matrix_runs(pm_ctx *c, int64_t *nruns) {
  if (!c || !nruns) return PM_ERR_ARG;
  if (!c->have_matrix) return PM_ERR_STATE;
  *nruns = c->A.nchunks;
  return PM_OK;
}

extern "C" int pm_bench_spmv(pm_ctx *c, int use32, int reps, double *ms_per_launch) {
  if (!c || reps <= 0 || !ms_per_launch) return PM_ERR_ARG;
  if (!c->have_matrix || !c->A.values_valid) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  if (use32 && !c->A.have32) PM_TRY(pm_make_val32(c));
  PM_TRY(pm_ensure_workspace(c, c->A.neq));
  double *x = c->d_wk[0], *y = c->d_wk[1];
  PM_CHECK_HIP(hipMemsetAsync(x, 0, sizeof(double) * c->A.neq, c->stream));
  PM_TRY(pm_spmv_dev(c, x, y, use32));  // warm-up
  PM_CHECK_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++) PM_TRY(pm_spmv_dev(c, (r & 1) ? y : x, (r & 1) ? x : y, use32));
  PM_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
  PM_CHECK_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  PM_CHECK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_per_launch = ms / reps;
  return PM_OK;
}

// The SpMV shape tuner: times every (LANES, UNR) shape on the current matrix,
// in the form the Krylov loops launch it (fused dot and finisher), and keeps
// the fastest for that value precision.  Shapes sum in different
// orders, so results move at the rounding level when the shape changes; the
// choice stays fixed until the next call.
extern "C" int pm_tune_spmv(pm_ctx *c, int use32, int reps, int *best, double *ms_per_shape) {
  if (!c || reps <= 0) return PM_ERR_ARG;
  if (!c->have_matrix || !c->A.values_valid) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  if (use32 && !c->A.have32) PM_TRY(pm_make_val32(c));
  PM_TRY(pm_ensure_workspace(c, c->A.neq));
  double *x = c->d_wk[0], *y = c->d_wk[1];
  PM_CHECK_HIP(hipMemsetAsync(x, 0, sizeof(double) * c->A.neq, c->stream));
  PmScalars z;
  memset(&z, 0, sizeof(z));  // the fused form skips its work once done is set
  PM_CHECK_HIP(hipMemcpyAsync(c->d_scal, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
  const int p = use32 ? 1 : 0, keep = c->spmv_cfg[p], ordered = c->ordered;
  c->ordered = 0;
  int arg = -1;
  double tbest = 0.0;
  for (int k = 0; k < PM_SPMV_NCFG; k++) {
    c->spmv_cfg[p] = k;
    auto spmv = [&](const double *a, double *b) {
      return c->batch ? pm_batch_spmv_dev(c, a, b, use32) : pm_spmv_fused_dev(c, a, b, use32);
    };
    int rc = spmv(x, y);  // warm-up
    if (rc == PM_OK) rc = hipEventRecord(c->ev0, c->stream) == hipSuccess ? PM_OK : PM_ERR_HIP;
    for (int r = 0; r < reps && rc == PM_OK; r++) rc = spmv((r & 1) ? y : x, (r & 1) ? x : y);
    float ms = 0.f;
    if (rc == PM_OK && (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
                        hipEventSynchronize(c->ev1) != hipSuccess ||
                        hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess))
      rc = PM_ERR_HIP;
    if (rc != PM_OK) {
      c->spmv_cfg[p] = keep;
      c->ordered = ordered;
      return rc;
    }
    const double per = ms / reps;
    if (ms_per_shape) ms_per_shape[k] = per;
    if (arg < 0 || per < tbest) {
      arg = k;
      tbest = per;
    }
  }
  c->ordered = ordered;
  c->spmv_cfg[p] = arg;
  c->spmv_tuned[p] = 1;
  if (best) *best = arg;
  return PM_OK;
}

extern "C" int pm_set_spmv_shape(pm_ctx *c, int use32, int shape) {
  if (!c || shape < 0 || shape >= PM_SPMV_NCFG) return PM_ERR_ARG;
  c->spmv_cfg[use32 ? 1 : 0] = shape;
  return PM_OK;
}

__global__ __launch_bounds__(256) void k_copy_bench(int64_t n, const double4 *__restrict__ a,
                                                   double4 *__restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    b[i] = a[i];
}

extern "C" int pm_bench_copy(pm_ctx *c, int64_t bytes, int reps, double *ms_per_launch) {
  if (!c || bytes < 32 || reps <= 0 || !ms_per_launch) return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  const int64_t n = bytes / 32;
  double4 *a = nullptr, *b = nullptr;
  PM_CHECK_HIP(hipMalloc(&a, n * 32));
  PM_CHECK_HIP(hipMalloc(&b, n * 32));
  PM_CHECK_HIP(hipMemsetAsync(a, 0, n * 32, c->stream));
  const int g = 8 * c->cu_count;
  hipLaunchKernelGGL(k_copy_bench, dim3(g), dim3(256), 0, c->stream, n, a, b);
  PM_CHECK_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy_bench, dim3(g), dim3(256), 0, c->stream, n, a, b);
  PM_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
  PM_CHECK_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  PM_CHECK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_per_launch = ms / reps;
  (void)hipFree(a);
  (void)hipFree(b);
  return PM_OK;
}

extern "C" int pm_batch_define(pm_ctx *c, int nbox, const int *elem_off) {
  if (!c || !elem_off) return PM_ERR_ARG;
  if (c->vprop[0]) return PM_ERR_UNSUPPORTED;
  PM_CHECK_HIP(hipSetDevice(c->device));
  return pm_batch_define_dev(c, nbox, elem_off);
}

extern "C" int pm_batch_iters(pm_ctx *c, int *iters) {
  if (!c || !iters) return PM_ERR_ARG;
  return pm_batch_last_iters(c, iters);
}

extern "C" int pm_batch_graphs(pm_ctx *c, int on) {
  if (!c || !c->batch) return PM_ERR_STATE;
  pm_batch_set_graph(c, on);
  return PM_OK;
}
