// pm_noalloc.hip -- pm_solve_noalloc: pm_solve with its scratch memory in a
// static device array, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).  No counterpart in
// the reference.
//
// The scratch holds what pm_solve's kernels use beyond the matrix, b and x:
// the 16 Krylov work vectors, the Jacobi vector, the workgroup partial sums
// of the reductions, and device copies of b and x.  Each GPU holds its own
// copy of the array from the first launch of this library's code on it
// until the process ends; its address on each device is taken once, with
// that device current, by hipGetSymbolAddress.  Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "pm_internal.h"

int pm_solve_dev(pm_ctx *c, const pm_solver_cfg *cfg, const double *b, double *x,
                 pm_solve_info *info);

namespace {

constexpr int64_t kMaxNeq = PM_NOALLOC_MAX_NEQ;
constexpr int kVectors = 16 + 3;                  // work vectors, m, b, x
constexpr int64_t kPartials = 8192 * 4;           // as the context's d_partials
constexpr int64_t kWords = kVectors * kMaxNeq + kPartials;

__device__ __align__(256) double scratch[kWords];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_ptr[kMaxDevices] = {nullptr};

// the array's address on `device` (current), looked up once per device
int scratch_ptr(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return PM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      pm_set_last_error(hipGetErrorString(e), __FILE__, __LINE__);
      return PM_ERR_HIP;
    }
    g_ptr[device] = static_cast<double *>(p);
  }
  *out = g_ptr[device];
  return PM_OK;
}

}  // namespace

extern "C" int pm_solve_noalloc(pm_ctx *c, const pm_solver_cfg *cfg, const double *b, double *x,
                                pm_solve_info *info) {
  if (!c || !cfg) return PM_ERR_ARG;
  if (!c->have_matrix || !c->A.values_valid) return PM_ERR_STATE;
  const int n = c->A.neq;
  if (n <= 0 || (int64_t)n > kMaxNeq) return PM_ERR_ARG;  // beyond the scratch's vectors
  if (c->batch) return PM_ERR_UNSUPPORTED;                 // the batching form has its own buffers
  if (cfg->solver == PM_SOLVER_GMRES) return PM_ERR_UNSUPPORTED;  // its basis is (ngram + 1) x neq
  if (cfg->precision == PM_PREC_FP32IR && !c->A.have32) return PM_ERR_STATE;  // FP32 values not made
  if (!b && !c->d_b) return PM_ERR_STATE;
  if (!x && !c->d_x0) return PM_ERR_STATE;
  PM_CHECK_HIP(hipSetDevice(c->device));
  double *base = nullptr;
  PM_TRY(scratch_ptr(c->device, &base));
  double *vec[kVectors];
  for (int i = 0; i < kVectors; i++) vec[i] = base + (int64_t)i * kMaxNeq;
  double *partials = base + (int64_t)kVectors * kMaxNeq;
  double *db = b ? vec[17] : c->d_b;
  double *dx = x ? vec[18] : c->d_x0;
  if (b)
    PM_CHECK_HIP(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));

  // the context's workspace pointers point into the scratch for this call
  double *keep_wk[16];
  std::memcpy(keep_wk, c->d_wk, sizeof(keep_wk));
  const int keep_wk_n = c->wk_n;
  double *keep_m = c->d_m, *keep_partials = c->d_partials;
  const int keep_time = c->time_spmv;
  for (int i = 0; i < 16; i++) c->d_wk[i] = vec[i];
  c->wk_n = (int)kMaxNeq;
  c->d_m = vec[16];
  c->d_partials = partials;
  c->time_spmv = 0;  // live SpMV timing creates events: off for this call
  c->no_alloc = true;
  pm_solve_info local;
  pm_solve_info *inf = info ? info : &local;
  int rc = pm_solve_dev(c, cfg, db, dx, inf);
  if (rc == PM_OK && x) {
    if (hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = PM_ERR_HIP;
  }
  (void)hipStreamSynchronize(c->stream);
  std::memcpy(c->d_wk, keep_wk, sizeof(keep_wk));
  c->wk_n = keep_wk_n;
  c->d_m = keep_m;
  c->d_partials = keep_partials;
  c->time_spmv = keep_time;
  c->no_alloc = false;
  return rc;
}
