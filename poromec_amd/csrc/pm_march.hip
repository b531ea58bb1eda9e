// pm_march.hip -- one backward-Euler step of the mef `solv` macro
// (Main.f:876-1123) on the device: pload_pm (Pload.f:1-92), the Delta
// predictor and update (Predict.f:117-215), cal_residuo_pm (Residuo.f:39-136),
// porosity_form (Pform.f:969-1176) and update_res_v2 (Pload.f:483-541).
#include <cmath>
#include <cstring>
#include <vector>

#include "pm_device.h"

__device__ double dev_tload(int nc, double t, const double *xn, double u, const int *load,
                            const double *fload, double gmod) {
  if (nc <= 0) return 0.0;
  const int itype = load[3 * (nc - 1)], nparc = load[3 * (nc - 1) + 1];
  const double *F = fload + (size_t)(nc - 1) * PM_NTERMS * PM_NPAR;
  auto interp = [&](int tx, int ty, int n) {
    const double *xs = F + (size_t)tx * PM_NPAR, *ys = F + (size_t)ty * PM_NPAR;
    if (t <= xs[0]) return ys[0];
    if (t >= xs[n - 1]) return ys[n - 1];
    for (int i = 0; i < n - 1; i++)
      if (t >= xs[i] && t <= xs[i + 1])
        return ys[i] + (ys[i + 1] - ys[i]) * (t - xs[i]) / (xs[i + 1] - xs[i]);
    return 0.0;
  };
  const double DOISPI = 6.283185307;
  double c = 0.0;
  switch (itype) {
    case 1: return F[0];
    case 2: return F[0] + F[1] * t;
    case 3: return (F[0] - u) * F[1];
    case 4:
    case 5:
      for (int i = 0; i < nparc; i++) {
        const double a = F[(size_t)i * PM_NPAR], w = 1.0 / F[(size_t)i * PM_NPAR + 1],
                     t0 = F[(size_t)i * PM_NPAR + 2];
        const double ti = fmax(t - t0, 0.0);
        c = c + a * (itype == 4 ? sin(DOISPI * w * ti) : cos(DOISPI * w * ti));
      }
      return c;
    case 6:
      for (int i = 0; i < nparc; i++) {
        const double a = F[(size_t)i * PM_NPAR], w = 1.0 / F[(size_t)i * PM_NPAR + 1],
                     t0 = F[(size_t)i * PM_NPAR + 2];
        c = c + a * cos(DOISPI * w * t + t0);
      }
      return c;
    case 7: return interp(1, 2, nparc);
    case 8: return interp(0, 1, nparc);
    case 39:
    case 41:
    case 43: return interp(0, 1, load[3 * (nc - 1) + 2]);
    case 42: {
      const int n = load[3 * (nc - 1) + 2];
      const double a = interp(0, 1, n), b = interp(0, 2, n), w = interp(0, 3, n),
                   t0 = interp(0, 4, n);
      return t0 * (b + a * (w - xn[2]) * gmod);
    }
    default: return 0.0;
  }
}

struct MarchArgs {
  int nnode, nnodev;
  const int *eq;       // (nnode, 4)
  const int *nload;    // (nnode, 4) or null
  double *f;           // (nnode, 4) or null
  const double *xv;    // (nnodev, 3)
  const int *load;
  const double *fload;
  double *u, *u0, *dp, *b0;
  const double *x0;
  double t, dt, gmod;
};

__global__ void k_pload(MarchArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nnode) return;
  const bool vert = i < A.nnodev;
  for (int j = 0; j < 4; j++) {
    if (j == 3 && !vert) break;
    double fv = 0.0;
    if (A.f) {
      if (A.nload) {
        const int nc = A.nload[(size_t)i * 4 + j];
        if (nc > 0)
          A.f[(size_t)i * 4 + j] =
              dev_tload(nc, A.t, A.xv + (size_t)(vert ? i : 0) * 3, A.u[(size_t)i * 4 + j],
                        A.load, A.fload, A.gmod);
      }
      fv = A.f[(size_t)i * 4 + j];
    }
    const int k = A.eq[(size_t)i * 4 + j];
    if (k >= 0) A.b0[k] = (j == 3) ? -A.dt * fv : fv;
    else A.u[(size_t)i * 4 + j] = fv;
  }
}

__global__ void k_predict(MarchArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nnode) return;
  const int nd = i < A.nnodev ? 4 : 3;
  for (int j = 0; j < nd; j++) A.u[(size_t)i * 4 + j] = A.u[(size_t)i * 4 + j] - A.u0[(size_t)i * 4 + j];
}

__global__ void k_delta_update(MarchArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nnode) return;
  const int nd = i < A.nnodev ? 4 : 3;
  for (int j = 0; j < nd; j++) {
    const int k = A.eq[(size_t)i * 4 + j];
    if (k >= 0) A.u[(size_t)i * 4 + j] = A.u[(size_t)i * 4 + j] + A.x0[k];
  }
}

__global__ void k_update_res(MarchArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nnode) return;
  for (int j = 0; j < 3; j++) {
    const double v = A.u0[(size_t)i * 4 + j] + A.u[(size_t)i * 4 + j];
    A.u[(size_t)i * 4 + j] = v;
    A.u0[(size_t)i * 4 + j] = v;
  }
  if (i < A.nnodev) {
    A.dp[i] = A.dp[i] + A.u[(size_t)i * 4 + 3];
    const double v = A.u0[(size_t)i * 4 + 3] + A.u[(size_t)i * 4 + 3];
    A.u[(size_t)i * 4 + 3] = v;
    A.u0[(size_t)i * 4 + 3] = v;
  }
}

// ||R_u||^2 and ||R_p||^2 over the free equations (pressure flag per eq)
__global__ __launch_bounds__(PM_WG) void k_resnorm(int n, const double *b, const unsigned char *isp,
                                                   double *partials) {
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t a, e;
  pm_chunk(n, lb, gridDim.x, a, e);
  double v[2] = {0.0, 0.0};
  for (int64_t i = a + threadIdx.x; i < e; i += PM_WG) {
    const double bi = b[i];
    v[isp[i]] += bi * bi;
  }
  pm_block_partials<2>(v, partials);
}

__global__ __launch_bounds__(PM_WG) void k_resnorm_sum(int nblk, const double *partials,
                                                       double *out) {
  double r[2];
  pm_sum_partials<2>(nblk, partials, r);
  if (threadIdx.x == 0) {
    out[0] = r[0];
    out[1] = r[1];
  }
}

__global__ void k_copy(int n, double *dst, const double *src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static MarchArgs march_args(pm_ctx *c, double dt) {
  MarchArgs A;
  A.nnode = c->mesh.nnode;
  A.nnodev = c->mesh.nnodev;
  A.eq = c->d_eq;
  A.nload = c->have_nodal_loads ? c->d_nload : nullptr;
  A.f = c->have_nodal_loads ? c->d_f : nullptr;
  A.xv = c->d_xv;
  A.load = c->d_load;
  A.fload = c->d_fload;
  A.u = c->d_u;
  A.u0 = c->d_u0;
  A.dp = c->d_dp;
  A.b0 = c->d_b0;
  A.x0 = c->d_x0;
  A.t = c->t;
  A.dt = dt;
  A.gmod = c->gravity_mod;
  return A;
}

// The same three sums in the reference's order (cal_residuo_pm,
// Residuo.f:39-136): b.b over the equations, then ||R_u||^2 and ||R_p||^2
// node by node.  One thread; reference-order mode only.
__global__ void k_resnorm_ref(int nnode, int neq, const int *eq, const double *b, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double all = 0.0, ru = 0.0, rp = 0.0;
  for (int i = 0; i < neq; i++) all = all + b[i] * b[i];
  for (int i = 0; i < nnode; i++)
    for (int j = 0; j < 3; j++) {
      const int q = eq[(size_t)i * 4 + j];
      if (q >= 0) ru = ru + b[q] * b[q];
    }
  for (int i = 0; i < nnode; i++) {
    const int q = eq[(size_t)i * 4 + 3];
    if (q >= 0) rp = rp + b[q] * b[q];
  }
  out[0] = ru;
  out[1] = rp;
  out[2] = all;
}

// residual norms on the host: ||R_u||^2, ||R_p||^2 and ||R||^2
static int residual_norms(pm_ctx *c, const double *b, double *ru2, double *rp2, double *r2) {
  int n = c->mesh.neq;
  double *out = c->d_wk[11];
  if (c->ordered) {
    hipLaunchKernelGGL(k_resnorm_ref, dim3(1), dim3(64), 0, c->stream, c->mesh.nnode, n, c->d_eq,
                       b, out);
  } else {
    int g = pm_grid_for(n);
    hipLaunchKernelGGL(k_resnorm, dim3(g), dim3(PM_WG), 0, c->stream, n, b, c->d_isp,
                       c->d_partials);
    hipLaunchKernelGGL(k_resnorm_sum, dim3(1), dim3(PM_WG), 0, c->stream, g, c->d_partials, out);
  }
  double h[3];
  PM_CHECK_HIP(hipMemcpyAsync(h, out, sizeof(double) * (c->ordered ? 3 : 2),
                              hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  *ru2 = h[0];
  *rp2 = h[1];
  *r2 = c->ordered ? h[2] : h[0] + h[1];
  return PM_OK;
}

static float elapsed(pm_ctx *c, hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  hipEventSynchronize(b);
  hipEventElapsedTime(&ms, a, b);
  return ms;
}

int pm_step_dev(pm_ctx *c, double dt, const pm_solver_cfg *cfg, pm_step_info *info) {
  memset(info, 0, sizeof(*info));
  if (!c->have_mesh || !c->have_materials || !c->have_matrix) return PM_ERR_STATE;
  const int nn = c->mesh.nnode, neq = c->mesh.neq;
  const int tb = 256, gn = (nn + tb - 1) / tb, ge = (neq + tb - 1) / tb;
  PM_TRY(pm_ensure_workspace(c, neq));
  hipEvent_t e0, e1, ea, eb;
  PM_CHECK_HIP(hipEventCreate(&e0));
  PM_CHECK_HIP(hipEventCreate(&e1));
  PM_CHECK_HIP(hipEventCreate(&ea));
  PM_CHECK_HIP(hipEventCreate(&eb));
  int rc = PM_OK;
  double ms_elem = 0.0, ms_solve = 0.0;
  hipEventRecord(e0, c->stream);
  c->t = c->t + dt;
  MarchArgs M = march_args(c, dt);
  hipLaunchKernelGGL(k_pload, dim3(gn), dim3(tb), 0, c->stream, M);
  hipEventRecord(ea, c->stream);
  rc = pm_pform_dev(c, 4, 0, 1, dt, c->t, c->d_b0);
  hipEventRecord(eb, c->stream);
  ms_elem += elapsed(c, ea, eb);
  if (rc == PM_OK) {
    hipLaunchKernelGGL(k_predict, dim3(gn), dim3(tb), 0, c->stream, M);
    int lhs = 1, i = 1;
    double s0 = 0.0, s1 = 0.0;  // saved resid0 values (res0 / resu0, resp0)
    const int nbox = pm_batch_nbox(c);
    std::vector<double> bs0(nbox, 0.0), bs1(nbox, 0.0);  // the same, per box
    for (;;) {
      hipLaunchKernelGGL(k_copy, dim3(ge), dim3(tb), 0, c->stream, neq, c->d_b, c->d_b0);
      hipEventRecord(ea, c->stream);
      rc = pm_pform_dev(c, 2, lhs, 1, dt, c->t, c->d_b);
      hipEventRecord(eb, c->stream);
      ms_elem += elapsed(c, ea, eb);
      if (rc) break;
      info->nlit = i;
      if (nbox > 0) {  // every box runs its own nonlinear stop test
        std::vector<double> ru2v, rp2v;
        if ((rc = pm_batch_resnorms(c, c->d_b, ru2v, rp2v))) break;
        std::vector<int> active(nbox, 0);
        int nactive = 0;
        for (int k = 0; k < nbox; k++) {
          int ist = 0;
          if (cfg->stop_crit == 1) {
            const double res = sqrt(ru2v[k] + rp2v[k]);
            if (i == 1) bs0[k] = res;
            else if (i == 2) bs0[k] = fmax(bs0[k], res);
            if (res / bs0[k] < cfg->nltol) ist = 2;
          } else {
            const double gu = sqrt(ru2v[k]), gp = sqrt(rp2v[k]);
            if (i == 1) {
              bs0[k] = gu;
              bs1[k] = gp;
            } else if (i == 2) {
              bs0[k] = fmax(bs0[k], gu);
              bs1[k] = fmax(bs1[k], gu);
            }
            if (gu / bs0[k] < cfg->nltol) ist++;
            if (gp / bs1[k] < cfg->nltol) ist++;
          }
          active[k] = ist != 2;
          nactive += active[k];
        }
        if (nactive == 0) break;
        if ((rc = pm_batch_mask_rhs(c, c->d_b, active))) break;
        pm_solve_info si;
        rc = pm_solve_dev(c, cfg, c->d_b, c->d_x0, &si);
        if (info->nsolve < 16) {
          info->iters[info->nsolve] = si.iters;
          info->used[info->nsolve] = si.used_solver;
          info->true_rel[info->nsolve] = si.true_rel;
        }
        info->nsolve++;
        ms_solve += si.ms;
        if (rc) break;
        M = march_args(c, dt);
        hipLaunchKernelGGL(k_delta_update, dim3(gn), dim3(tb), 0, c->stream, M);
        if (i >= cfg->maxnlit) {
          if (!cfg->maxnlit_accept) rc = PM_ERR_NLIT;
          break;
        }
        i++;
        lhs = 0;
        continue;
      }
      double ru2, rp2, r2;
      if ((rc = residual_norms(c, c->d_b, &ru2, &rp2, &r2))) break;
      int istop = 0;
      if (cfg->stop_crit == 1) {
        const double res = sqrt(r2);
        if (i == 1) s0 = res;
        else if (i == 2) s0 = fmax(s0, res);
        if (res / s0 < cfg->nltol) istop = 2;
      } else {
        const double gu = sqrt(ru2), gp = sqrt(rp2);
        if (i == 1) {
          s0 = gu;
          s1 = gp;
        } else if (i == 2) {
          s0 = fmax(s0, gu);
          s1 = fmax(s1, gu);  // resp0 = max(resp0, gresu), Residuo.f:119
        }
        if (gu / s0 < cfg->nltol) istop++;
        if (gp / s1 < cfg->nltol) istop++;
      }
      if (istop == 2) break;
      pm_solve_info si;
      rc = pm_solve_dev(c, cfg, c->d_b, c->d_x0, &si);
      if (info->nsolve < 16) {
        info->iters[info->nsolve] = si.iters;
        info->used[info->nsolve] = si.used_solver;
        info->true_rel[info->nsolve] = si.true_rel;
      }
      info->nsolve++;
      ms_solve += si.ms;
      if (rc) break;
      M = march_args(c, dt);
      hipLaunchKernelGGL(k_delta_update, dim3(gn), dim3(tb), 0, c->stream, M);
      if (i >= cfg->maxnlit) {
        if (!cfg->maxnlit_accept) rc = PM_ERR_NLIT;
        break;
      }
      i++;
      lhs = 0;
    }
  }
  if (rc == PM_OK) {
    hipEventRecord(ea, c->stream);
    // update_prop (isw 10) before the nodal porosity (Main.f:1070-1093)
    rc = pm_update_prop_dev(c, dt);
    if (rc == PM_OK) rc = pm_porosity_dev(c, dt);
    hipEventRecord(eb, c->stream);
    ms_elem += elapsed(c, ea, eb);
  }
  if (rc == PM_OK) {
    M = march_args(c, dt);
    hipLaunchKernelGGL(k_update_res, dim3(gn), dim3(tb), 0, c->stream, M);
  }
  hipEventRecord(e1, c->stream);
  info->ms_total = elapsed(c, e0, e1);
  info->ms_element = ms_elem;
  info->ms_solve = ms_solve;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(ea);
  hipEventDestroy(eb);
  if (rc == PM_OK && hipGetLastError() != hipSuccess) rc = PM_ERR_HIP;
  return rc;
}
