// pm_adjoint.hip -- hand-written adjoint of the backward-Euler Biot march
// (the differentiation form; the reference has no counterpart).
//
// With the linear elastic element and one solve per step, one step of the
// solv macro (Main.f:876-1123) is  K(theta) s_n = F_n + C s_{n-1}, where
//   K = [Kuu(E)  Kup; Kpu  -(S + Lk(k))]   (elmt17_pm isw 2, Elmlib03.f:2053-2154)
//   C s = [0; Kpu u - S p]                 (what isw 4 leaves after cancelling K s)
// S = int N N / M, Lk = dt perm int grad N . grad N.  Only K depends on the
// per-element parameters: dK/dlogE_e = Kuu_e, dK/dlogk_e = -Lk_e.  For a
// misfit J = sum_n j_n(s_n) the adjoint states solve (K symmetric)
//   K lambda_n = dj_n/ds_n + C^T lambda_{n+1},   lambda_{N+1} = 0,
// and dJ/dtheta = -sum_n lambda_n^T (dK/dtheta) s_n, i.e. per element
//   dJ/dlogE_e = -sum_n int eps(lambda_n)^T D eps(u_n)
//   dJ/dlogk_e = +sum_n dt perm int grad lambda_p,n . grad p_n.
// The forward states are recomputed from checkpoints kept every k steps.
// The kernels serve the Q1/Q1 element (element type 7) too, with its 8
// trilinear displacement nodes and the 2 x 2 x 2 rule.
#include <algorithm>
#include <cstring>
#include <vector>

#include "pm_elemlib.h"
#include "pm_q1.h"
#include "pm_internal.h"

struct AdjArgs {
  const int *ix;        // (numel, nen): 20 hexa20/8 nodes or 8 Q1/Q1 vertices
  const double *xv;     // (nnodev, 3)
  const int *eq;        // (nnode, 4)
  const int *mat;       // (numel)
  const double *e;      // (numat, 15)
  const double *lam;    // (neq)
  const double *s;      // (nnode, 4) nodal state, prescribed values included
  double *out;          // (neq)   C^T lambda
  double *gE, *gk;      // (numel) gradients
  const int *elems;
  int nelem;
  double dt, gmod;
  double grav[3];
};

// Gauss point g of the element's volume rule: 4 x 4 x 4 for hexa20/8, the
// 2 x 2 x 2 rule (weights 1) for Q1/Q1; (lz, ly, lx) with lx fastest
template <bool Q1>
__device__ __forceinline__ void adj_point(int g, double &r, double &s, double &t, double &wg) {
  if (Q1) {
    r = c_q1_pg[g & 1];
    s = c_q1_pg[(g >> 1) & 1];
    t = c_q1_pg[g >> 2];
    wg = 1.0;
  } else {
    const int lx = g & 3, ly = (g >> 2) & 3, lz = g >> 4;
    r = c_pg4[lx];
    s = c_pg4[ly];
    t = c_pg4[lz];
    wg = c_wg4[lx] * c_wg4[ly] * c_wg4[lz];
  }
}

// global derivatives of displacement node k (20 serendipity or 8 trilinear)
template <bool Q1>
__device__ __forceinline__ void adj_ugrad(int k, double r, double s, double t,
                                          const double xji[3][3], double &gx, double &gy,
                                          double &gz) {
  double h, dr, ds, dt;
  if (Q1) dn8(k, r, s, t, dr, ds, dt);
  else dn20(k, r, s, t, h, dr, ds, dt);
  to_global(xji, dr, ds, dt, gx, gy, gz);
}

// C^T lambda = [Kup lambda_p; -S lambda_p], element by element, colour order.
// One workgroup of 64 threads per element, one thread per Gauss point.
template <bool Q1>
__global__ __launch_bounds__(64) void k_adj_coupling(AdjArgs A) {
  constexpr int NU = Q1 ? 8 : 20, NP = 3 * NU, NS = NP + 8, NG = Q1 ? 8 : 64;
  __shared__ double s_xv[8][3];
  __shared__ double s_lp[8];
  __shared__ double s_c[NG][NS + 1];
  __shared__ int s_ld[NS];
  __shared__ double s_mat[PM_PROP];
  const int tid = threadIdx.x;
  const int el = A.elems[blockIdx.x];
  const int *nodes = A.ix + (size_t)el * NU;
  if (tid < 24) s_xv[tid / 3][tid % 3] = A.xv[(size_t)nodes[tid / 3] * 3 + tid % 3];
  if (tid < 8) {
    const int q = A.eq[(size_t)nodes[tid] * 4 + 3];
    s_lp[tid] = q >= 0 ? A.lam[q] : 0.0;
  }
  if (tid < PM_PROP) s_mat[tid] = A.e[(size_t)A.mat[el] * PM_PROP + tid];
  for (int i = tid; i < NS; i += 64)
    s_ld[i] = i < NP ? A.eq[(size_t)nodes[i / 3] * 4 + i % 3] : A.eq[(size_t)nodes[i - NP] * 4 + 3];
  __syncthreads();
  const ElemMat M = elem_material(s_mat, A.dt, A.gmod);
  const int g = tid;
  if (g < NG) {
    double r, s, t, wg;
    adj_point<Q1>(g, r, s, t, wg);
    double xji[3][3];
    const double det = jacobian(s_xv, r, s, t, xji);
    const double w = wg * det;
    double q = 0.0, hp[8];
    for (int k = 0; k < 8; k++) {
      hp[k] = n8(k, r, s, t);
      q += hp[k] * s_lp[k];
    }
    for (int k = 0; k < NU; k++) {
      double gx, gy, gz;
      adj_ugrad<Q1>(k, r, s, t, xji, gx, gy, gz);
      const double f = -M.coef_biot * w * q;
      s_c[g][3 * k] = f * gx;
      s_c[g][3 * k + 1] = f * gy;
      s_c[g][3 * k + 2] = f * gz;
    }
    for (int k = 0; k < 8; k++) s_c[g][NP + k] = -M.imod_biot * w * hp[k] * q;
  }
  __syncthreads();
  for (int i = tid; i < NS; i += 64) {
    double v = 0.0;
    for (int p = 0; p < NG; p++) v += s_c[p][i];
    const int row = s_ld[i];
    if (row >= 0) A.out[row] += v;
  }
}

// per-element gradient contributions of one step
template <bool Q1>
__global__ __launch_bounds__(64) void k_adj_grad(AdjArgs A) {
  constexpr int NU = Q1 ? 8 : 20, NP = 3 * NU, NS = NP + 8, NG = Q1 ? 8 : 64;
  __shared__ double s_xv[8][3];
  __shared__ double s_l[NS], s_s[NS];
  __shared__ double s_mat[PM_PROP];
  __shared__ double s_g[NG][2];
  const int tid = threadIdx.x;
  const int el = blockIdx.x;
  const int *nodes = A.ix + (size_t)el * NU;
  if (tid < 24) s_xv[tid / 3][tid % 3] = A.xv[(size_t)nodes[tid / 3] * 3 + tid % 3];
  if (tid < PM_PROP) s_mat[tid] = A.e[(size_t)A.mat[el] * PM_PROP + tid];
  for (int i = tid; i < NS; i += 64) {
    const size_t ns = i < NP ? (size_t)nodes[i / 3] * 4 + i % 3 : (size_t)nodes[i - NP] * 4 + 3;
    const int q = A.eq[ns];
    s_l[i] = q >= 0 ? A.lam[q] : 0.0;
    s_s[i] = A.s[ns];
  }
  __syncthreads();
  const ElemMat M = elem_material(s_mat, A.dt, A.gmod);
  const int g = tid;
  if (g < NG) {
    double r, s, t, wg;
    adj_point<Q1>(g, r, s, t, wg);
    double xji[3][3];
    const double det = jacobian(s_xv, r, s, t, xji);
    const double w = wg * det;
    double el6[6] = {0, 0, 0, 0, 0, 0}, es6[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < NU; k++) {
      double gx, gy, gz;
      adj_ugrad<Q1>(k, r, s, t, xji, gx, gy, gz);
      const double lu = s_l[3 * k], lv = s_l[3 * k + 1], lw = s_l[3 * k + 2];
      const double su = s_s[3 * k], sv = s_s[3 * k + 1], sw = s_s[3 * k + 2];
      el6[0] += gx * lu; el6[1] += gy * lv; el6[2] += gz * lw;
      el6[3] += gy * lu + gx * lv; el6[4] += gz * lv + gy * lw; el6[5] += gz * lu + gx * lw;
      es6[0] += gx * su; es6[1] += gy * sv; es6[2] += gz * sw;
      es6[3] += gy * su + gx * sv; es6[4] += gz * sv + gy * sw; es6[5] += gz * su + gx * sw;
    }
    double tx[6];
    tx[0] = (es6[0] + M.b * es6[1] + M.b * es6[2]) * M.a;
    tx[1] = (M.b * es6[0] + es6[1] + M.b * es6[2]) * M.a;
    tx[2] = (M.b * es6[0] + M.b * es6[1] + es6[2]) * M.a;
    tx[3] = es6[3] * M.c * M.a;
    tx[4] = es6[4] * M.c * M.a;
    tx[5] = es6[5] * M.c * M.a;
    double eDe = 0.0;
    for (int i = 0; i < 6; i++) eDe += el6[i] * tx[i];
    double glx = 0, gly = 0, glz = 0, gsx = 0, gsy = 0, gsz = 0;
    for (int k = 0; k < 8; k++) {
      double dr, ds, dt, gx, gy, gz;
      dn8(k, r, s, t, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx, gy, gz);
      glx += gx * s_l[NP + k]; gly += gy * s_l[NP + k]; glz += gz * s_l[NP + k];
      gsx += gx * s_s[NP + k]; gsy += gy * s_s[NP + k]; gsz += gz * s_s[NP + k];
    }
    // d r_p / dlog k = -Lk p + G: the flux term and the gravity flux G
    // (wt1 grad N . g in the isw 4 residual) both scale with the permeability
    s_g[g][0] = -w * eDe;
    s_g[g][1] = w * M.dt_perm * (glx * gsx + gly * gsy + glz * gsz) -
                w * M.dt_fluid_perm * (glx * A.grav[0] + gly * A.grav[1] + glz * A.grav[2]);
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int p = 0; p < NG; p++) {
      a += s_g[p][0];
      b += s_g[p][1];
    }
    A.gE[el] += a;
    A.gk[el] += b;
  }
}

// sensor values of the state after a step: out[i] = u at the node/dof of equation seq[i]
__global__ void k_sensor_read(int nsens, const int *snode, const double *u, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsens) out[i] = u[snode[i]];
}

// misfit gradient 2 (value - target) into the adjoint right-hand side
__global__ void k_sensor_rhs(int nsens, const int *seq, const double *val, const double *target,
                             double *rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsens) rhs[seq[i]] += 2.0 * (val[i] - target[i]);
}

__global__ void k_zero(size_t n, double *x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

static AdjArgs adj_args(pm_ctx *c, double dt) {
  AdjArgs A;
  memset(&A, 0, sizeof(A));
  A.ix = c->d_ix;
  A.xv = c->d_xv;
  A.eq = c->d_eq;
  A.mat = c->d_mat;
  A.e = c->d_e;
  A.dt = dt;
  A.gmod = c->gravity_mod;
  for (int d = 0; d < 3; d++) A.grav[d] = c->gravity[d];
  return A;
}

// Forward march of nsteps from the current state with checkpoints, the
// misfit J = sum_n sum_i (u[sensor_i] after step n - target[n * nsens + i])^2
// over nodal dofs sensor_i = 4 * node + dof (free dofs only), then the
// backward pass.  Gradients are summed per material.  The state is left at
// the end of the march, as nsteps calls of pm_step would leave it.
int pm_adjoint_dev(pm_ctx *c, int nsteps, double dt, const pm_solver_cfg *cfg, int nsens,
                   const int *sensor_dof, const double *target, int ckpt_every, double *J,
                   double *grad_logE, double *grad_logk) {
  if (!c->have_mesh || !c->have_materials || !c->have_matrix) return PM_ERR_STATE;
  if (nsteps < 1 || nsens < 0 || ckpt_every < 1 || (nsens > 0 && (!sensor_dof || !target)))
    return PM_ERR_ARG;
  if (pm_batch_nbox(c) > 0) return PM_ERR_UNSUPPORTED;
  const int nn = c->mesh.nnode, neq = c->mesh.neq, numel = c->mesh.numel;
  const bool q1 = c->mesh.nen == 8;
  std::vector<int> seq(nsens);
  for (int i = 0; i < nsens; i++) {
    const int d = sensor_dof[i];
    if (d < 0 || d >= 4 * nn) return PM_ERR_ARG;
    seq[i] = c->mesh.id[d] - 1;  // id is (4, nnode): dof + 4 * node
    if (seq[i] < 0) return PM_ERR_ARG;
  }
  PM_TRY(pm_ensure_workspace(c, neq));
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  const size_t nst = (size_t)nn * 10;  // u (4), u0 (4), dp, porosity
  const size_t nu = (size_t)nn * 4;
  std::vector<void *> bufs;
  auto dalloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return p;
  };
  double *ck = (double *)dalloc(sizeof(double) * nst * (nck + 1));  // + the final state
  double *seg = (double *)dalloc(sizeof(double) * nu * ckpt_every);
  double *lam = (double *)dalloc(sizeof(double) * neq);
  double *lam_next = (double *)dalloc(sizeof(double) * neq);
  double *rhs = (double *)dalloc(sizeof(double) * neq);
  double *gE = (double *)dalloc(sizeof(double) * numel);
  double *gk = (double *)dalloc(sizeof(double) * numel);
  int *d_sdof = (int *)dalloc(sizeof(int) * nsens);
  int *d_seq = (int *)dalloc(sizeof(int) * nsens);
  double *d_val = (double *)dalloc(sizeof(double) * nsens * nsteps);
  double *d_tgt = (double *)dalloc(sizeof(double) * nsens * nsteps);
  int rc = PM_OK;
  for (void *p : bufs)
    if (!p) rc = PM_ERR_HIP;
  if (bufs.size() != 11) rc = PM_ERR_HIP;
  if (rc == PM_OK && nsens > 0) {
    if (hipMemcpy(d_sdof, sensor_dof, sizeof(int) * nsens, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_seq, seq.data(), sizeof(int) * nsens, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tgt, target, sizeof(double) * nsens * nsteps, hipMemcpyHostToDevice) !=
            hipSuccess)
      rc = PM_ERR_HIP;
  }
  std::vector<double> tck(nck + 1, 0.0);
  auto copy = [&](double *dst, const double *src, size_t n) {
    (void)hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream);
  };
  auto save = [&](int k) {
    double *p = ck + nst * k;
    copy(p, c->d_u, nu);
    copy(p + nu, c->d_u0, nu);
    copy(p + 2 * nu, c->d_dp, nn);
    copy(p + 2 * nu + nn, c->d_poro, nn);
    tck[k] = c->t;
  };
  auto restore = [&](int k) {
    const double *p = ck + nst * k;
    copy(c->d_u, p, nu);
    copy(c->d_u0, p + nu, nu);
    copy(c->d_dp, p + 2 * nu, nn);
    copy(c->d_poro, p + 2 * nu + nn, nn);
    c->t = tck[k];
  };
  auto zero = [&](double *x, size_t n) {
    if (n) hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, x);
  };
  const int gs = (nsens + 255) / 256;
  pm_step_info si;
  // forward, checkpoints at steps 0, k, 2k, ...
  for (int n = 0; n < nsteps && rc == PM_OK; n++) {
    if (n % ckpt_every == 0) save(n / ckpt_every);
    rc = pm_step_dev(c, dt, cfg, &si);
    if (rc == PM_OK && nsens > 0)
      hipLaunchKernelGGL(k_sensor_read, dim3(gs), dim3(256), 0, c->stream, nsens, d_sdof, c->d_u,
                         d_val + (size_t)n * nsens);
  }
  if (rc == PM_OK) save(nck);
  double Jv = 0.0;
  if (rc == PM_OK && nsens > 0) {
    std::vector<double> hv((size_t)nsens * nsteps);
    if (hipMemcpyAsync(hv.data(), d_val, sizeof(double) * hv.size(), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = PM_ERR_HIP;
    for (size_t i = 0; i < hv.size(); i++) {
      const double d = hv[i] - target[i];
      Jv += d * d;
    }
  }
  if (J) *J = Jv;
  // backward pass, segment by segment, from the last one
  if (rc == PM_OK) {
    zero(lam_next, neq);
    zero(gE, numel);
    zero(gk, numel);
    AdjArgs A = adj_args(c, dt);
    const int ncolour = (int)c->colour_ptr.size() - 1;
    for (int sg = nck - 1; sg >= 0 && rc == PM_OK; sg--) {
      const int n0 = sg * ckpt_every, n1 = std::min(nsteps, n0 + ckpt_every);
      restore(sg);
      for (int n = n0; n < n1 && rc == PM_OK; n++) {  // the segment's states s_n
        rc = pm_step_dev(c, dt, cfg, &si);
        if (rc == PM_OK) copy(seg + (size_t)(n - n0) * nu, c->d_u, nu);
      }
      for (int n = n1 - 1; n >= n0 && rc == PM_OK; n--) {
        // rhs = dJ/ds_n + C^T lambda_{n+1}
        zero(rhs, neq);
        A.lam = lam_next;
        A.out = rhs;
        for (int k = 0; k < ncolour; k++) {
          A.elems = c->d_colour_elems + c->colour_ptr[k];
          A.nelem = c->colour_ptr[k + 1] - c->colour_ptr[k];
          if (A.nelem == 0) continue;
          if (q1) hipLaunchKernelGGL(k_adj_coupling<true>, dim3(A.nelem), dim3(64), 0, c->stream, A);
          else hipLaunchKernelGGL(k_adj_coupling<false>, dim3(A.nelem), dim3(64), 0, c->stream, A);
        }
        if (nsens > 0)
          hipLaunchKernelGGL(k_sensor_rhs, dim3(gs), dim3(256), 0, c->stream, nsens, d_seq,
                             d_val + (size_t)n * nsens, d_tgt + (size_t)n * nsens, rhs);
        // K is symmetric, so K^T lambda_n = rhs is the forward operator's solve
        zero(lam, neq);
        pm_solve_info inf;
        rc = pm_solve_dev(c, cfg, rhs, lam, &inf);
        if (rc) break;
        A.lam = lam;
        A.s = seg + (size_t)(n - n0) * nu;
        A.gE = gE;
        A.gk = gk;
        if (q1) hipLaunchKernelGGL(k_adj_grad<true>, dim3(numel), dim3(64), 0, c->stream, A);
        else hipLaunchKernelGGL(k_adj_grad<false>, dim3(numel), dim3(64), 0, c->stream, A);
        std::swap(lam, lam_next);
      }
    }
  }
  if (rc == PM_OK) {
    std::vector<double> hE(numel), hk(numel);
    std::vector<int> hm(numel);
    if (hipMemcpyAsync(hE.data(), gE, sizeof(double) * numel, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipMemcpyAsync(hk.data(), gk, sizeof(double) * numel, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipMemcpyAsync(hm.data(), c->d_mat, sizeof(int) * numel, hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = PM_ERR_HIP;
    if (rc == PM_OK) {
      if (grad_logE) memset(grad_logE, 0, sizeof(double) * c->numat);
      if (grad_logk) memset(grad_logk, 0, sizeof(double) * c->numat);
      for (int e = 0; e < numel; e++) {
        if (grad_logE) grad_logE[hm[e]] += hE[e];
        if (grad_logk) grad_logk[hm[e]] += hk[e];
      }
    }
    restore(nck);  // the state at the end of the march
    // the matrix and diagonal are those of the same dt, so nothing to redo
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PM_OK) rc = PM_ERR_HIP;
  for (void *p : bufs) (void)hipFree(p);
  return rc;
}
