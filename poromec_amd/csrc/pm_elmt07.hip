// pm_elmt07.hip -- the equal-order Q1/Q1 hexa8 Biot element (element type 7)
// on the device: isw 2 (matrix, internal force, scatter), isw 4 (volume,
// previous-step and face terms) and isw 7 (vertex porosity change).
//
// The reference reserves this element (ELMT07_PM, Elmlib03.f:22-23, "NAO
// IMPLEMENTADO") and sizes a maxno-8 hexa8 mesh for it (nst = 32,
// Rdata.f:650-660); it is elmt17_pm's weak form (Elmlib03.f:2022-2256) with
// trilinear displacements: Kuu = int B^T D B, Kpu = -alpha int N div N_u,
// Kpp = -(int N N / M + dt perm int grad N . grad N), and optionally a
// pressure stabilization from the material's e(9) (k_q1_isw2).
//
// An element has 8 vertices, so one wavefront takes one element and a lane
// one vertex pair (a, b): its 4 x 4 block (3 x 3 Kuu, 3 Kup, 3 Kpu, Kpp)
// summed over the 8 Gauss points in the order lz, ly, lx.  A workgroup runs
// 4 elements of one colour (no shared vertices), so the pair blocks go into
// the node-block matrix with plain additions, like the hexa20/8 kernels.
#include <cmath>

#include "pm_elemargs.h"
#include "pm_elemlib.h"
#include "pm_q1.h"

#define Q1_EPW 4  // elements per workgroup of 256 lanes: one wavefront each

// 4 elements' shared data
struct Q1Shared {
  double xv[8][3];
  double u[Q1_NST];
  double mat[PM_PROP];
  double sh[8][8][4];  // Gauss point g, vertex k: N, dN/dx, dN/dy, dN/dz
  double w[8];         // det J at g (Gauss weights 1)
  int ld[Q1_NST];      // equation of each local dof (-1: none)
};

// loads one element's vertices, local state and material (lanes 0..31)
__device__ __forceinline__ void q1_load(const ElemArgs &A, int el, int lane, Q1Shared &S,
                                        double *dp8) {
  const int *nodes = A.ix + (size_t)el * 8;
  if (lane < 24) S.xv[lane / 3][lane % 3] = A.xv[(size_t)nodes[lane / 3] * 3 + lane % 3];
  if (lane < Q1_NST) {
    const int no = nodes[lane < 24 ? lane / 3 : lane - 24];
    const int d = lane < 24 ? lane % 3 : 3;
    S.u[lane] = A.u[(size_t)no * 4 + d];
    S.ld[lane] = A.eq[(size_t)no * 4 + d];
  }
  if (lane < PM_PROP) S.mat[lane] = A.e[(size_t)A.mat[el] * PM_PROP + lane];
  if (dp8 && lane < 8) dp8[lane] = A.dp[nodes[lane]];
}

// lane (g, k): vertex k's shape value and global derivatives at Gauss point g
__device__ __forceinline__ void q1_point_shapes(const ElemArgs &A, int lane, Q1Shared &S) {
  const int g = lane >> 3, k = lane & 7;
  const double r = c_q1_pg[g & 1], s = c_q1_pg[(g >> 1) & 1], t = c_q1_pg[g >> 2];
  double h[8], hr[8], hs[8], ht[8], ji[3][3];
  q1_shape(r, s, t, h, hr, hs, ht);
  const double det = q1_jacobian(S.xv, hr, hs, ht, ji);
  if (det <= 1.0e-14 && k == 0) atomicExch(A.err, PM_ERR_JACOBIAN);
  double gx, gy, gz;
  q1_grad(ji, hr[k], hs[k], ht[k], gx, gy, gz);
  S.sh[g][k][0] = h[k];
  S.sh[g][k][1] = gx;
  S.sh[g][k][2] = gy;
  S.sh[g][k][3] = gz;
  if (k == 0) S.w[g] = det;
}

// ---------------------------------------------------------------------------
// isw = 2: element matrix, internal force b -= K_e u_e, scatter
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_q1_isw2(ElemArgs A) {
  __shared__ Q1Shared sm[Q1_EPW];
  __shared__ double s_K[Q1_EPW][Q1_NST][Q1_NST + 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int le = blockIdx.x * Q1_EPW + wv;
  const bool live = le < A.nelem;  // every wave reaches every barrier
  const int el = live ? A.elems[le] : 0;
  Q1Shared &S = sm[wv];
  if (live) q1_load(A, el, lane, S, nullptr);
  __syncthreads();
  if (live) q1_point_shapes(A, lane, S);
  __syncthreads();
  const int a = lane >> 3, b = lane & 7;
  if (live) {
    const ElemMat M = elem_material(S.mat, A.dt, A.gmod);
    // pressure stabilization, e(9) = beta (0: none): -tau int grad N . grad N
    // in Kpp, tau = beta h^2 / (lambda + 2 mu), h = V^(1/3) -- the perturbation
    // tau d/dt (div grad p) of the flow equation (Aguilar, Gaspar, Lisbona,
    // Rodrigo, IJNME 75 (2008) 1282-1300), on the step's pressure change like
    // the storage term
    double tau = 0.0;
    if (S.mat[8] != 0.0) {
      double vol = 0.0;
      for (int g = 0; g < 8; g++) vol = vol + S.w[g];
      tau = S.mat[8] * cbrt(vol) * cbrt(vol) / M.a;
    }
    double kuu[9], kup[3], kpu[3], kpp = 0.0;
    for (int i = 0; i < 9; i++) kuu[i] = 0.0;
    for (int i = 0; i < 3; i++) kup[i] = kpu[i] = 0.0;
    for (int g = 0; g < 8; g++) {
      const double w = S.w[g];
      const double ha = S.sh[g][a][0], xi = S.sh[g][a][1], yi = S.sh[g][a][2], zi = S.sh[g][a][3];
      const double hb = S.sh[g][b][0], xj = S.sh[g][b][1], yj = S.sh[g][b][2], zj = S.sh[g][b][3];
      const double wt1 = w * M.a, bb = M.b, c = M.c;
      kuu[0] = kuu[0] + (xi * xj + c * (yi * yj + zi * zj)) * wt1;
      kuu[1] = kuu[1] + (bb * xi * yj + c * yi * xj) * wt1;
      kuu[2] = kuu[2] + (bb * xi * zj + c * zi * xj) * wt1;
      kuu[3] = kuu[3] + (bb * yi * xj + c * xi * yj) * wt1;
      kuu[4] = kuu[4] + (yi * yj + c * (xi * xj + zi * zj)) * wt1;
      kuu[5] = kuu[5] + (bb * yi * zj + c * zi * yj) * wt1;
      kuu[6] = kuu[6] + (bb * zi * xj + c * xi * zj) * wt1;
      kuu[7] = kuu[7] + (bb * zi * yj + c * yi * zj) * wt1;
      kuu[8] = kuu[8] + (zi * zj + c * (yi * yj + xi * xj)) * wt1;
      // pressure row of a, displacement columns of b: -alpha N_a grad N_b
      const double wqa = ha * (w * M.coef_biot);
      kpu[0] = kpu[0] - xj * wqa;
      kpu[1] = kpu[1] - yj * wqa;
      kpu[2] = kpu[2] - zj * wqa;
      // its transpose: displacement rows of a, pressure column of b
      const double wqb = hb * (w * M.coef_biot);
      kup[0] = kup[0] - xi * wqb;
      kup[1] = kup[1] - yi * wqb;
      kup[2] = kup[2] - zi * wqb;
      const double w1 = w * M.imod_biot, w2 = w * M.dt_perm;
      kpp = kpp - (w1 * ha * hb + w2 * (xi * xj + yi * yj + zi * zj));
      if (tau != 0.0) kpp = kpp - (w * tau) * (xi * xj + yi * yj + zi * zj);
    }
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) s_K[wv][3 * a + i][3 * b + j] = kuu[3 * i + j];
      s_K[wv][24 + a][3 * b + i] = kpu[i];
      s_K[wv][3 * a + i][24 + b] = kup[i];
    }
    s_K[wv][24 + a][24 + b] = kpp;
  }
  __syncthreads();
  if (!live) return;  // no barrier follows
  // internal force p = K u, each row summed over its 32 columns in order
  if (A.rhs && lane < Q1_NST) {
    double p = 0.0;
    for (int j = 0; j < Q1_NST; j++) p = p + s_K[wv][lane][j] * S.u[j];
    const int k = S.ld[lane];
    if (k >= 0) A.b[k] = A.b[k] - p;
  }
  if (!A.lhs) return;
  // the pair block into row block of vertex a, column block of vertex b
  const int la[4] = {3 * a, 3 * a + 1, 3 * a + 2, 24 + a};
  const int lb[4] = {3 * b, 3 * b + 1, 3 * b + 2, 24 + b};
  int rowfirst = -1, colfirst = -1;
  for (int i = 0; i < 4; i++)
    if (S.ld[la[i]] >= 0) {
      rowfirst = S.ld[la[i]];
      break;
    }
  for (int j = 0; j < 4; j++)
    if (S.ld[lb[j]] >= 0) {
      colfirst = S.ld[lb[j]];
      break;
    }
  if (rowfirst < 0 || colfirst < 0) return;
  const int kb = A.eq_block[rowfirst];
  const int64_t cp = A.colptr[kb];
  const int cnt = (int)(A.colptr[kb + 1] - cp);
  int lo = 0, hi = cnt;  // lower_bound of colfirst
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A.ncol[cp + mid] < colfirst) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= cnt || A.ncol[cp + lo] != colfirst) {  // pattern does not hold the pair
    atomicExch(A.err, PM_ERR_STATE);
    return;
  }
  const int64_t vp = A.valptr[kb];
  const int f = A.fr[kb];
  for (int i = 0; i < 4; i++) {
    const int row = S.ld[la[i]];
    if (row < 0) continue;
    double *vrow = A.val + vp + (int64_t)(row - f) * cnt + lo;
    for (int j = 0; j < 4; j++) {
      const int col = S.ld[lb[j]];
      if (col < 0) continue;
      const int off = col - colfirst;
      if (lo + off >= cnt) {
        atomicExch(A.err, PM_ERR_STATE);
        continue;
      }
      vrow[off] += s_K[wv][la[i]][lb[j]];
    }
  }
}

// ---------------------------------------------------------------------------
// isw = 4: volume terms at t+dt from the state at t, face loads
// ---------------------------------------------------------------------------
// the face's vertices in its plane (origin vertex 1, x along edge 1-2, z the
// right-hand normal (1-2) x (1-3), returned in nrm)
__device__ void q1_face_frame(const double (*xf)[3], double x2[4][2], double *nrm) {
  const double e1[3] = {xf[1][0] - xf[0][0], xf[1][1] - xf[0][1], xf[1][2] - xf[0][2]};
  const double e2[3] = {xf[2][0] - xf[0][0], xf[2][1] - xf[0][1], xf[2][2] - xf[0][2]};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
  const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  const double ax[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
  const double az[3] = {n[0] / ln, n[1] / ln, n[2] / ln};
  const double ay[3] = {az[1] * ax[2] - az[2] * ax[1], az[2] * ax[0] - az[0] * ax[2],
                        az[0] * ax[1] - az[1] * ax[0]};
  for (int i = 0; i < 4; i++) {
    const double d[3] = {xf[i][0] - xf[0][0], xf[i][1] - xf[0][1], xf[i][2] - xf[0][2]};
    x2[i][0] = d[0] * ax[0] + d[1] * ax[1] + d[2] * ax[2];
    x2[i][1] = d[0] * ay[0] + d[1] * ay[1] + d[2] * ay[2];
  }
  for (int d = 0; d < 3; d++) nrm[d] = az[d];
}

static __constant__ int c_q1_face[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 4, 5, 1},
                                           {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 5, 6, 2}};
static __constant__ double c_q1_qr[4] = {1, -1, -1, 1};
static __constant__ double c_q1_qs[4] = {1, 1, -1, -1};

// face integrals of one element, subtracted from p (types 40 and 41:
// traction on the displacement rows; 43: flux on the pressure rows)
__device__ void q1_face_loads(const ElemArgs &A, const int *iqm, const int *iqh,
                              const double (*xv)[3], double *p) {
  for (int j = 0; j < 6; j++) {
    const int nm = iqm[j], nh = iqh[j];
    if (nm <= 0 && nh <= 0) continue;
    double xf[4][3], x2[4][2], nrm[3];
    for (int i = 0; i < 4; i++)
      for (int d = 0; d < 3; d++) xf[i][d] = xv[c_q1_face[j][i]][d];
    q1_face_frame(xf, x2, nrm);
    double f[3] = {0.0, 0.0, 0.0}, qh = 0.0;
    if (nm > 0) {
      const int type = A.load[3 * (nm - 1)], nparc = A.load[3 * (nm - 1) + 2];
      if (type == 40) {
        for (int d = 0; d < 3; d++)
          f[d] = dev_interpol(DFLOAD(1, 1, nm), DFLOAD(1, 2 + d, nm), A.t, nparc);
      } else if (type == 41) {  // the value along the outward normal
        const double q = dev_interpol(DFLOAD(1, 1, nm), DFLOAD(1, 2, nm), A.t, nparc);
        const double sg = j == 1 ? 1.0 : -1.0;
        for (int d = 0; d < 3; d++) f[d] = -sg * nrm[d] * q;
      } else {
        atomicExch(A.err, PM_ERR_UNSUPPORTED);
        return;
      }
    }
    if (nh > 0) {
      const int nparc = A.load[3 * (nh - 1) + 2];
      qh = dev_interpol(DFLOAD(1, 1, nh), DFLOAD(1, 2, nh), A.t, nparc);
    }
    for (int ly = 0; ly < 2; ly++)
      for (int lx = 0; lx < 2; lx++) {
        const double r = c_q1_pg[lx], s = c_q1_pg[ly];
        double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
        for (int k = 0; k < 4; k++) {
          const double dr = 0.25 * c_q1_qr[k] * (1.0 + c_q1_qs[k] * s);
          const double ds = 0.25 * (1.0 + c_q1_qr[k] * r) * c_q1_qs[k];
          j00 = j00 + dr * x2[k][0];
          j01 = j01 + dr * x2[k][1];
          j10 = j10 + ds * x2[k][0];
          j11 = j11 + ds * x2[k][1];
        }
        const double w = j00 * j11 - j01 * j10;
        if (w <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
        for (int i = 0; i < 4; i++) {
          const double hi = 0.25 * (1.0 + c_q1_qr[i] * r) * (1.0 + c_q1_qs[i] * s) * w;
          const int no = c_q1_face[j][i];
          if (nm > 0)
            for (int d = 0; d < 3; d++) p[3 * no + d] = p[3 * no + d] - f[d] * hi;
          if (nh > 0) p[24 + no] = p[24 + no] - A.dt * qh * hi;
        }
      }
  }
}

__global__ __launch_bounds__(256) void k_q1_isw4(ElemArgs A) {
  __shared__ Q1Shared sm[Q1_EPW];
  __shared__ double s_dp[Q1_EPW][8];
  __shared__ double s_tx[Q1_EPW][8][6];  // D eps - alpha dp at each point
  __shared__ double s_gp[Q1_EPW][8][3];  // grad p at each point
  __shared__ double s_p[Q1_EPW][Q1_NST];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int le = blockIdx.x * Q1_EPW + wv;
  const bool live = le < A.nelem;
  const int el = live ? A.elems[le] : 0;
  Q1Shared &S = sm[wv];
  if (live) q1_load(A, el, lane, S, s_dp[wv]);
  __syncthreads();
  if (live) q1_point_shapes(A, lane, S);
  __syncthreads();
  const ElemMat M = elem_material(S.mat, A.dt, A.gmod);
  if (live && lane < 8) {  // the point's stress and pressure gradient
    const int g = lane;
    double hx[8], hy[8], hz[8], eps[6], tx[6];
    for (int k = 0; k < 8; k++) {
      hx[k] = S.sh[g][k][1];
      hy[k] = S.sh[g][k][2];
      hz[k] = S.sh[g][k][3];
    }
    q1_strain(hx, hy, hz, S.u, eps);
    q1_stress(M.a, M.b, M.c, eps, tx);
    double dpm = 0.0;
    for (int k = 0; k < 8; k++) dpm = dpm + S.sh[g][k][0] * s_dp[wv][k];
    dpm = M.coef_biot * dpm;
    tx[0] = tx[0] - dpm;
    tx[1] = tx[1] - dpm;
    tx[2] = tx[2] - dpm;
    double gpx = 0, gpy = 0, gpz = 0;
    for (int k = 0; k < 8; k++) {
      gpx = gpx + hx[k] * S.u[24 + k];
      gpy = gpy + hy[k] * S.u[24 + k];
      gpz = gpz + hz[k] * S.u[24 + k];
    }
    for (int i = 0; i < 6; i++) s_tx[wv][g][i] = tx[i];
    s_gp[wv][g][0] = gpx;
    s_gp[wv][g][1] = gpy;
    s_gp[wv][g][2] = gpz;
  }
  __syncthreads();
  if (live && lane < Q1_NST) {  // row lane, two updates per point as the element adds them
    double p = 0.0;
    for (int g = 0; g < 8; g++) {
      const double w = S.w[g];
      if (lane < 24) {
        const int i = lane / 3, d = lane % 3;
        const double xi = S.sh[g][i][1], yi = S.sh[g][i][2], zi = S.sh[g][i][3];
        const double *tx = s_tx[wv][g];
        const double sd = d == 0 ? xi * tx[0] + yi * tx[3] + zi * tx[5]
                        : d == 1 ? xi * tx[3] + yi * tx[1] + zi * tx[4]
                                 : xi * tx[5] + yi * tx[4] + zi * tx[2];
        p = p + w * sd;
        const double wt = S.sh[g][i][0] * (w * M.pm_d);
        p = p - wt * A.grav[d];
      } else {
        const int i = lane - 24;
        const double xi = S.sh[g][i][1], yi = S.sh[g][i][2], zi = S.sh[g][i][3];
        p = p + w * M.dt_fluid_perm * (xi * A.grav[0] + yi * A.grav[1] + zi * A.grav[2]);
        p = p - w * M.dt_perm *
                    (xi * s_gp[wv][g][0] + yi * s_gp[wv][g][1] + zi * s_gp[wv][g][2]);
      }
    }
    s_p[wv][lane] = p;
  }
  __syncthreads();
  if (live && lane == 0) {
    const int *eq14 = A.eload + (size_t)el * 14;
    int any = 0;
    for (int j = 0; j < 6; j++) any |= eq14[j] | eq14[7 + j];
    if (any) q1_face_loads(A, eq14, eq14 + 7, S.xv, s_p[wv]);
  }
  __syncthreads();
  if (live && A.rhs && lane < Q1_NST) {
    const int k = S.ld[lane];
    if (k >= 0) A.b[k] = A.b[k] - s_p[wv][lane];
  }
}

// ---------------------------------------------------------------------------
// isw = 7: alpha tr(eps) + p / M at the 8 vertices, added per vertex
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_q1_isw7(ElemArgs A, double *dporo) {
  const int tid = threadIdx.x;
  const int le = blockIdx.x * 32 + (tid >> 3);
  const int v = tid & 7;
  if (le >= A.nelem) return;
  const int el = A.elems[le];
  const int *nodes = A.ix + (size_t)el * 8;
  double xv[8][3], u[24];
  for (int k = 0; k < 8; k++)
    for (int d = 0; d < 3; d++) {
      xv[k][d] = A.xv[(size_t)nodes[k] * 3 + d];
      u[3 * k + d] = A.u[(size_t)nodes[k] * 4 + d];
    }
  const double *e = A.e + (size_t)A.mat[el] * PM_PROP;
  double h[8], hr[8], hs[8], ht[8], ji[3][3], hx[8], hy[8], hz[8], eps[6];
  q1_shape(c_q1_rn[v], c_q1_sn[v], c_q1_tn[v], h, hr, hs, ht);
  const double det = q1_jacobian(xv, hr, hs, ht, ji);
  if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
  for (int k = 0; k < 8; k++) q1_grad(ji, hr[k], hs[k], ht[k], hx[k], hy[k], hz[k]);
  q1_strain(hx, hy, hz, u, eps);
  const int no = nodes[v];
  const double val = e[4] * (eps[0] + eps[1] + eps[2]) + (1.0 / e[3]) * A.u[(size_t)no * 4 + 3];
  dporo[no] += val;  // colour order: no two elements of a colour share a vertex
}

int pm_q1_elements(const ElemArgs &A, int isw, hipStream_t stream) {
  const dim3 grid((A.nelem + Q1_EPW - 1) / Q1_EPW);
  if (isw == 2) hipLaunchKernelGGL(k_q1_isw2, grid, dim3(256), 0, stream, A);
  else if (isw == 4) hipLaunchKernelGGL(k_q1_isw4, grid, dim3(256), 0, stream, A);
  else return PM_ERR_ARG;
  return hipGetLastError() == hipSuccess ? PM_OK : PM_ERR_HIP;
}

int pm_q1_porosity(const ElemArgs &A, double *dporo, hipStream_t stream) {
  hipLaunchKernelGGL(k_q1_isw7, dim3((A.nelem + 31) / 32), dim3(256), 0, stream, A, dporo);
  return hipGetLastError() == hipSuccess ? PM_OK : PM_ERR_HIP;
}
