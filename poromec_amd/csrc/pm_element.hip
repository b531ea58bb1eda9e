// pm_element.hip -- the hexa20/8 Biot element elmt17_pm (Elmlib03.f:1822-2636)
// on the device, with the pform_pm element loop (Pform.f:1-429) and the
// assembly of assbly_pm/csr (Assbly.f:89-304) fused into it.
//
// One workgroup per element.  Elements run colour by colour (no two elements
// of a colour share a node), so the additions into the matrix and the
// right-hand side need no atomics and are deterministic.  Instead of the
// reference's linear search of each column in its row (Assbly.f:283-288) a
// node pair (a, b) finds the column block of b in the row block of a once by
// binary search.
//
// Shape functions, Jacobian, strain and stress follow sfhexa8_m/sfhexa20_m
// (Elmlib.f:1510-1807), jacob3d_m (:2301-2396), deform3d (:2774) and
// stress3d (:2906); Gauss points are the reference's 15-digit table
// (Elmlib.f:3837-3903); integrals are summed point by point in the
// reference's (lz, ly, lx) order.
#include <cmath>

#include "pm_device.h"

#include <algorithm>

#include "pm_elemargs.h"
#include "pm_elemlib.h"

__device__ __forceinline__ bool is_elmt18(const ElemArgs &A, int el) {
  return A.ie && A.ie[A.mat[el]] == 18;
}

#define GP_CHUNK 16

// ---------------------------------------------------------------------------
// isw = 2: stiffness, internal force p = K_e u_e, scatter
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PM_WG) void k_elmt_isw2(ElemArgs A) {
  __shared__ double s_xv[8][3];
  __shared__ double s_u[PM_NST];
  __shared__ int s_ld[PM_NST];
  __shared__ double s_xji[GP_CHUNK][9];
  __shared__ double s_w[GP_CHUNK];
  __shared__ double s_hu[GP_CHUNK][20][3];  // global derivatives of the 20 u nodes
  __shared__ double s_hp[GP_CHUNK][8][4];   // value + global derivatives of 8 p nodes
  __shared__ double s_K[PM_NST][PM_NST + 1];
  __shared__ double s_mat[PM_PROP];
  __shared__ double s_gp[GP_CHUNK][6];  // elmt18_pm: a, b, c, alpha, 1/M, perm per point

  const int tid = threadIdx.x;
  const int el = A.elems[blockIdx.x];
  const int *nodes = A.ix + (size_t)el * 20;
  if (tid < 24) s_xv[tid / 3][tid % 3] = A.xv[(size_t)nodes[tid / 3] * 3 + tid % 3];
  if (tid < 60) {
    int no = nodes[tid / 3];
    s_u[tid] = A.u[(size_t)no * 4 + tid % 3];
    s_ld[tid] = A.eq[(size_t)no * 4 + tid % 3];
  } else if (tid < 68) {
    int no = nodes[tid - 60];
    s_u[tid] = A.u[(size_t)no * 4 + 3];
    s_ld[tid] = A.eq[(size_t)no * 4 + 3];
  }
  if (tid < PM_PROP) s_mat[tid] = A.e[(size_t)A.mat[el] * PM_PROP + tid];
  __syncthreads();
  const ElemMat M = elem_material(s_mat, A.dt, A.gmod);
  const bool v18 = is_elmt18(A, el);

  // work split: Kuu node pairs p0 = tid, p1 = tid + 256 (< 400);
  // Kpu (vertex i, node j) pair tid < 160; Kpp entry tid - 160 in [0, 64)
  double k0[9], k1[9], kp[3], kpp = 0.0;
  for (int i = 0; i < 9; i++) k0[i] = k1[i] = 0.0;
  kp[0] = kp[1] = kp[2] = 0.0;
  const int pa0 = tid / 20, pb0 = tid % 20;
  const int p1 = tid + 256, pa1 = p1 / 20, pb1 = p1 % 20;
  const bool has1 = p1 < 400;
  const int qa = tid / 20, qb = tid % 20;  // Kpu: vertex qa (< 8), node qb
  const bool hasq = tid < 160;
  const int ra = (tid - 160) >> 3, rb = (tid - 160) & 7;
  const bool hasr = tid >= 160 && tid < 224;
  const double wt_a = M.a;

  for (int lz = 0; lz < 4; lz++) {
    // geometry of the 16 points of this lz slice
    if (tid < GP_CHUNK) {
      const int lx = tid & 3, ly = tid >> 2;
      double xji[3][3];
      const double d = jacobian(s_xv, c_pg4[lx], c_pg4[ly], c_pg4[lz], xji);
      if (d <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
      for (int a = 0; a < 3; a++)
        for (int j = 0; j < 3; j++) s_xji[tid][3 * a + j] = xji[a][j];
      s_w[tid] = c_wg4[lx] * c_wg4[ly] * c_wg4[lz] * d;
      if (v18) {  // prop_porosity of point lz * 16 + tid (Elmlib03.f:2895)
        const GpProp P =
            prop_porosity(A.vp + ((size_t)el * 64 + lz * 16 + tid) * 7, M.fluid_d * A.gmod);
        s_gp[tid][0] = P.a;
        s_gp[tid][1] = P.b;
        s_gp[tid][2] = P.c;
        s_gp[tid][3] = P.cbiot;
        s_gp[tid][4] = P.ibiot;
        s_gp[tid][5] = P.perm;
      }
    }
    __syncthreads();
    for (int q = tid; q < GP_CHUNK * 28; q += PM_WG) {
      const int g = q / 28, k = q % 28;
      const int lx = g & 3, ly = g >> 2;
      const double r = c_pg4[lx], s = c_pg4[ly], t = c_pg4[lz];
      double xji[3][3];
      for (int a = 0; a < 3; a++)
        for (int j = 0; j < 3; j++) xji[a][j] = s_xji[g][3 * a + j];
      double dr, ds, dt, h, gx, gy, gz;
      if (k < 8) {
        dn8(k, r, s, t, dr, ds, dt);
        to_global(xji, dr, ds, dt, gx, gy, gz);
        s_hp[g][k][0] = n8(k, r, s, t);
        s_hp[g][k][1] = gx;
        s_hp[g][k][2] = gy;
        s_hp[g][k][3] = gz;
      } else {
        dn20(k - 8, r, s, t, h, dr, ds, dt);
        to_global(xji, dr, ds, dt, gx, gy, gz);
        s_hu[g][k - 8][0] = gx;
        s_hu[g][k - 8][1] = gy;
        s_hu[g][k - 8][2] = gz;
      }
    }
    __syncthreads();
    for (int g = 0; g < GP_CHUNK; g++) {
      const double w = s_w[g];
      const double wt1 = w * (v18 ? s_gp[g][0] : wt_a);
      const double gb = v18 ? s_gp[g][1] : M.b, gc = v18 ? s_gp[g][2] : M.c;
      {
        const double xi = s_hu[g][pa0][0], yi = s_hu[g][pa0][1], zi = s_hu[g][pa0][2];
        const double xj = s_hu[g][pb0][0], yj = s_hu[g][pb0][1], zj = s_hu[g][pb0][2];
        const double b = gb, c = gc;
        k0[0] += (xi * xj + c * (yi * yj + zi * zj)) * wt1;
        k0[1] += (b * xi * yj + c * yi * xj) * wt1;
        k0[2] += (b * xi * zj + c * zi * xj) * wt1;
        k0[3] += (b * yi * xj + c * xi * yj) * wt1;
        k0[4] += (yi * yj + c * (xi * xj + zi * zj)) * wt1;
        k0[5] += (b * yi * zj + c * zi * yj) * wt1;
        k0[6] += (b * zi * xj + c * xi * zj) * wt1;
        k0[7] += (b * zi * yj + c * yi * zj) * wt1;
        k0[8] += (zi * zj + c * (yi * yj + xi * xj)) * wt1;
      }
      if (has1) {
        const double xi = s_hu[g][pa1][0], yi = s_hu[g][pa1][1], zi = s_hu[g][pa1][2];
        const double xj = s_hu[g][pb1][0], yj = s_hu[g][pb1][1], zj = s_hu[g][pb1][2];
        const double b = gb, c = gc;
        k1[0] += (xi * xj + c * (yi * yj + zi * zj)) * wt1;
        k1[1] += (b * xi * yj + c * yi * xj) * wt1;
        k1[2] += (b * xi * zj + c * zi * xj) * wt1;
        k1[3] += (b * yi * xj + c * xi * yj) * wt1;
        k1[4] += (yi * yj + c * (xi * xj + zi * zj)) * wt1;
        k1[5] += (b * yi * zj + c * zi * yj) * wt1;
        k1[6] += (b * zi * xj + c * xi * zj) * wt1;
        k1[7] += (b * zi * yj + c * yi * zj) * wt1;
        k1[8] += (zi * zj + c * (yi * yj + xi * xj)) * wt1;
      }
      if (hasq) {
        const double wq = s_hp[g][qa][0] * (w * (v18 ? s_gp[g][3] : M.coef_biot));
        kp[0] -= s_hu[g][qb][0] * wq;
        kp[1] -= s_hu[g][qb][1] * wq;
        kp[2] -= s_hu[g][qb][2] * wq;
      }
      if (hasr) {
        // elmt18_pm weights Kpp by w dt perm (Elmlib03.f:3031)
        const double w1 = w * (v18 ? s_gp[g][4] : M.imod_biot);
        const double w2 = v18 ? w * A.dt * s_gp[g][5] : w * M.dt_perm;
        kpp = kpp - w1 * s_hp[g][ra][0] * s_hp[g][rb][0] -
              w2 * (s_hp[g][ra][1] * s_hp[g][rb][1] + s_hp[g][ra][2] * s_hp[g][rb][2] +
                    s_hp[g][ra][3] * s_hp[g][rb][3]);
      }
    }
    __syncthreads();
  }
  // element matrix into LDS
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) s_K[3 * pa0 + a][3 * pb0 + b] = k0[3 * a + b];
  if (has1)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) s_K[3 * pa1 + a][3 * pb1 + b] = k1[3 * a + b];
  if (hasq)
    for (int b = 0; b < 3; b++) {
      s_K[60 + qa][3 * qb + b] = kp[b];
      s_K[3 * qb + b][60 + qa] = kp[b];
    }
  if (hasr) s_K[60 + ra][60 + rb] = kpp;
  __syncthreads();

  // internal force p = K u (lku_m), subtracted from the right-hand side
  if (A.rhs && tid < PM_NST) {
    double p = 0.0;
    for (int j = 0; j < PM_NST; j++) p = p + s_K[tid][j] * s_u[j];
    if (A.pe) {
      A.pe[(size_t)el * PM_NST + tid] = p;
    } else {
      const int k = s_ld[tid];
      if (k >= 0) A.b[k] = A.b[k] - p;
    }
  }
  if (!A.lhs) return;
  if (A.ke) {  // element order assembly: keep K_e for the gather
    double *ke = A.ke + (size_t)el * PM_NST * PM_NST;
    for (int q = tid; q < PM_NST * PM_NST; q += PM_WG) ke[q] = s_K[q / PM_NST][q % PM_NST];
    return;
  }
  // scatter node-pair blocks: rows of node a, columns of node b
  for (int pr = tid; pr < 400; pr += PM_WG) {
    const int a = pr / 20, b = pr % 20;
    // local dofs of a node: u 3n..3n+2, p 60+n for vertices
    int la[4], lb[4], na = 0, nb = 0;
    for (int d = 0; d < 3; d++) {
      la[na++] = 3 * a + d;
      lb[nb++] = 3 * b + d;
    }
    if (a < 8) la[na++] = 60 + a;
    if (b < 8) lb[nb++] = 60 + b;
    // first free equation of each node (equations of a node are consecutive)
    int rowfirst = -1, colfirst = -1;
    for (int i = 0; i < na; i++)
      if (s_ld[la[i]] >= 0) { rowfirst = s_ld[la[i]]; break; }
    for (int j = 0; j < nb; j++)
      if (s_ld[lb[j]] >= 0) { colfirst = s_ld[lb[j]]; break; }
    if (rowfirst < 0 || colfirst < 0) continue;
    const int kb = A.eq_block[rowfirst];
    const int64_t cp = A.colptr[kb];
    const int cnt = (int)(A.colptr[kb + 1] - cp);
    int lo = 0, hi = cnt;  // lower_bound of colfirst
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (A.ncol[cp + mid] < colfirst) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= cnt || A.ncol[cp + lo] != colfirst) {  // pattern does not hold the pair
      atomicExch(A.err, PM_ERR_STATE);
      continue;
    }
    const int64_t vp = A.valptr[kb];
    const int f = A.fr[kb];
    for (int i = 0; i < na; i++) {
      const int row = s_ld[la[i]];
      if (row < 0) continue;
      double *vrow = A.val + vp + (int64_t)(row - f) * cnt + lo;
      for (int j = 0; j < nb; j++) {
        const int col = s_ld[lb[j]];
        if (col < 0) continue;
        const int off = col - colfirst;
        if (lo + off >= cnt) {
          atomicExch(A.err, PM_ERR_STATE);
          continue;
        }
        vrow[off] += s_K[la[i]][lb[j]];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// isw = 4: volume terms at t+dt from the state at t, face loads
// ---------------------------------------------------------------------------
__device__ void rot_face(double xf[4][3]) {
  double v1x = xf[1][0] - xf[0][0], v1y = xf[1][1] - xf[0][1], v1z = xf[1][2] - xf[0][2];
  double v2x = xf[2][0] - xf[0][0], v2y = xf[2][1] - xf[0][1], v2z = xf[2][2] - xf[0][2];
  double v3x = v1y * v2z - v1z * v2y, v3y = v1z * v2x - v1x * v2z, v3z = v1x * v2y - v1y * v2x;
  v2x = v1z * v3y - v1y * v3z;
  v2y = v1x * v3z - v1z * v3x;
  v2z = v1y * v3x - v1x * v3y;
  double n1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z);
  double n2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z);
  double n3 = sqrt(v3x * v3x + v3y * v3y + v3z * v3z);
  double r[3][3] = {{v1x / n1, v1y / n1, v1z / n1}, {v2x / n2, v2y / n2, v2z / n2},
                    {v3x / n3, v3y / n3, v3z / n3}};
  for (int i = 0; i < 4; i++) {
    double a = xf[i][0], b = xf[i][1], c = xf[i][2];
    xf[i][0] = r[0][0] * a + r[0][1] * b + r[0][2] * c;
    xf[i][1] = r[1][0] * a + r[1][1] * b + r[1][2] * c;
    xf[i][2] = r[2][0] * a + r[2][1] * b + r[2][2] * c;
  }
}

// face integrals of one element (Elmlib03.f:2373-2518), added to p
__device__ void face_loads(const ElemArgs &A, const int *iqm, const int *iqh,
                           const double (*xv)[3], double *p) {
  int cargm = 0, cargh = 0;
  double face_f[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < 6; j++) {
    if (!(iqm[j] > 0 || iqh[j] > 0)) continue;
    double xf[4][3];
    for (int i = 0; i < 4; i++)
      for (int d = 0; d < 3; d++) xf[i][d] = xv[c_face20[j][i]][d];
    if (iqm[j] != 0) cargm = A.load[3 * (iqm[j] - 1)];
    if (iqh[j] != 0) cargh = A.load[3 * (iqh[j] - 1)];
    if (cargm == 42) {
      atomicExch(A.err, PM_ERR_UNSUPPORTED);
      return;
    }
    if (cargm == 40 || cargm == 41) {
      if (cargm == 41) {
        double v1[3] = {xf[1][0] - xf[0][0], xf[1][1] - xf[0][1], xf[1][2] - xf[0][2]};
        double v2[3] = {xf[2][0] - xf[0][0], xf[2][1] - xf[0][1], xf[2][2] - xf[0][2]};
        face_f[0] = v2[1] * v1[2] - v2[2] * v1[1];
        face_f[1] = v2[2] * v1[0] - v2[0] * v1[2];
        face_f[2] = v2[0] * v1[1] - v2[1] * v1[0];
        double tmp = 1.0 / sqrt(face_f[0] * face_f[0] + face_f[1] * face_f[1] + face_f[2] * face_f[2]);
        for (int d = 0; d < 3; d++) face_f[d] *= tmp;
        if (j != 1)
          for (int d = 0; d < 3; d++) face_f[d] = -face_f[d];
      }
      rot_face(xf);
      const int nc = iqm[j];
      if (nc > 0) {
        const int nparc = A.load[3 * (nc - 1) + 2];
        if (cargm == 41) {
          double nca = dev_interpol(DFLOAD(1, 1, nc), DFLOAD(1, 2, nc), A.t, nparc);
          for (int d = 0; d < 3; d++) face_f[d] = nca * face_f[d];
        } else {
          for (int d = 0; d < 3; d++)
            face_f[d] = dev_interpol(DFLOAD(1, 1, nc), DFLOAD(1, 2 + d, nc), A.t, nparc);
        }
      }
      for (int ly = 0; ly < 3; ly++)
        for (int lx = 0; lx < 3; lx++) {
          const double r = c_pg3[lx], s = c_pg3[ly];
          const double hx[4] = {(1 + s) * 0.25, -(1 + s) * 0.25, -(1 - s) * 0.25, (1 - s) * 0.25};
          const double hy[4] = {(1 + r) * 0.25, (1 - r) * 0.25, -(1 - r) * 0.25, -(1 + r) * 0.25};
          double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
          for (int k = 0; k < 4; k++) {
            j00 += hx[k] * xf[k][0];
            j01 += hx[k] * xf[k][1];
            j10 += hy[k] * xf[k][0];
            j11 += hy[k] * xf[k][1];
          }
          const double det = j00 * j11 - j10 * j01;
          if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
          const double w = c_wg3[lx] * c_wg3[ly] * det;
          const double pr = 1 + r, ps = 1 + s, mr = 1 - r, ms = 1 - s;
          double h[8];
          h[4] = (1 - r * r) * ps * 0.5;
          h[5] = (1 - s * s) * mr * 0.5;
          h[6] = (1 - r * r) * ms * 0.5;
          h[7] = (1 - s * s) * pr * 0.5;
          h[0] = pr * ps * 0.25 - h[4] * 0.5 - h[7] * 0.5;
          h[1] = mr * ps * 0.25 - h[4] * 0.5 - h[5] * 0.5;
          h[2] = mr * ms * 0.25 - h[5] * 0.5 - h[6] * 0.5;
          h[3] = pr * ms * 0.25 - h[6] * 0.5 - h[7] * 0.5;
          for (int i = 0; i < 8; i++) {
            const int no = c_face20[j][i];
            const double wt1 = h[i] * w;
            for (int d = 0; d < 3; d++) p[3 * no + d] = p[3 * no + d] - face_f[d] * wt1;
          }
        }
    }
    if (cargh == 43) {
      rot_face(xf);
      double nca = 0.0;
      const int nc = iqh[j];
      if (nc > 0) {
        const int nparc = A.load[3 * (nc - 1) + 2];
        nca = dev_interpol(DFLOAD(1, 1, nc), DFLOAD(1, 2, nc), A.t, nparc);
      }
      for (int ly = 0; ly < 3; ly++)
        for (int lx = 0; lx < 3; lx++) {
          const double r = c_pg3[lx], s = c_pg3[ly];
          const double hx[4] = {(1 + s) * 0.25, -(1 + s) * 0.25, -(1 - s) * 0.25, (1 - s) * 0.25};
          const double hy[4] = {(1 + r) * 0.25, (1 - r) * 0.25, -(1 - r) * 0.25, -(1 + r) * 0.25};
          const double h4[4] = {(1 + r) * (1 + s) * 0.25, (1 - r) * (1 + s) * 0.25,
                                (1 - r) * (1 - s) * 0.25, (1 + r) * (1 - s) * 0.25};
          double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
          for (int k = 0; k < 4; k++) {
            j00 += hx[k] * xf[k][0];
            j01 += hx[k] * xf[k][1];
            j10 += hy[k] * xf[k][0];
            j11 += hy[k] * xf[k][1];
          }
          const double det = j00 * j11 - j10 * j01;
          if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
          const double w = c_wg3[lx] * c_wg3[ly] * det * A.dt * nca;
          for (int i = 0; i < 4; i++) {
            const int no = c_face20[j][i];
            p[60 + no] = p[60 + no] - h4[i] * w;
          }
        }
    }
  }
}

__global__ __launch_bounds__(PM_WG) void k_elmt_isw4(ElemArgs A) {
  __shared__ double s_xv[8][3];
  __shared__ double s_u[PM_NST];
  __shared__ double s_dp[8];
  __shared__ int s_ld[PM_NST];
  __shared__ double s_mat[PM_PROP];
  // per Gauss point, the two terms the reference adds and subtracts in turn:
  // stress (flux) term, then gravity (permeation) term
  __shared__ double s_c[64][PM_NST + 1];
  __shared__ double s_c2[64][PM_NST + 1];
  __shared__ double s_p[PM_NST];

  const int tid = threadIdx.x;
  const int el = A.elems[blockIdx.x];
  const int *nodes = A.ix + (size_t)el * 20;
  if (tid < 24) s_xv[tid / 3][tid % 3] = A.xv[(size_t)nodes[tid / 3] * 3 + tid % 3];
  if (tid < 60) {
    int no = nodes[tid / 3];
    s_u[tid] = A.u[(size_t)no * 4 + tid % 3];
    s_ld[tid] = A.eq[(size_t)no * 4 + tid % 3];
  } else if (tid < 68) {
    int no = nodes[tid - 60];
    s_u[tid] = A.u[(size_t)no * 4 + 3];
    s_ld[tid] = A.eq[(size_t)no * 4 + 3];
    s_dp[tid - 60] = A.dp[no];
  }
  if (tid < PM_PROP) s_mat[tid] = A.e[(size_t)A.mat[el] * PM_PROP + tid];
  __syncthreads();
  const ElemMat M = elem_material(s_mat, A.dt, A.gmod);

  // 4 threads per Gauss point: sub 0..2 own u nodes [7*sub, 7*sub+7), sub 3 the p rows
  const int g = tid >> 2, sub = tid & 3;
  const int lx = g & 3, ly = (g >> 2) & 3, lz = g >> 4;
  const double r = c_pg4[lx], s = c_pg4[ly], t = c_pg4[lz];
  double xji[3][3];
  const double det = jacobian(s_xv, r, s, t, xji);
  if (det <= 1.0e-14 && sub == 0) atomicExch(A.err, PM_ERR_JACOBIAN);
  const double w = c_wg4[lx] * c_wg4[ly] * c_wg4[lz] * det;
  // elmt18_pm: the point's properties, where the point index advances once
  // per ly (Elmlib03.f:3171-3189), so points read those of 1..16, 4 each
  const bool v18 = is_elmt18(A, el);
  GpProp G{M.a, M.b, M.c, M.imod_biot, M.coef_biot, 0.0};
  double gpm_d = M.pm_d;
  if (v18) {
    const double *vq = A.vp + ((size_t)el * 64 + lz * 4 + ly) * 7;
    G = prop_porosity(vq, M.fluid_d * A.gmod);
    gpm_d = vq[2];
  }
  if (sub < 3) {
    // strain of the 20-node field (deform3d), stress (stress3d), -alpha dp
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
    double hv[20], hgx[20], hgy[20], hgz[20];
    for (int k = 0; k < 20; k++) {
      double dr, ds, dt;
      dn20(k, r, s, t, hv[k], dr, ds, dt);
      to_global(xji, dr, ds, dt, hgx[k], hgy[k], hgz[k]);
      ux = ux + hgx[k] * s_u[3 * k];
      uy = uy + hgy[k] * s_u[3 * k];
      uz = uz + hgz[k] * s_u[3 * k];
      vx = vx + hgx[k] * s_u[3 * k + 1];
      vy = vy + hgy[k] * s_u[3 * k + 1];
      vz = vz + hgz[k] * s_u[3 * k + 1];
      wx = wx + hgx[k] * s_u[3 * k + 2];
      wy = wy + hgy[k] * s_u[3 * k + 2];
      wz = wz + hgz[k] * s_u[3 * k + 2];
    }
    const double e0 = ux, e1 = vy, e2 = wz, e3 = uy + vx, e4 = vz + wy, e5 = uz + wx;
    double tx[6];
    tx[0] = (e0 + G.b * e1 + G.b * e2) * G.a;
    tx[1] = (G.b * e0 + e1 + G.b * e2) * G.a;
    tx[2] = (G.b * e0 + G.b * e1 + e2) * G.a;
    tx[3] = e3 * G.c * G.a;
    tx[4] = e4 * G.c * G.a;
    tx[5] = e5 * G.c * G.a;
    double dpm = 0.0;
    for (int k = 0; k < 8; k++) dpm = dpm + n8(k, r, s, t) * s_dp[k];
    dpm = G.cbiot * dpm;
    tx[0] -= dpm;
    tx[1] -= dpm;
    tx[2] -= dpm;
    const double wt2 = w * gpm_d;
    const int k0 = 7 * sub, k1 = min(20, k0 + 7);
    for (int k = k0; k < k1; k++) {
      const double wt3 = hv[k] * wt2;
      s_c[g][3 * k] = w * (hgx[k] * tx[0] + hgy[k] * tx[3] + hgz[k] * tx[5]);
      s_c[g][3 * k + 1] = w * (hgx[k] * tx[3] + hgy[k] * tx[1] + hgz[k] * tx[4]);
      s_c[g][3 * k + 2] = w * (hgx[k] * tx[5] + hgy[k] * tx[4] + hgz[k] * tx[2]);
      s_c2[g][3 * k] = wt3 * A.grav[0];
      s_c2[g][3 * k + 1] = wt3 * A.grav[1];
      s_c2[g][3 * k + 2] = wt3 * A.grav[2];
    }
  } else {
    double gx[8], gy[8], gz[8];
    for (int k = 0; k < 8; k++) {
      double dr, ds, dt;
      dn8(k, r, s, t, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx[k], gy[k], gz[k]);
    }
    // elmt18_pm: w dt rho_f perm and w dt perm (Elmlib03.f:3242-3243)
    const double wt1 = v18 ? w * A.dt * M.fluid_d * G.perm : w * M.dt_fluid_perm;
    const double wt2 = v18 ? w * A.dt * G.perm : w * M.dt_perm;
    for (int i = 0; i < 8; i++) {
      double tmp = 0.0;
      for (int j = 0; j < 8; j++)
        tmp = tmp + wt2 * s_u[60 + j] * (gx[i] * gx[j] + gy[i] * gy[j] + gz[i] * gz[j]);
      s_c[g][60 + i] = wt1 * (gx[i] * A.grav[0] + gy[i] * A.grav[1] + gz[i] * A.grav[2]);
      s_c2[g][60 + i] = tmp;
    }
  }
  __syncthreads();
  if (tid < PM_NST) {
    double p = 0.0;
    for (int q = 0; q < 64; q++) {  // (lz, ly, lx) order, two updates per point
      p = p + s_c[q][tid];
      p = p - s_c2[q][tid];
    }
    s_p[tid] = p;
  }
  __syncthreads();
  if (tid == 0) {
    const int *eq14 = A.eload + (size_t)el * 14;
    int any = 0;
    for (int j = 0; j < 6; j++) any |= eq14[j] | eq14[7 + j];
    if (any) face_loads(A, eq14, eq14 + 7, s_xv, s_p);
  }
  __syncthreads();
  if (A.rhs && tid < PM_NST) {
    if (A.pe) {
      A.pe[(size_t)el * PM_NST + tid] = s_p[tid];
    } else {
      const int k = s_ld[tid];
      if (k >= 0) A.b[k] = A.b[k] - s_p[tid];
    }
  }
}

// ---------------------------------------------------------------------------
// isw = 7: nodal porosity change alpha tr(eps) + dp / M at the 8 vertices,
// summed per vertex (porosity_form, Pform.f:969-1176)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PM_WG) void k_elmt_isw7(ElemArgs A, double *dporo, double *pv) {
  const int tid = threadIdx.x;
  const int le = blockIdx.x * (PM_WG / 8) + (tid >> 3);
  const int v = tid & 7;
  if (le >= A.nelem) return;
  const int el = A.elems[le];
  const int *nodes = A.ix + (size_t)el * 20;
  double xv[8][3];
  for (int k = 0; k < 8; k++)
    for (int d = 0; d < 3; d++) xv[k][d] = A.xv[(size_t)nodes[k] * 3 + d];
  const double *e = A.e + (size_t)A.mat[el] * PM_PROP;
  const int no = nodes[v];
  if (is_elmt18(A, el)) {  // the points' porosity extrapolated (Elmlib03.f:3527-3538)
    double f[8];
    for (int j = 0; j < 8; j++) f[j] = A.vp[((size_t)el * 64 + c_gp_corner[j]) * 7];
    const double val = extrapol_vertex(f, v);
    if (pv) pv[(size_t)el * 8 + v] = val;
    else dporo[no] += val;
    return;
  }
  const double r = c_rn[v], s = c_sn[v], t = c_tn[v];
  double xji[3][3];
  const double det = jacobian(xv, r, s, t, xji);
  if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
  double ux = 0, vy = 0, wz = 0;
  for (int k = 0; k < 20; k++) {
    double h, dr, ds, dt, gx, gy, gz;
    dn20(k, r, s, t, h, dr, ds, dt);
    to_global(xji, dr, ds, dt, gx, gy, gz);
    const int no = nodes[k];
    ux = ux + gx * A.u[(size_t)no * 4];
    vy = vy + gy * A.u[(size_t)no * 4 + 1];
    wz = wz + gz * A.u[(size_t)no * 4 + 2];
  }
  const double val = e[4] * (ux + vy + wz) + (1.0 / e[3]) * A.u[(size_t)no * 4 + 3];
  if (pv) pv[(size_t)el * 8 + v] = val;  // gathered per node in element order
  else dporo[no] += val;  // colour order: no two elements of a colour share a vertex
}

// ---------------------------------------------------------------------------
// isw = 10 of elmt18_pm (Elmlib03.f:3558-3602) inside update_prop
// (Pform.f:1176-1330): one thread per integration point.  u holds the
// step's increments here.  The porosity gains alpha tr(du) + dp / M (the
// material's alpha and M); the permeability (Kozeny-Carman), density and
// Hashin-Shtrikman moduli of the new porosity are kept as the flags ask.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_elmt_isw10(ElemArgs A, int numel, int f2, int f3, int f4,
                                                    double *vp) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int el = (int)(gid >> 6), q = (int)(gid & 63);
  if (el >= numel || !is_elmt18(A, el)) return;
  const int lx = q & 3, ly = (q >> 2) & 3, lz = q >> 4;
  const int *nodes = A.ix + (size_t)el * 20;
  double xv[8][3];
  for (int k = 0; k < 8; k++)
    for (int d = 0; d < 3; d++) xv[k][d] = A.xv[(size_t)nodes[k] * 3 + d];
  const double *e = A.e + (size_t)A.mat[el] * PM_PROP;
  const double r = c_pg4[lx], s = c_pg4[ly], t = c_pg4[lz];
  double xji[3][3];
  const double det = jacobian(xv, r, s, t, xji);
  if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
  double ux = 0, vy = 0, wz = 0;
  for (int k = 0; k < 20; k++) {
    double h, dr, ds, dt, gx, gy, gz;
    dn20(k, r, s, t, h, dr, ds, dt);
    to_global(xji, dr, ds, dt, gx, gy, gz);
    const double *uk = A.u + (size_t)nodes[k] * 4;
    ux = ux + gx * uk[0];
    vy = vy + gy * uk[1];
    wz = wz + gz * uk[2];
  }
  double pi = 0.0;
  for (int k = 0; k < 8; k++) pi = pi + n8(k, r, s, t) * A.u[(size_t)nodes[k] * 4 + 3];
  pi = (1.0 / e[3]) * pi;
  const double tmp = e[4] * (ux + vy + wz) + pi;
  double *v = vp + ((size_t)el * 64 + q) * 7;
  const double poro = v[0] + tmp;
  v[0] = poro;
  if (f2) v[1] = konzey_carman(e[2], poro, e[7]);
  if (f3) v[2] = (1.0 - poro) * (e[5] * 1.0e-06) + poro * (e[6] * 1.0e-06);
  if (f4) hashin_shtrikman(poro, e[0], e[1], v[3], v[4], v[5], v[6]);
}

// ---------------------------------------------------------------------------
// element-order gathers (reference summation order)
// ---------------------------------------------------------------------------
__global__ void k_gather_vals(int64_t nnz, const int64_t *ptr, const int64_t *src,
                              const double *ke, double *val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  double v = 0.0;
  for (int64_t i = ptr[t]; i < ptr[t + 1]; i++) v = v + ke[src[i]];
  val[t] = v;
}

__global__ void k_gather_rhs(int neq, const int64_t *ptr, const int64_t *src, const double *pe,
                             double *b) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  double v = b[r];
  for (int64_t i = ptr[r]; i < ptr[r + 1]; i++) v = v - pe[src[i]];
  b[r] = v;
}

__global__ void k_gather_poro(int nnode, const int64_t *ptr, const int64_t *src, const double *pv,
                              double *dporo) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  double v = 0.0;
  for (int64_t i = ptr[n]; i < ptr[n + 1]; i++) v = v + pv[src[i]];
  dporo[n] = v;
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static ElemArgs elem_args(pm_ctx *c, double dt, double t, double *bvec, int lhs, int rhs) {
  ElemArgs A;
  A.ix = c->d_ix;
  A.xv = c->d_xv;
  A.eq = c->d_eq;
  A.mat = c->d_mat;
  A.e = c->d_e;
  A.eload = c->d_eload;
  A.load = c->d_load;
  A.fload = c->d_fload;
  A.u = c->d_u;
  A.dp = c->d_dp;
  A.b = bvec;
  A.err = c->d_elem_err;
  A.dt = dt;
  A.t = t;
  A.gmod = c->gravity_mod;
  for (int d = 0; d < 3; d++) A.grav[d] = c->gravity[d];
  A.eq_block = c->A.eq_block;
  A.fr = c->A.fr;
  A.colptr = c->A.colptr;
  A.valptr = c->A.valptr;
  A.ncol = c->A.ncol;
  A.val = c->A.val;
  A.lhs = lhs;
  A.rhs = rhs;
  A.ke = nullptr;
  A.pe = nullptr;
  A.ie = c->vprop[0] ? c->d_ie : nullptr;
  A.vp = c->d_vp;
  return A;
}

// element-order maps and element arrays for the reference summation order
static int ensure_ordered(pm_ctx *c) {
  if (c->have_ordered_maps) return PM_OK;
  const int64_t numel = c->mesh.numel;
  if (numel > 50000) return PM_ERR_UNSUPPORTED;  // element matrices would exceed ~1.8 GB
  PmHostPattern p;
  PM_TRY(pm_host_pattern(c->mesh, p));
  PmOrderedMaps o;
  PM_TRY(pm_host_ordered_maps(c->mesh, p, o));
  auto up = [&](int64_t **d, const std::vector<int64_t> &h) -> int {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    PM_CHECK_HIP(hipMalloc((void **)d, sizeof(int64_t) * std::max<size_t>(h.size(), 1)));
    PM_CHECK_HIP(hipMemcpy(*d, h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice));
    return PM_OK;
  };
  PM_TRY(up(&c->d_omptr, o.mptr));
  PM_TRY(up(&c->d_omsrc, o.msrc));
  PM_TRY(up(&c->d_orptr, o.rptr));
  PM_TRY(up(&c->d_orsrc, o.rsrc));
  PM_TRY(up(&c->d_opptr, o.pptr));
  PM_TRY(up(&c->d_opsrc, o.psrc));
  for (double **d : {&c->d_ke, &c->d_pe, &c->d_pv}) {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
  }
  PM_CHECK_HIP(hipMalloc((void **)&c->d_ke, sizeof(double) * numel * PM_NST * PM_NST));
  PM_CHECK_HIP(hipMalloc((void **)&c->d_pe, sizeof(double) * numel * PM_NST));
  PM_CHECK_HIP(hipMalloc((void **)&c->d_pv, sizeof(double) * numel * 8));
  c->have_ordered_maps = true;
  return PM_OK;
}

static int read_elem_err(pm_ctx *c) {
  int err = 0;
  PM_CHECK_HIP(hipMemcpyAsync(&err, c->d_elem_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return err;
}

int pm_pform_dev(pm_ctx *c, int isw, int lhs, int rhs, double dt, double t, double *bvec) {
  if (!c->have_mesh || !c->have_materials || !c->have_matrix) return PM_ERR_STATE;
  PM_CHECK_HIP(hipMemsetAsync(c->d_elem_err, 0, sizeof(int), c->stream));
  if (isw == 2 && lhs) {
    PM_CHECK_HIP(hipMemsetAsync(c->A.val, 0, sizeof(double) * (size_t)c->A.nnz, c->stream));
    c->A.values_valid = false;
    c->A.have32 = false;
  }
  ElemArgs A = elem_args(c, dt, t, bvec, lhs, rhs);
  const bool q1 = c->mesh.nen == 8;
  if (q1 && isw != 2 && isw != 4) return PM_ERR_ARG;
  if (c->ordered) {  // the Q1/Q1 element has no reference order to follow
    if (q1) return PM_ERR_UNSUPPORTED;
    PM_TRY(ensure_ordered(c));
    if (isw == 2 && lhs) A.ke = c->d_ke;
    if (rhs) A.pe = c->d_pe;
  }
  const int ncolour = (int)c->colour_ptr.size() - 1;
  for (int k = 0; k < ncolour; k++) {
    A.elems = c->d_colour_elems + c->colour_ptr[k];
    A.nelem = c->colour_ptr[k + 1] - c->colour_ptr[k];
    if (A.nelem == 0) continue;
    if (q1)
      PM_TRY(pm_q1_elements(A, isw, c->stream));
    else if (isw == 2)
      hipLaunchKernelGGL(k_elmt_isw2, dim3(A.nelem), dim3(PM_WG), 0, c->stream, A);
    else if (isw == 4)
      hipLaunchKernelGGL(k_elmt_isw4, dim3(A.nelem), dim3(PM_WG), 0, c->stream, A);
    else
      return PM_ERR_ARG;
  }
  if (c->ordered) {  // sums in element order, as the reference's assembly
    if (A.ke) {
      const int64_t nnz = c->A.nnz;
      hipLaunchKernelGGL(k_gather_vals, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0,
                         c->stream, nnz, c->d_omptr, c->d_omsrc, c->d_ke, c->A.val);
    }
    if (A.pe) {
      const int neq = c->mesh.neq;
      hipLaunchKernelGGL(k_gather_rhs, dim3((neq + 255) / 256), dim3(256), 0, c->stream, neq,
                         c->d_orptr, c->d_orsrc, c->d_pe, bvec);
    }
  }
  PM_CHECK_HIP(hipGetLastError());
  int err = read_elem_err(c);
  if (err) return err;
  if (isw == 2 && lhs) {
    PM_TRY(pm_extract_diag(c));
    c->A.values_valid = true;
  }
  return PM_OK;
}

// with variable properties the elements give the porosity itself, not its
// change (porosity_form, Pform.f:1158-1172)
__global__ void k_poro_update(int nnodev, const int *ic, const double *dporo, double *poro,
                              int absolute) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnodev) poro[i] = absolute ? dporo[i] / ic[i] : poro[i] + dporo[i] / ic[i];
}

int pm_porosity_dev(pm_ctx *c, double dt) {
  PM_CHECK_HIP(hipMemsetAsync(c->d_dporo, 0, sizeof(double) * c->mesh.nnode, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(c->d_elem_err, 0, sizeof(int), c->stream));
  ElemArgs A = elem_args(c, dt, c->t, nullptr, 0, 0);
  double *pv = nullptr;
  const bool q1 = c->mesh.nen == 8;
  if (c->ordered) {
    if (q1) return PM_ERR_UNSUPPORTED;
    PM_TRY(ensure_ordered(c));
    pv = c->d_pv;
  }
  const int ncolour = (int)c->colour_ptr.size() - 1;
  for (int k = 0; k < ncolour; k++) {
    A.elems = c->d_colour_elems + c->colour_ptr[k];
    A.nelem = c->colour_ptr[k + 1] - c->colour_ptr[k];
    if (A.nelem == 0) continue;
    if (q1) {
      PM_TRY(pm_q1_porosity(A, c->d_dporo, c->stream));
      continue;
    }
    int g = (A.nelem + PM_WG / 8 - 1) / (PM_WG / 8);
    hipLaunchKernelGGL(k_elmt_isw7, dim3(g), dim3(PM_WG), 0, c->stream, A, c->d_dporo, pv);
  }
  if (pv) {
    const int nn = c->mesh.nnode;
    hipLaunchKernelGGL(k_gather_poro, dim3((nn + 255) / 256), dim3(256), 0, c->stream, nn,
                       c->d_opptr, c->d_opsrc, pv, c->d_dporo);
  }
  int nv = c->mesh.nnodev;
  hipLaunchKernelGGL(k_poro_update, dim3((nv + 255) / 256), dim3(256), 0, c->stream, nv, c->d_ic,
                     c->d_dporo, c->d_poro, c->vprop[0]);
  PM_CHECK_HIP(hipGetLastError());
  return read_elem_err(c);
}

int pm_update_prop_dev(pm_ctx *c, double dt) {
  if (!c->vprop[0]) return PM_OK;
  PM_CHECK_HIP(hipMemsetAsync(c->d_elem_err, 0, sizeof(int), c->stream));
  ElemArgs A = elem_args(c, dt, c->t, nullptr, 0, 0);
  const int64_t nthreads = (int64_t)c->mesh.numel * 64;
  hipLaunchKernelGGL(k_elmt_isw10, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0,
                     c->stream, A, c->mesh.numel, c->vprop[1], c->vprop[2], c->vprop[3], c->d_vp);
  PM_CHECK_HIP(hipGetLastError());
  return read_elem_err(c);
}
