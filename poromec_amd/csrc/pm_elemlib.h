// pm_elemlib.h -- device element library of the hexa20/8 Biot element
// (elmt17_pm, Elmlib03.f:1822-2636): Gauss table (Elmlib.f:3837-3903), hexa8
// and serendipity hexa20 shape functions (Elmlib.f:1510-1807), Jacobian
// (jacob3d_m, Elmlib.f:2301-2396) and material constants.
#pragma once
#include "pm_device.h"

static __constant__ double c_pg4[4] = {-.861136311594053, -.339981043584856, .339981043584856,
                                .861136311594053};
static __constant__ double c_wg4[4] = {.347854845137454, .652145154862546, .652145154862546,
                                .347854845137454};
static __constant__ double c_pg3[3] = {-.774596669241483, 0.0, .774596669241483};
static __constant__ double c_wg3[3] = {.555555555555556, .888888888888889, .555555555555556};
static __constant__ int c_face20[6][8] = {{0, 1, 2, 3, 8, 9, 10, 11},   {4, 5, 6, 7, 12, 13, 14, 15},
                                   {0, 4, 5, 1, 16, 12, 17, 8},  {3, 2, 6, 7, 10, 18, 14, 19},
                                   {0, 3, 7, 4, 11, 19, 15, 16}, {1, 5, 6, 2, 17, 13, 18, 9}};
static __constant__ double c_rn[8] = {1, -1, -1, 1, 1, -1, -1, 1};
static __constant__ double c_sn[8] = {1, 1, -1, -1, 1, 1, -1, -1};
static __constant__ double c_tn[8] = {1, 1, 1, 1, -1, -1, -1, -1};

// natural derivatives of hexa8 node k at (r, s, t)
__device__ __forceinline__ void dn8(int k, double r, double s, double t, double &dr, double &ds,
                                    double &dt) {
  const double sr = c_rn[k], ss = c_sn[k], st = c_tn[k];
  dr = sr * (1.0 + ss * s) * (1.0 + st * t) * 0.125;
  ds = (1.0 + sr * r) * ss * (1.0 + st * t) * 0.125;
  dt = (1.0 + sr * r) * (1.0 + ss * s) * st * 0.125;
}
__device__ __forceinline__ double n8(int k, double r, double s, double t) {
  return (1.0 + c_rn[k] * r) * (1.0 + c_sn[k] * s) * (1.0 + c_tn[k] * t) * 0.125;
}

// natural derivatives (and value) of serendipity hexa20 node k (0..19)
__device__ void dn20(int k, double r, double s, double t, double &h, double &dr, double &ds,
                     double &dt) {
  if (k < 8) {
    const double a = c_rn[k], b = c_sn[k], c = c_tn[k];
    const double pr = 1.0 + a * r, ps = 1.0 + b * s, pt = 1.0 + c * t;
    const double g = a * r + b * s + c * t - 2.0;
    h = pr * ps * pt * g * 0.125;
    dr = a * ps * pt * (g + pr) * 0.125;
    ds = b * pr * pt * (g + ps) * 0.125;
    dt = c * pr * ps * (g + pt) * 0.125;
    return;
  }
  // mid-edge nodes: 8..11 top (t = +1), 12..15 bottom (t = -1), 16..19 (t = 0)
  const int e = k - 8;
  const int grp = e >> 2, idx = e & 3;
  // position on the top ring of each group: 9 (r=0,s=+), 10 (r=-,s=0), 11 (r=0,s=-), 12 (r=+,s=0)
  if (grp < 2) {
    const double c = grp == 0 ? 1.0 : -1.0, pt = 1.0 + c * t;
    if ((idx & 1) == 0) {  // r = 0 edge, s = +-1
      const double b = idx == 0 ? 1.0 : -1.0, ps = 1.0 + b * s, mrr = 1.0 - r * r;
      h = mrr * ps * pt * 0.25;
      dr = -2.0 * r * ps * pt * 0.25;
      ds = b * mrr * pt * 0.25;
      dt = c * mrr * ps * 0.25;
    } else {  // s = 0 edge, r = -+1
      const double a = idx == 1 ? -1.0 : 1.0, pr = 1.0 + a * r, mss = 1.0 - s * s;
      h = pr * mss * pt * 0.25;
      dr = a * mss * pt * 0.25;
      ds = -2.0 * s * pr * pt * 0.25;
      dt = c * pr * mss * 0.25;
    }
  } else {  // vertical edges under vertices 1..4
    const double a = c_rn[idx], b = c_sn[idx];
    const double pr = 1.0 + a * r, ps = 1.0 + b * s, mtt = 1.0 - t * t;
    h = pr * ps * mtt * 0.25;
    dr = a * ps * mtt * 0.25;
    ds = b * pr * mtt * 0.25;
    dt = -2.0 * t * pr * ps * 0.25;
  }
}

struct ElemMat {
  double a, b, c;           // D-matrix constants
  double coef_biot, imod_biot, dt_perm, dt_fluid_perm, pm_d, fluid_d;
};

__device__ __forceinline__ ElemMat elem_material(const double *e, double dt, double gmod) {
  ElemMat m;
  const double scale = 1.0e-06;
  m.pm_d = e[5] * scale;
  m.fluid_d = e[6] * scale;
  const double fluid_sw = m.fluid_d * gmod;
  const double ym = e[0], ps = e[1];
  const double perm = e[2] / fluid_sw;
  m.imod_biot = 1.0 / e[3];
  m.coef_biot = e[4];
  m.dt_perm = perm * dt;
  m.dt_fluid_perm = m.fluid_d * m.dt_perm;
  const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
  m.a = (ym * a1) / (a3 * a2);
  m.b = ps / a1;
  m.c = 0.5 * (a2 / a1);
  return m;
}

// Jacobian of the 8-vertex geometry at (r, s, t): returns det and the
// inverse (xji[a][j], the reference's xji(a, j))
__device__ __forceinline__ double jacobian(const double (*xv)[3], double r, double s, double t,
                                           double xji[3][3]) {
  double xj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int k = 0; k < 8; k++) {
    double dr, ds, dt;
    dn8(k, r, s, t, dr, ds, dt);
    for (int j = 0; j < 3; j++) {
      xj[0][j] += dr * xv[k][j];
      xj[1][j] += ds * xv[k][j];
      xj[2][j] += dt * xv[k][j];
    }
  }
  double d = xj[0][0] * xj[1][1] * xj[2][2] + xj[0][1] * xj[1][2] * xj[2][0] +
             xj[0][2] * xj[1][0] * xj[2][1] - xj[2][0] * xj[1][1] * xj[0][2] -
             xj[1][0] * xj[0][1] * xj[2][2] - xj[0][0] * xj[2][1] * xj[1][2];
  xji[0][0] = (xj[1][1] * xj[2][2] - xj[1][2] * xj[2][1]) / d;
  xji[1][0] = -(xj[1][0] * xj[2][2] - xj[1][2] * xj[2][0]) / d;
  xji[2][0] = (xj[1][0] * xj[2][1] - xj[1][1] * xj[2][0]) / d;
  xji[0][1] = -(xj[0][1] * xj[2][2] - xj[0][2] * xj[2][1]) / d;
  xji[1][1] = (xj[0][0] * xj[2][2] - xj[0][2] * xj[2][0]) / d;
  xji[2][1] = -(xj[0][0] * xj[2][1] - xj[0][1] * xj[2][0]) / d;
  xji[0][2] = (xj[0][1] * xj[1][2] - xj[0][2] * xj[1][1]) / d;
  xji[1][2] = -(xj[0][0] * xj[1][2] - xj[0][2] * xj[1][0]) / d;
  xji[2][2] = (xj[0][0] * xj[1][1] - xj[0][1] * xj[1][0]) / d;
  return d;
}

__device__ __forceinline__ void to_global(const double xji[3][3], double a, double b, double c,
                                          double &gx, double &gy, double &gz) {
  gx = xji[0][0] * a + xji[0][1] * b + xji[0][2] * c;
  gy = xji[1][0] * a + xji[1][1] * b + xji[1][2] * c;
  gz = xji[2][0] * a + xji[2][1] * b + xji[2][2] * c;
}

