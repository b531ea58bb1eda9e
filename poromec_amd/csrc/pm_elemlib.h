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

// value and natural derivatives of serendipity hexa20 node k (0..19): the
// reference's per-node expressions of sfhexa20 (Elmlib.f:1590-1807) term
// for term, so that with contraction off they round as the reference does.
// Nodes 8..11 are the mid-edges of the t = +1 face, 12..15 of t = -1, and
// 16..19 the vertical edges under vertices 0..3.
__device__ void dn20(int k, double r, double s, double t, double &h, double &dr, double &ds,
                     double &dt) {
  const double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + t;
  const double mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - t;
  const double mrr = 1.0 - r * r, mss = 1.0 - s * s, mtt = 1.0 - t * t;
  switch (k) {
    case 0:
      h = pr * ps * pt * (r + s + t - 2.0) * 0.125;
      dr = ps * pt * (2.0 * r + s + t - 1.0) * 0.125;
      ds = pr * pt * (r + 2.0 * s + t - 1.0) * 0.125;
      dt = pr * ps * (r + s + 2.0 * t - 1.0) * 0.125;
      break;
    case 1:
      h = mr * ps * pt * (-r + s + t - 2.0) * 0.125;
      dr = ps * pt * (2.0 * r - s - t + 1.0) * 0.125;
      ds = mr * pt * (-r + 2.0 * s + t - 1.0) * 0.125;
      dt = mr * ps * (-r + s + 2.0 * t - 1.0) * 0.125;
      break;
    case 2:
      h = mr * ms * pt * (-r - s + t - 2.0) * 0.125;
      dr = ms * pt * (2.0 * r + s - t + 1.0) * 0.125;
      ds = mr * pt * (r + 2.0 * s - t + 1.0) * 0.125;
      dt = mr * ms * (-r - s + 2.0 * t - 1.0) * 0.125;
      break;
    case 3:
      h = pr * ms * pt * (r - s + t - 2.0) * 0.125;
      dr = ms * pt * (2.0 * r - s + t - 1.0) * 0.125;
      ds = pr * pt * (-r + 2.0 * s - t + 1.0) * 0.125;
      dt = pr * ms * (r - s + 2.0 * t - 1.0) * 0.125;
      break;
    case 4:
      h = pr * ps * mt * (r + s - t - 2.0) * 0.125;
      dr = ps * mt * (2.0 * r + s - t - 1.0) * 0.125;
      ds = pr * mt * (r + 2.0 * s - t - 1.0) * 0.125;
      dt = pr * ps * (-r - s + 2.0 * t + 1.0) * 0.125;
      break;
    case 5:
      h = mr * ps * mt * (-r + s - t - 2.0) * 0.125;
      dr = ps * mt * (2.0 * r - s + t + 1.0) * 0.125;
      ds = mr * mt * (-r + 2.0 * s - t - 1.0) * 0.125;
      dt = mr * ps * (r - s + 2.0 * t + 1.0) * 0.125;
      break;
    case 6:
      h = mr * ms * mt * (-r - s - t - 2.0) * 0.125;
      dr = ms * mt * (2.0 * r + s + t + 1.0) * 0.125;
      ds = mr * mt * (r + 2.0 * s + t + 1.0) * 0.125;
      dt = mr * ms * (r + s + 2.0 * t + 1.0) * 0.125;
      break;
    case 7:
      h = pr * ms * mt * (r - s - t - 2.0) * 0.125;
      dr = ms * mt * (2.0 * r - s - t - 1.0) * 0.125;
      ds = pr * mt * (-r + 2.0 * s + t + 1.0) * 0.125;
      dt = pr * ms * (-r + s + 2.0 * t + 1.0) * 0.125;
      break;
    case 8:
      h = mrr * ps * pt * 0.25;
      dr = -r * ps * pt * 0.50;
      ds = mrr * pt * 0.25;
      dt = mrr * ps * 0.25;
      break;
    case 9:
      h = mr * mss * pt * 0.25;
      dr = -mss * pt * 0.25;
      ds = -s * mr * pt * 0.50;
      dt = mr * mss * 0.25;
      break;
    case 10:
      h = mrr * ms * pt * 0.25;
      dr = -r * ms * pt * 0.50;
      ds = -mrr * pt * 0.25;
      dt = mrr * ms * 0.25;
      break;
    case 11:
      h = pr * mss * pt * 0.25;
      dr = mss * pt * 0.25;
      ds = -s * pr * pt * 0.50;
      dt = pr * mss * 0.25;
      break;
    case 12:
      h = mrr * ps * mt * 0.25;
      dr = -r * ps * mt * 0.50;
      ds = mrr * mt * 0.25;
      dt = -mrr * ps * 0.25;
      break;
    case 13:
      h = mr * mss * mt * 0.25;
      dr = -mss * mt * 0.25;
      ds = -s * mr * mt * 0.50;
      dt = -mr * mss * 0.25;
      break;
    case 14:
      h = mrr * ms * mt * 0.25;
      dr = -r * ms * mt * 0.50;
      ds = -mrr * mt * 0.25;
      dt = -mrr * ms * 0.25;
      break;
    case 15:
      h = pr * mss * mt * 0.25;
      dr = mss * mt * 0.25;
      ds = -s * pr * mt * 0.50;
      dt = -pr * mss * 0.25;
      break;
    case 16:
      h = pr * ps * mtt * 0.25;
      dr = ps * mtt * 0.25;
      ds = pr * mtt * 0.25;
      dt = -t * pr * ps * 0.50;
      break;
    case 17:
      h = mr * ps * mtt * 0.25;
      dr = -ps * mtt * 0.25;
      ds = mr * mtt * 0.25;
      dt = -t * mr * ps * 0.50;
      break;
    case 18:
      h = mr * ms * mtt * 0.25;
      dr = -ms * mtt * 0.25;
      ds = -mr * mtt * 0.25;
      dt = -t * mr * ms * 0.50;
      break;
    case 19:
      h = pr * ms * mtt * 0.25;
      dr = ms * mtt * 0.25;
      ds = -pr * mtt * 0.25;
      dt = -t * pr * ms * 0.50;
      break;
    default:
      h = dr = ds = dt = 0.0;
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


// ---------------------------------------------------------------------------
// variable properties (elmt18_pm, Elmlib03.f:2640-3640; Poromec.f)
// ---------------------------------------------------------------------------
// prop_porosity (Poromec.f:165-200): one integration point's D-matrix
// constants, Biot moduli and permeability from its properties
// vp[0..6] = porosity, permeability, density, K, mu, 1/M, alpha; div43 and
// div23 are the reference's 15-digit constants
struct GpProp {
  double a, b, c, ibiot, cbiot, perm;
};
__device__ __forceinline__ GpProp prop_porosity(const double *vp, double fluid_sw) {
  const double div43 = 0.133333333333333e+01, div23 = 0.666666666666667;
  GpProp g;
  const double modK = vp[3], mu = vp[4];
  g.ibiot = vp[5];
  g.cbiot = vp[6];
  g.a = modK + div43 * mu;
  g.b = (modK - div23 * mu) / g.a;
  g.c = mu / g.a;
  g.perm = vp[1] / fluid_sw;
  return g;
}

// konzey_carman (Poromec.f:18-30)
__device__ __forceinline__ double konzey_carman(double perm0, double poro, double poro0) {
  const double tmp1 = (poro * poro * poro) / ((1.0 - poro) * (1.0 - poro));
  const double tmp2 = ((1.0 - poro0) * (1.0 - poro0)) / (poro0 * poro0 * poro0);
  return perm0 * tmp1 * tmp2;
}

// hashin_shtrikman (Poromec.f:113-163): drained K, mu, 1/M, alpha of a solid
// of moduli ks, mus at porosity poro
__device__ __forceinline__ void hashin_shtrikman(double poro, double ks, double mus, double &k,
                                                 double &mu, double &im_biot, double &c_biot) {
  const double ha = 1.0 - poro;
  const double hb = 9.0 + 6.0 * poro;
  const double hc = 8.0 + 12.0 * poro;
  k = (4.0 * ks * mus * ha) / (3.0 * ks * poro + 4.0 * mus);
  mu = (mus * ha * (9.0 * ks + 8 * mus)) / (ks * hb + mus * hc);
  c_biot = 1.0 - k / ks;
  im_biot = (c_biot - poro) / ks;
}

// extrapol_hexa20_v2 (Elmlib.f:3540-3590): the trilinear fit through the
// values f[0..7] at the corner integration points 1, 4, 13, 16, 49, 52, 61,
// 64, evaluated at vertex v
__device__ __forceinline__ double extrapol_vertex(const double *f, int v) {
  const double c1 = 0.125000000000000, c2 = 0.145157042290566;
  const double c3 = 0.168564535400043, c4 = 0.195746635145486;
  const double a1 = c1 * (f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7]);
  const double a2 = c2 * (f[1] + f[3] + f[5] + f[7] - (f[0] + f[2] + f[4] + f[6]));
  const double a3 = c2 * (f[2] + f[3] + f[6] + f[7] - (f[0] + f[1] + f[4] + f[5]));
  const double a4 = c2 * (f[4] + f[5] + f[6] + f[7] - (f[0] + f[1] + f[2] + f[3]));
  const double a5 = c3 * (f[0] + f[3] + f[4] + f[7] - (f[1] + f[2] + f[5] + f[6]));
  const double a6 = c3 * (f[0] + f[1] + f[6] + f[7] - (f[2] + f[3] + f[4] + f[5]));
  const double a7 = c3 * (f[0] + f[2] + f[5] + f[7] - (f[1] + f[3] + f[4] + f[6]));
  const double a8 = c4 * (f[1] + f[2] + f[4] + f[7] - (f[0] + f[3] + f[5] + f[6]));
  const double rnj = c_rn[v], snj = c_sn[v], tnj = c_tn[v];
  return a1 + a2 * rnj + a3 * snj + a4 * tnj + a5 * rnj * snj + a6 * snj * tnj +
         a7 * rnj * tnj + a8 * rnj * snj * tnj;
}
// the corner integration points (0-based) of the 4x4x4 rule, lx fastest
static __constant__ int c_gp_corner[8] = {0, 3, 12, 15, 48, 51, 60, 63};
