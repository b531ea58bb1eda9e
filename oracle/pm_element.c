/*
 * pm_element.c -- oracle restatement of the hexa20/8 Biot element elmt17_pm.
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 * Every routine keeps the reference's loop order and expression grouping so
 * that, built with -ffp-contract=off, it reproduces the reference's own
 * floating-point results.  Sources followed:
 *   sfhexa8_m    Elmlib.f:1510-1617     sfhexa20_m   Elmlib.f:1619-1807
 *   jacob3d_m    Elmlib.f:2301-2396     jacob2d_m    Elmlib.f:2128-2211
 *   sfquad4_m    Elmlib.f:1174-1230     sfquad8_m    Elmlib.f:1318-1420
 *   deform3d     Elmlib.f:2774-2815     stress3d     Elmlib.f:2906-2936
 *   rotmatrix    Elmlib.f:2397-2450     rotate       Elmlib.f:2518-2545
 *   face_normal_vector Elmlib.f:3678-3727, vet Matvec.f
 *   Gauss table  Elmlib.f:3837-3903 (15-digit constants)
 *   elmt17_pm    Elmlib03.f:1822-2636 (isw 2, 3, 4, 7), darcy_flux Elmlib.f:3178
 *   tload        Pload.f:220-377,  interpol Interpolation.f:1-40
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "pm_oracle.h"

/* Gauss-Legendre points and weights exactly as tabulated in Elmlib.f:3837-3903 */
static const double PG4[4] = {-.861136311594053, -.339981043584856, .339981043584856,
                              .861136311594053};
static const double WG4[4] = {.347854845137454, .652145154862546, .652145154862546,
                              .347854845137454};
static const double PG3[3] = {-.774596669241483, 0.0, .774596669241483};
static const double WG3[3] = {.555555555555556, .888888888888889, .555555555555556};

/* hexa20 face -> local node table (Elmlib03.f data hexa_face_node20), 1-based */
static const int FACE20[6][8] = {{1, 2, 3, 4, 9, 10, 11, 12},   {5, 6, 7, 8, 13, 14, 15, 16},
                                 {1, 5, 6, 2, 17, 13, 18, 9},   {4, 3, 7, 8, 11, 19, 15, 20},
                                 {1, 4, 8, 5, 12, 20, 16, 17},  {2, 6, 7, 3, 18, 14, 19, 10}};
/* natural coordinates of the 8 vertices (Elmlib03.f data rn/sn/tn) */
static const double RN[8] = {1, -1, -1, 1, 1, -1, -1, 1};
static const double SN[8] = {1, 1, -1, -1, 1, 1, -1, -1};
static const double TN[8] = {1, 1, 1, 1, -1, -1, -1, -1};

static void sfhexa8(double *h, double *hx, double *hy, double *hz, double r, double s,
                    double t, int afl, int bfl) {
  double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + t;
  double mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - t;
  if (afl) {
    h[0] = pr * ps * pt * 0.125; h[1] = mr * ps * pt * 0.125;
    h[2] = mr * ms * pt * 0.125; h[3] = pr * ms * pt * 0.125;
    h[4] = pr * ps * mt * 0.125; h[5] = mr * ps * mt * 0.125;
    h[6] = mr * ms * mt * 0.125; h[7] = pr * ms * mt * 0.125;
  }
  if (bfl) {
    hx[0] = ps * pt * 0.125;  hx[1] = -ps * pt * 0.125;
    hx[2] = -ms * pt * 0.125; hx[3] = ms * pt * 0.125;
    hx[4] = ps * mt * 0.125;  hx[5] = -ps * mt * 0.125;
    hx[6] = -ms * mt * 0.125; hx[7] = ms * mt * 0.125;
    hy[0] = pr * pt * 0.125;  hy[1] = mr * pt * 0.125;
    hy[2] = -mr * pt * 0.125; hy[3] = -pr * pt * 0.125;
    hy[4] = pr * mt * 0.125;  hy[5] = mr * mt * 0.125;
    hy[6] = -mr * mt * 0.125; hy[7] = -pr * mt * 0.125;
    hz[0] = pr * ps * 0.125;  hz[1] = mr * ps * 0.125;
    hz[2] = mr * ms * 0.125;  hz[3] = pr * ms * 0.125;
    hz[4] = -pr * ps * 0.125; hz[5] = -mr * ps * 0.125;
    hz[6] = -mr * ms * 0.125; hz[7] = -pr * ms * 0.125;
  }
}

static void sfhexa20(double *h, double *hx, double *hy, double *hz, double r, double s,
                     double t, int afl, int bfl) {
  double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + t;
  double mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - t;
  double mrr = 1.0 - r * r, mss = 1.0 - s * s, mtt = 1.0 - t * t;
  if (afl) {
    h[0] = pr * ps * pt * (r + s + t - 2.0) * 0.125;
    h[1] = mr * ps * pt * (-r + s + t - 2.0) * 0.125;
    h[2] = mr * ms * pt * (-r - s + t - 2.0) * 0.125;
    h[3] = pr * ms * pt * (r - s + t - 2.0) * 0.125;
    h[4] = pr * ps * mt * (r + s - t - 2.0) * 0.125;
    h[5] = mr * ps * mt * (-r + s - t - 2.0) * 0.125;
    h[6] = mr * ms * mt * (-r - s - t - 2.0) * 0.125;
    h[7] = pr * ms * mt * (r - s - t - 2.0) * 0.125;
    h[8] = mrr * ps * pt * 0.25;  h[9] = mr * mss * pt * 0.25;
    h[10] = mrr * ms * pt * 0.25; h[11] = pr * mss * pt * 0.25;
    h[12] = mrr * ps * mt * 0.25; h[13] = mr * mss * mt * 0.25;
    h[14] = mrr * ms * mt * 0.25; h[15] = pr * mss * mt * 0.25;
    h[16] = pr * ps * mtt * 0.25; h[17] = mr * ps * mtt * 0.25;
    h[18] = mr * ms * mtt * 0.25; h[19] = pr * ms * mtt * 0.25;
  }
  if (bfl) {
    hx[0] = ps * pt * (2.0 * r + s + t - 1.0) * 0.125;
    hx[1] = ps * pt * (2.0 * r - s - t + 1.0) * 0.125;
    hx[2] = ms * pt * (2.0 * r + s - t + 1.0) * 0.125;
    hx[3] = ms * pt * (2.0 * r - s + t - 1.0) * 0.125;
    hx[4] = ps * mt * (2.0 * r + s - t - 1.0) * 0.125;
    hx[5] = ps * mt * (2.0 * r - s + t + 1.0) * 0.125;
    hx[6] = ms * mt * (2.0 * r + s + t + 1.0) * 0.125;
    hx[7] = ms * mt * (2.0 * r - s - t - 1.0) * 0.125;
    hx[8] = -r * ps * pt * 0.50;  hx[9] = -mss * pt * 0.25;
    hx[10] = -r * ms * pt * 0.50; hx[11] = mss * pt * 0.25;
    hx[12] = -r * ps * mt * 0.50; hx[13] = -mss * mt * 0.25;
    hx[14] = -r * ms * mt * 0.50; hx[15] = mss * mt * 0.25;
    hx[16] = ps * mtt * 0.25;     hx[17] = -ps * mtt * 0.25;
    hx[18] = -ms * mtt * 0.25;    hx[19] = ms * mtt * 0.25;

    hy[0] = pr * pt * (r + 2.0 * s + t - 1.0) * 0.125;
    hy[1] = mr * pt * (-r + 2.0 * s + t - 1.0) * 0.125;
    hy[2] = mr * pt * (r + 2.0 * s - t + 1.0) * 0.125;
    hy[3] = pr * pt * (-r + 2.0 * s - t + 1.0) * 0.125;
    hy[4] = pr * mt * (r + 2.0 * s - t - 1.0) * 0.125;
    hy[5] = mr * mt * (-r + 2.0 * s - t - 1.0) * 0.125;
    hy[6] = mr * mt * (r + 2.0 * s + t + 1.0) * 0.125;
    hy[7] = pr * mt * (-r + 2.0 * s + t + 1.0) * 0.125;
    hy[8] = mrr * pt * 0.25;      hy[9] = -s * mr * pt * 0.50;
    hy[10] = -mrr * pt * 0.25;    hy[11] = -s * pr * pt * 0.50;
    hy[12] = mrr * mt * 0.25;     hy[13] = -s * mr * mt * 0.50;
    hy[14] = -mrr * mt * 0.25;    hy[15] = -s * pr * mt * 0.50;
    hy[16] = pr * mtt * 0.25;     hy[17] = mr * mtt * 0.25;
    hy[18] = -mr * mtt * 0.25;    hy[19] = -pr * mtt * 0.25;

    hz[0] = pr * ps * (r + s + 2.0 * t - 1.0) * 0.125;
    hz[1] = mr * ps * (-r + s + 2.0 * t - 1.0) * 0.125;
    hz[2] = mr * ms * (-r - s + 2.0 * t - 1.0) * 0.125;
    hz[3] = pr * ms * (r - s + 2.0 * t - 1.0) * 0.125;
    hz[4] = pr * ps * (-r - s + 2.0 * t + 1.0) * 0.125;
    hz[5] = mr * ps * (r - s + 2.0 * t + 1.0) * 0.125;
    hz[6] = mr * ms * (r + s + 2.0 * t + 1.0) * 0.125;
    hz[7] = pr * ms * (-r + s + 2.0 * t + 1.0) * 0.125;
    hz[8] = mrr * ps * 0.25;      hz[9] = mr * mss * 0.25;
    hz[10] = mrr * ms * 0.25;     hz[11] = pr * mss * 0.25;
    hz[12] = -mrr * ps * 0.25;    hz[13] = -mr * mss * 0.25;
    hz[14] = -mrr * ms * 0.25;    hz[15] = -pr * mss * 0.25;
    hz[16] = -t * pr * ps * 0.50; hz[17] = -t * mr * ps * 0.50;
    hz[18] = -t * mr * ms * 0.50; hz[19] = -t * pr * ms * 0.50;
  }
}

/* jacob3d_m: xj(a,j) = sum_k d_a N_k x(j,k); stored xj[a][j]; xji likewise */
static int jacob3d(double *hx, double *hy, double *hz, double xj[3][3], double xji[3][3],
                   const double *x, double *det, int nen, int afl) {
  if (afl) {
    for (int j = 0; j < 3; j++) {
      xj[0][j] = 0.0; xj[1][j] = 0.0; xj[2][j] = 0.0;
      for (int k = 0; k < nen; k++) {
        xj[0][j] = xj[0][j] + hx[k] * x[3 * k + j];
        xj[1][j] = xj[1][j] + hy[k] * x[3 * k + j];
        xj[2][j] = xj[2][j] + hz[k] * x[3 * k + j];
      }
    }
    double d = xj[0][0] * xj[1][1] * xj[2][2] + xj[0][1] * xj[1][2] * xj[2][0] +
               xj[0][2] * xj[1][0] * xj[2][1] - xj[2][0] * xj[1][1] * xj[0][2] -
               xj[1][0] * xj[0][1] * xj[2][2] - xj[0][0] * xj[2][1] * xj[1][2];
    *det = d;
    if (d <= 1.0e-14) return -1;
    xji[0][0] = (xj[1][1] * xj[2][2] - xj[1][2] * xj[2][1]) / d;
    xji[1][0] = -(xj[1][0] * xj[2][2] - xj[1][2] * xj[2][0]) / d;
    xji[2][0] = (xj[1][0] * xj[2][1] - xj[1][1] * xj[2][0]) / d;
    xji[0][1] = -(xj[0][1] * xj[2][2] - xj[0][2] * xj[2][1]) / d;
    xji[1][1] = (xj[0][0] * xj[2][2] - xj[0][2] * xj[2][0]) / d;
    xji[2][1] = -(xj[0][0] * xj[2][1] - xj[0][1] * xj[2][0]) / d;
    xji[0][2] = (xj[0][1] * xj[1][2] - xj[0][2] * xj[1][1]) / d;
    xji[1][2] = -(xj[0][0] * xj[1][2] - xj[0][2] * xj[1][0]) / d;
    xji[2][2] = (xj[0][0] * xj[1][1] - xj[0][1] * xj[1][0]) / d;
  }
  for (int k = 0; k < nen; k++) {
    double a = hx[k], b = hy[k], c = hz[k];
    hx[k] = xji[0][0] * a + xji[0][1] * b + xji[0][2] * c;
    hy[k] = xji[1][0] * a + xji[1][1] * b + xji[1][2] * c;
    hz[k] = xji[2][0] * a + xji[2][1] * b + xji[2][2] * c;
  }
  return 0;
}

static int jacob2d(double *hx, double *hy, double xj[2][2], double xji[2][2], const double *x,
                   double *det, int nen) {
  for (int j = 0; j < 2; j++) {
    xj[0][j] = 0.0; xj[1][j] = 0.0;
    for (int k = 0; k < nen; k++) {
      xj[0][j] = xj[0][j] + hx[k] * x[3 * k + j];
      xj[1][j] = xj[1][j] + hy[k] * x[3 * k + j];
    }
  }
  double d = xj[0][0] * xj[1][1] - xj[1][0] * xj[0][1];
  *det = d;
  if (d <= 1.0e-14) return -1;
  xji[0][0] = xj[1][1] / d; xji[0][1] = -xj[0][1] / d;
  xji[1][0] = -xj[1][0] / d; xji[1][1] = xj[0][0] / d;
  for (int k = 0; k < nen; k++) {
    double a = hx[k], b = hy[k];
    hx[k] = xji[0][0] * a + xji[0][1] * b;
    hy[k] = xji[1][0] * a + xji[1][1] * b;
  }
  return 0;
}

static void sfquad4(double *h, double *hx, double *hy, double r, double s, int afl, int bfl) {
  double pr = 1.0 + r, ps = 1.0 + s, mr = 1.0 - r, ms = 1.0 - s;
  if (afl) {
    h[0] = pr * ps * 0.25; h[1] = mr * ps * 0.25; h[2] = mr * ms * 0.25; h[3] = pr * ms * 0.25;
  }
  if (bfl) {
    hx[0] = ps * 0.25; hx[1] = -ps * 0.25; hx[2] = -ms * 0.25; hx[3] = ms * 0.25;
    hy[0] = pr * 0.25; hy[1] = mr * 0.25; hy[2] = -mr * 0.25; hy[3] = -pr * 0.25;
  }
}

/* sfquad8_m, shape values only (the element calls it with afl=T, bfl=F) */
static void sfquad8_h(double *h, double r, double s) {
  double pr = 1.0 + r, ps = 1.0 + s, mr = 1.0 - r, ms = 1.0 - s;
  double mrr = 1.0 - r * r, mss = 1.0 - s * s;
  h[4] = mrr * ps * 0.5; h[5] = mss * mr * 0.5; h[6] = mrr * ms * 0.5; h[7] = mss * pr * 0.5;
  double a5 = h[4] * 0.5, a6 = h[5] * 0.5, a7 = h[6] * 0.5, a8 = h[7] * 0.5;
  h[0] = pr * ps * 0.25 - a5 - a8;
  h[1] = mr * ps * 0.25 - a5 - a6;
  h[2] = mr * ms * 0.25 - a6 - a7;
  h[3] = pr * ms * 0.25 - a7 - a8;
}

static void deform3d(const double *hx, const double *hy, const double *hz, const double *u,
                     double *eps, int nen) {
  double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
  for (int i = 0; i < nen; i++) {
    ux = ux + hx[i] * u[3 * i];     uy = uy + hy[i] * u[3 * i];     uz = uz + hz[i] * u[3 * i];
    vx = vx + hx[i] * u[3 * i + 1]; vy = vy + hy[i] * u[3 * i + 1]; vz = vz + hz[i] * u[3 * i + 1];
    wx = wx + hx[i] * u[3 * i + 2]; wy = wy + hy[i] * u[3 * i + 2]; wz = wz + hz[i] * u[3 * i + 2];
  }
  eps[0] = ux; eps[1] = vy; eps[2] = wz;
  eps[3] = uy + vx; eps[4] = vz + wy; eps[5] = uz + wx;
}

static void stress3d(double a, double b, double c, const double *e, double *t) {
  t[0] = (e[0] + b * e[1] + b * e[2]) * a;
  t[1] = (b * e[0] + e[1] + b * e[2]) * a;
  t[2] = (b * e[0] + b * e[1] + e[2]) * a;
  t[3] = e[3] * c * a;
  t[4] = e[4] * c * a;
  t[5] = e[5] * c * a;
}

static void rotmatrix(const double *x, double r[3][3]) {
  double v1x = x[3] - x[0], v1y = x[4] - x[1], v1z = x[5] - x[2];
  double v2x = x[6] - x[0], v2y = x[7] - x[1], v2z = x[8] - x[2];
  double v3x = v1y * v2z - v1z * v2y, v3y = v1z * v2x - v1x * v2z, v3z = v1x * v2y - v1y * v2x;
  v2x = v1z * v3y - v1y * v3z; v2y = v1x * v3z - v1z * v3x; v2z = v1y * v3x - v1x * v3y;
  double v1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z);
  double v2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z);
  double v3 = sqrt(v3x * v3x + v3y * v3y + v3z * v3z);
  r[0][0] = v1x / v1; r[0][1] = v1y / v1; r[0][2] = v1z / v1;
  r[1][0] = v2x / v2; r[1][1] = v2y / v2; r[1][2] = v2z / v2;
  r[2][0] = v3x / v3; r[2][1] = v3y / v3; r[2][2] = v3z / v3;
}

static void rotate(double *x, double r[3][3]) {
  double x1 = r[0][0] * x[0] + r[0][1] * x[1] + r[0][2] * x[2];
  double x2 = r[1][0] * x[0] + r[1][1] * x[1] + r[1][2] * x[2];
  double x3 = r[2][0] * x[0] + r[2][1] * x[1] + r[2][2] * x[2];
  x[0] = x1; x[1] = x2; x[2] = x3;
}

static void face_normal(const double *xf, double *nrm) {
  double v1[3] = {xf[3] - xf[0], xf[4] - xf[1], xf[5] - xf[2]};
  double v2[3] = {xf[6] - xf[0], xf[7] - xf[1], xf[8] - xf[2]};
  /* vet(v2, v1, normal): normal = v2 x v1 */
  nrm[0] = v2[1] * v1[2] - v2[2] * v1[1];
  nrm[1] = v2[2] * v1[0] - v2[0] * v1[2];
  nrm[2] = v2[0] * v1[1] - v2[1] * v1[0];
  double tmp = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  tmp = 1.0 / tmp;
  nrm[0] *= tmp; nrm[1] *= tmp; nrm[2] *= tmp;
}

/* interpol (Interpolation.f:1): piecewise-linear table lookup, clamped */
static double interpol(const double *xs, const double *ys, double xi, int n) {
  if (xi <= xs[0]) return ys[0];
  if (xi >= xs[n - 1]) return ys[n - 1];
  double c = 0.0;
  for (int i = 0; i < n - 1; i++) {
    double x1 = xs[i], x2 = xs[i + 1], y1 = ys[i], y2 = ys[i + 1];
    if (xi >= x1 && xi <= x2) { c = y1 + (y2 - y1) * (xi - x1) / (x2 - x1); break; }
  }
  return c;
}

#define LOAD(k, nc) (env->load[3 * ((nc) - 1) + (k) - 1])
#define FLOAD(i, term, nc) \
  (env->fload + (size_t)((nc) - 1) * ORC_NTERMS * ORC_NPAR + (size_t)((term) - 1) * ORC_NPAR + (i) - 1)

/* tload (Pload.f:220-377).  Returns c; vc receives vector loads (type 40). */
double orc_tload(int nc, double t, const double *x, double u, double *vc, const orc_env *env) {
  const double DOISPI = 6.283185307;
  double c = 0.0;
  if (nc <= 0) return c;
  int itype = LOAD(1, nc), nparc = LOAD(2, nc);
  if (itype == 1) {
    c = *FLOAD(1, 1, nc);
  } else if (itype == 2) {
    c = *FLOAD(1, 1, nc) + *FLOAD(2, 1, nc) * t;
  } else if (itype == 3) {
    c = (*FLOAD(1, 1, nc) - u) * *FLOAD(2, 1, nc);
  } else if (itype == 4 || itype == 5) {
    for (int i = 1; i <= nparc; i++) {
      double a = *FLOAD(1, i, nc), w = 1.0 / *FLOAD(2, i, nc), t0 = *FLOAD(3, i, nc);
      double ti = fmax(t - t0, 0.0);
      c = c + a * (itype == 4 ? sin(DOISPI * w * ti) : cos(DOISPI * w * ti));
    }
  } else if (itype == 6) {
    for (int i = 1; i <= nparc; i++) {
      double a = *FLOAD(1, i, nc), w = 1.0 / *FLOAD(2, i, nc), t0 = *FLOAD(3, i, nc);
      c = c + a * cos(DOISPI * w * t + t0);
    }
  } else if (itype == 7) {
    c = interpol(FLOAD(1, 2, nc), FLOAD(1, 3, nc), t, nparc);
  } else if (itype == 8) {
    c = interpol(FLOAD(1, 1, nc), FLOAD(1, 2, nc), t, nparc);
  } else if (itype == 39 || itype == 41 || itype == 43) {
    nparc = LOAD(3, nc);
    c = interpol(FLOAD(1, 1, nc), FLOAD(1, 2, nc), t, nparc);
  } else if (itype == 40) {
    nparc = LOAD(3, nc);
    vc[0] = interpol(FLOAD(1, 1, nc), FLOAD(1, 2, nc), t, nparc);
    vc[1] = interpol(FLOAD(1, 1, nc), FLOAD(1, 3, nc), t, nparc);
    vc[2] = interpol(FLOAD(1, 1, nc), FLOAD(1, 4, nc), t, nparc);
  } else if (itype == 42) {
    nparc = LOAD(3, nc);
    double a = interpol(FLOAD(1, 1, nc), FLOAD(1, 2, nc), t, nparc);
    double b = interpol(FLOAD(1, 1, nc), FLOAD(1, 3, nc), t, nparc);
    double w = interpol(FLOAD(1, 1, nc), FLOAD(1, 4, nc), t, nparc);
    double t0 = interpol(FLOAD(1, 1, nc), FLOAD(1, 5, nc), t, nparc);
    double h = (w - x[2]);
    c = t0 * (b + a * h * env->gravity_mod);
  }
  return c;
}

#define S(i, j) s[(size_t)((j) - 1) * ORC_NST + (i) - 1]
#define P(i) p[(i) - 1]
#define U(i) u[(i) - 1]

/* elmt17_pm (Elmlib03.f:1822-2636), isw = 2, 3, 4, 7.
 * Returns 0, -1 on a non-positive Jacobian (jacob3d_m stops there), -2 on an
 * unsupported face-load type.  p must hold >= 68 values (120 for isw 3), s
 * 68*68. */
/* prop_porosity (Poromec.f:165-200): D-matrix constants, Biot moduli and
 * permeability of one integration point from its variable properties
 * vp(1..7) = porosity, permeability, density, K, mu, 1/M, alpha.  div43 and
 * div23 are the reference's 15-digit constants. */
static void prop_porosity(double *a, double *b, double *c, double *ibiot, double *cbiot,
                          double *perm, const double *vp, double fluid_sw) {
  const double div43 = 0.133333333333333e+01, div23 = 0.666666666666667;
  double modK = vp[3], mu = vp[4];
  *ibiot = vp[5];
  *cbiot = vp[6];
  *a = modK + div43 * mu;
  *b = (modK - div23 * mu) / *a;
  *c = mu / *a;
  *perm = vp[1] / fluid_sw;
}

/* The hexa20/8 Biot element, isw 2, 3, 4, 7: elmt17_pm when vp is NULL;
 * with vp the per-integration-point properties (7 x 64) of elmt18_pm, whose
 * isw 2 and 4 differ from elmt17_pm only there (Elmlib03.f:2872-3067 against
 * 2022-2205, 3167-3293 against 2263-2372): prop_porosity at each point, Kpp
 * weighted by w dt perm, the isw 4 flux by w dt rho_f perm and w dt perm,
 * and in isw 4 the point index advancing once per ly (Elmlib03.f:3171-3173),
 * so the 64 points read the properties of points 1..16, 4 each. */
static int elmt_hexa20(const double *e, const int *iqm, const int *iqh, const double *x,
                       const double *u, const double *dp, double *p, double *s, int isw,
                       const orc_env *env, const double *vp) {
  double hu[20], hux[20], huy[20], huz[20], hp[8], hpx[8], hpy[8], hpz[8];
  double xj[3][3], xji[3][3], det, w, wt1, wt2, wt3;
  double epsi[6], txi[6];
  const double scale = 1.0e-06;
  double gl[3] = {env->gravity[0], env->gravity[1], env->gravity[2]};
  double pm_d = e[5] * scale, fluid_d = e[6] * scale;
  double fluid_sw = fluid_d * env->gravity_mod;
  double ym = e[0], ps = e[1];
  double perm = e[2] / fluid_sw;
  double imod_biot = 1.0 / e[3], coef_biot = e[4];
  double dt_perm = perm * env->dt;
  double dt_fluid_perm = fluid_d * dt_perm;
  double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
  double a = (ym * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
  const int nst = ORC_NST;

  if (isw == 2) {
    for (int i = 1; i <= nst; i++) {
      for (int j = 1; j <= nst; j++) S(j, i) = 0.0;
      P(i) = 0.0;
    }
    int inpi = 0;
    for (int lz = 0; lz < 4; lz++) {
      double ti = PG4[lz];
      for (int ly = 0; ly < 4; ly++) {
        double si = PG4[ly];
        for (int lx = 0; lx < 4; lx++) {
          inpi++;
          double ri = PG4[lx];
          if (vp)
            prop_porosity(&a, &b, &c, &imod_biot, &coef_biot, &perm, vp + 7 * (inpi - 1),
                          fluid_sw);
          sfhexa8(hp, hpx, hpy, hpz, ri, si, ti, 1, 1);
          if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
          w = WG4[lx] * WG4[ly] * WG4[lz] * det;
          sfhexa20(hu, hux, huy, huz, ri, si, ti, 0, 1);
          jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
          wt1 = w * a;
          for (int i = 1; i <= 20; i++) {
            int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
            double xi_ = hux[i - 1], yi = huy[i - 1], zi = huz[i - 1];
            for (int j = 1; j <= 20; j++) {
              int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
              double xj_ = hux[j - 1], yj = huy[j - 1], zj = huz[j - 1];
              S(l1, k1) = S(l1, k1) + (xi_ * xj_ + c * (yi * yj + zi * zj)) * wt1;
              S(l1, k2) = S(l1, k2) + (b * xi_ * yj + c * yi * xj_) * wt1;
              S(l1, k3) = S(l1, k3) + (b * xi_ * zj + c * zi * xj_) * wt1;
              S(l2, k1) = S(l2, k1) + (b * yi * xj_ + c * xi_ * yj) * wt1;
              S(l2, k2) = S(l2, k2) + (yi * yj + c * (xi_ * xj_ + zi * zj)) * wt1;
              S(l2, k3) = S(l2, k3) + (b * yi * zj + c * zi * yj) * wt1;
              S(l3, k1) = S(l3, k1) + (b * zi * xj_ + c * xi_ * zj) * wt1;
              S(l3, k2) = S(l3, k2) + (b * zi * yj + c * yi * zj) * wt1;
              S(l3, k3) = S(l3, k3) + (zi * zj + c * (yi * yj + xi_ * xj_)) * wt1;
            }
          }
          wt2 = w * coef_biot;
          for (int i = 1; i <= 8; i++) {
            int l = i + 60;
            wt1 = hp[i - 1] * wt2;
            for (int j = 1; j <= 20; j++) {
              int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
              S(l, k1) = S(l, k1) - hux[j - 1] * wt1;
              S(l, k2) = S(l, k2) - huy[j - 1] * wt1;
              S(l, k3) = S(l, k3) - huz[j - 1] * wt1;
            }
          }
          wt1 = w * imod_biot;
          wt2 = vp ? w * env->dt * perm : w * dt_perm;
          for (int i = 1; i <= 8; i++) {
            int l = i + 60;
            for (int j = 1; j <= 8; j++) {
              int k = j + 60;
              S(l, k) = S(l, k) - wt1 * hp[i - 1] * hp[j - 1] -
                        wt2 * (hpx[i - 1] * hpx[j - 1] + hpy[i - 1] * hpy[j - 1] +
                               hpz[i - 1] * hpz[j - 1]);
            }
          }
        }
      }
    }
    for (int i = 1; i <= 8; i++) {
      int l = i + 60;
      for (int j = 1; j <= 20; j++) {
        int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
        S(k1, l) = S(l, k1); S(k2, l) = S(l, k2); S(k3, l) = S(l, k3);
      }
    }
    /* lku_m (Elmlib.f:695-727): p = s.u, sweeping columns */
    for (int i = 1; i <= nst; i++) P(i) = 0.0;
    for (int j = 1; j <= nst; j++)
      for (int i = 1; i <= nst; i++) P(i) = P(i) + S(i, j) * U(j);
    return 0;
  }

  if (isw == 4) {
    int inpi = 0;
    for (int lz = 0; lz < 4; lz++) {
      double ti = PG4[lz];
      for (int ly = 0; ly < 4; ly++) {
        inpi++;
        double si = PG4[ly];
        for (int lx = 0; lx < 4; lx++) {
          double ri = PG4[lx];
          if (vp) {
            prop_porosity(&a, &b, &c, &imod_biot, &coef_biot, &perm, vp + 7 * (inpi - 1),
                          fluid_sw);
            pm_d = vp[7 * (inpi - 1) + 2];
          }
          sfhexa8(hp, hpx, hpy, hpz, ri, si, ti, 1, 1);
          if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
          w = WG4[lx] * WG4[ly] * WG4[lz] * det;
          sfhexa20(hu, hux, huy, huz, ri, si, ti, 1, 1);
          jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
          deform3d(hux, huy, huz, u, epsi, 20);
          stress3d(a, b, c, epsi, txi);
          double dpm = hp[0] * dp[0] + hp[1] * dp[1] + hp[2] * dp[2] + hp[3] * dp[3] +
                       hp[4] * dp[4] + hp[5] * dp[5] + hp[6] * dp[6] + hp[7] * dp[7];
          dpm = coef_biot * dpm;
          txi[0] = txi[0] - dpm; txi[1] = txi[1] - dpm; txi[2] = txi[2] - dpm;
          wt1 = w;
          wt2 = w * pm_d;
          for (int i = 1; i <= 20; i++) {
            int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
            P(l1) = P(l1) + wt1 * (hux[i - 1] * txi[0] + huy[i - 1] * txi[3] + huz[i - 1] * txi[5]);
            P(l2) = P(l2) + wt1 * (hux[i - 1] * txi[3] + huy[i - 1] * txi[1] + huz[i - 1] * txi[4]);
            P(l3) = P(l3) + wt1 * (hux[i - 1] * txi[5] + huy[i - 1] * txi[4] + huz[i - 1] * txi[2]);
            wt3 = hu[i - 1] * wt2;
            P(l1) = P(l1) - wt3 * gl[0];
            P(l2) = P(l2) - wt3 * gl[1];
            P(l3) = P(l3) - wt3 * gl[2];
          }
          wt1 = vp ? w * env->dt * fluid_d * perm : w * dt_fluid_perm;
          wt2 = vp ? w * env->dt * perm : w * dt_perm;
          for (int i = 1; i <= 8; i++) {
            int l = i + 60;
            P(l) = P(l) + wt1 * (hpx[i - 1] * gl[0] + hpy[i - 1] * gl[1] + hpz[i - 1] * gl[2]);
            double tmp = 0.0;
            for (int j = 1; j <= 8; j++) {
              int k = j + 60;
              tmp = tmp + wt2 * U(k) *
                              (hpx[i - 1] * hpx[j - 1] + hpy[i - 1] * hpy[j - 1] +
                               hpz[i - 1] * hpz[j - 1]);
            }
            P(l) = P(l) - tmp;
          }
        }
      }
    }
    /* face loads (Elmlib03.f:2373-2518) */
    int tpm = 0, tph = 0;
    for (int i = 0; i < 6; i++) { tpm += iqm[i]; tph += iqh[i]; }
    if (tpm > 0 || tph > 0) {
      /* cargm, cargh and face_f persist from face to face as in the
         reference, which sets them only when the face carries that load */
      int cargm = 0, cargh = 0;
      double face_f[3] = {0, 0, 0}, mp[3] = {0, 0, 0}, n_carg = 0.0;
      for (int j = 1; j <= 6; j++) {
        if (!(iqm[j - 1] > 0 || iqh[j - 1] > 0)) continue;
        double xf[12], r[3][3];
        for (int i = 0; i < 4; i++) {
          int no = FACE20[j - 1][i];
          xf[3 * i] = x[3 * (no - 1)];
          xf[3 * i + 1] = x[3 * (no - 1) + 1];
          xf[3 * i + 2] = x[3 * (no - 1) + 2];
        }
        if (iqm[j - 1] != 0) cargm = LOAD(1, iqm[j - 1]);
        if (iqh[j - 1] != 0) cargh = LOAD(1, iqh[j - 1]);
        if (cargm == 42) return -2; /* the reference reads an unset normal here */
        if (cargm == 40 || cargm == 41) {
          if (cargm == 41) {
            face_normal(xf, face_f);
            if (j != 2) { face_f[0] = -face_f[0]; face_f[1] = -face_f[1]; face_f[2] = -face_f[2]; }
          }
          rotmatrix(xf, r);
          for (int i = 0; i < 4; i++) rotate(xf + 3 * i, r);
          if (cargm == 41) {
            n_carg = orc_tload(iqm[j - 1], env->t, mp, 0.0, NULL, env);
            face_f[0] = n_carg * face_f[0]; face_f[1] = n_carg * face_f[1];
            face_f[2] = n_carg * face_f[2];
          } else {
            orc_tload(iqm[j - 1], env->t, mp, 0.0, face_f, env);
          }
          for (int ly = 0; ly < 3; ly++) {
            double si = PG3[ly];
            for (int lx = 0; lx < 3; lx++) {
              double ri = PG3[lx];
              double h4[4], h4x[4], h4y[4], h8[8], xj2[2][2], xji2[2][2];
              sfquad4(h4, h4x, h4y, ri, si, 0, 1);
              if (jacob2d(h4x, h4y, xj2, xji2, xf, &det, 4)) return -1;
              w = WG3[lx] * WG3[ly] * det;
              sfquad8_h(h8, ri, si);
              for (int i = 1; i <= 8; i++) {
                int no = FACE20[j - 1][i - 1];
                wt1 = h8[i - 1] * w;
                int l1 = 3 * no - 2, l2 = l1 + 1, l3 = l2 + 1;
                P(l1) = P(l1) - face_f[0] * wt1;
                P(l2) = P(l2) - face_f[1] * wt1;
                P(l3) = P(l3) - face_f[2] * wt1;
              }
            }
          }
        }
        if (cargh == 43) {
          /* rotates xf again, in place, as the reference does */
          rotmatrix(xf, r);
          for (int i = 0; i < 4; i++) rotate(xf + 3 * i, r);
          n_carg = orc_tload(iqh[j - 1], env->t, mp, 0.0, NULL, env);
          for (int ly = 0; ly < 3; ly++) {
            double si = PG3[ly];
            for (int lx = 0; lx < 3; lx++) {
              double ri = PG3[lx];
              double h4[4], h4x[4], h4y[4], xj2[2][2], xji2[2][2];
              sfquad4(h4, h4x, h4y, ri, si, 1, 1);
              if (jacob2d(h4x, h4y, xj2, xji2, xf, &det, 4)) return -1;
              w = WG3[lx] * WG3[ly] * det;
              w = w * env->dt * n_carg;
              for (int i = 1; i <= 4; i++) {
                int no = FACE20[j - 1][i - 1];
                wt1 = h4[i - 1] * w;
                P(60 + no) = P(60 + no) - wt1;
              }
            }
          }
        }
      }
    }
    return 0;
  }

  /* isw 3 (Elmlib03.f:2206-2256): at each vertex the total stress
   * D eps - alpha dp (p(1:48)), Biot's effective stress, total + alpha p
   * (p(49:96)), and the Darcy flux perm (rho_f g - grad p) (darcy_flux,
   * Elmlib.f:3178-3200; p(97:120)) */
  if (isw == 3) {
    for (int i = 1; i <= 8; i++) {
      int tp = 6 * i - 5, tp1 = 6 * i + 43, l = i + 60;
      sfhexa8(hp, hpx, hpy, hpz, RN[i - 1], SN[i - 1], TN[i - 1], 0, 1);
      if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
      sfhexa20(hu, hux, huy, huz, RN[i - 1], SN[i - 1], TN[i - 1], 0, 1);
      jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
      deform3d(hux, huy, huz, u, epsi, 20);
      stress3d(a, b, c, epsi, &P(tp));
      double dpm = coef_biot * dp[i - 1];
      double pm = coef_biot * U(l);
      P(tp) = P(tp) - dpm;
      P(tp + 1) = P(tp + 1) - dpm;
      P(tp + 2) = P(tp + 2) - dpm;
      P(tp1) = P(tp) + pm;
      P(tp1 + 1) = P(tp + 1) + pm;
      P(tp1 + 2) = P(tp + 2) + pm;
      P(tp1 + 3) = P(tp + 3);
      P(tp1 + 4) = P(tp + 4);
      P(tp1 + 5) = P(tp + 5);
    }
    for (int i = 1; i <= 8; i++) {
      int tp = 3 * i + 94;
      sfhexa8(hp, hpx, hpy, hpz, RN[i - 1], SN[i - 1], TN[i - 1], 1, 1);
      if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
      double *q = &P(tp);
      q[0] = fluid_d * gl[0];
      q[1] = fluid_d * gl[1];
      q[2] = fluid_d * gl[2];
      for (int k = 0; k < 8; k++) {
        q[0] = q[0] - hpx[k] * U(61 + k);
        q[1] = q[1] - hpy[k] * U(61 + k);
        q[2] = q[2] - hpz[k] * U(61 + k);
      }
      q[0] = perm * q[0];
      q[1] = perm * q[1];
      q[2] = perm * q[2];
    }
    return 0;
  }

  if (isw == 7) {
    for (int i = 1; i <= 8; i++) {
      sfhexa8(hp, hpx, hpy, hpz, RN[i - 1], SN[i - 1], TN[i - 1], 0, 1);
      if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
      sfhexa20(hu, hux, huy, huz, RN[i - 1], SN[i - 1], TN[i - 1], 0, 1);
      jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
      deform3d(hux, huy, huz, u, epsi, 20);
      double pi = imod_biot * U(i + 60);
      P(i) = coef_biot * (epsi[0] + epsi[1] + epsi[2]) + pi;
    }
    return 0;
  }
  return -3;
}

int orc_elmt17(const double *e, const int *iqm, const int *iqh, const double *x,
               const double *u, const double *dp, double *p, double *s, int isw,
               const orc_env *env) {
  return elmt_hexa20(e, iqm, iqh, x, u, dp, p, s, isw, env, NULL);
}

/* ---- variable properties (Poromec.f) ---- */

/* konzey_carman (Poromec.f:18-30) */
double orc_konzey_carman(double perm0, double poro, double poro0) {
  double tmp1 = (poro * poro * poro) / ((1.0 - poro) * (1.0 - poro));
  double tmp2 = ((1.0 - poro0) * (1.0 - poro0)) / (poro0 * poro0 * poro0);
  return perm0 * tmp1 * tmp2;
}

/* hashin_shtrikman (Poromec.f:113-163): drained K, mu, 1/M and alpha of a
 * porous solid of moduli ks, mus at porosity poro */
void orc_hashin_shtrikman(double poro, double ks, double mus, double *k, double *mu,
                          double *im_biot, double *c_biot) {
  double ha = 1.0 - poro;
  double hb = 9.0 + 6.0 * poro;
  double hc = 8.0 + 12.0 * poro;
  *k = (4.0 * ks * mus * ha) / (3.0 * ks * poro + 4.0 * mus);
  *mu = (mus * ha * (9.0 * ks + 8 * mus)) / (ks * hb + mus * hc);
  *c_biot = 1.0 - *k / ks;
  *im_biot = (*c_biot - poro) / ks;
}

/* ks_and_mus (Poromec.f:48-111): the solid moduli whose Hashin-Shtrikman
 * bound at poro is the drained (k, mu); returns -1 where the reference
 * stops (no real root, negative moduli) */
int orc_ks_and_mus(double ym, double ps, double poro, double *k, double *mu) {
  double ha = 1.0 - poro;
  double hb = 9.0 + 6.0 * poro;
  double hc = 8.0 + 12.0 * poro;
  double a = 32.0 * ha * ha;
  double b = 4.0 * ha * (9.0 * *k - *mu * hc - 6.0 * poro * *k);
  double c = 3.0 * *mu * hc * poro * *k - 4.0 * *k * *mu * hb;
  double delta = b * b - 4.0 * a * c, mus;
  if (delta > 0) {
    delta = sqrt(delta);
    mus = 0.5 * (delta - b) / a;
    if (mus < 0) return -1;
  } else {
    return -1;
  }
  double ks = 4.0 * *k * mus / (4.0 * mus * ha - 3.0 * poro * *k);
  if (ks < 0) return -1;
  *k = ks;
  *mu = mus;
  (void)ym;
  (void)ps;
  return 0;
}

/* extrapol_hexa20_v2 (Elmlib.f:3540-3590): the trilinear fit through the 8
 * corner integration points (1, 4, 13, 16, 49, 52, 61, 64 of the 4x4x4
 * rule), evaluated at the 8 vertices.  up: ndf x 64, ue: ndf x 8. */
void orc_extrapol_hexa20(const double *up, double *ue, int ndf) {
  static const int IP[8] = {1, 4, 13, 16, 49, 52, 61, 64};
  const double c1 = 0.125000000000000, c2 = 0.145157042290566;
  const double c3 = 0.168564535400043, c4 = 0.195746635145486;
  for (int i = 0; i < ndf; i++) {
    double f[9], a[9];
    for (int j = 1; j <= 8; j++) f[j] = up[(IP[j - 1] - 1) * ndf + i];
    a[1] = c1 * (f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8]);
    a[2] = c2 * (f[2] + f[4] + f[6] + f[8] - (f[1] + f[3] + f[5] + f[7]));
    a[3] = c2 * (f[3] + f[4] + f[7] + f[8] - (f[1] + f[2] + f[5] + f[6]));
    a[4] = c2 * (f[5] + f[6] + f[7] + f[8] - (f[1] + f[2] + f[3] + f[4]));
    a[5] = c3 * (f[1] + f[4] + f[5] + f[8] - (f[2] + f[3] + f[6] + f[7]));
    a[6] = c3 * (f[1] + f[2] + f[7] + f[8] - (f[3] + f[4] + f[5] + f[6]));
    a[7] = c3 * (f[1] + f[3] + f[6] + f[8] - (f[2] + f[4] + f[5] + f[7]));
    a[8] = c4 * (f[2] + f[3] + f[5] + f[8] - (f[1] + f[4] + f[6] + f[7]));
    for (int j = 0; j < 8; j++) {
      double rnj = RN[j], snj = SN[j], tnj = TN[j];
      ue[j * ndf + i] = a[1] + a[2] * rnj + a[3] * snj + a[4] * tnj + a[5] * rnj * snj +
                        a[6] * snj * tnj + a[7] * rnj * tnj + a[8] * rnj * snj * tnj;
    }
  }
}

/* elmt18_pm (Elmlib03.f:2640-3640): the hexa20/8 Biot element with
 * properties per integration point, vp (7 x 64).  e(1), e(2) are the
 * solid's K and mu after initial_prop, e(3) the initial permeability, e(8)
 * the initial porosity.  isw 2, 4: the elmt17_pm integrals with the point's
 * properties; isw 3: stresses and flux at the points, extrapolated to the
 * vertices; isw 7: the points' porosity extrapolated to the vertices;
 * isw 10: porosity += alpha tr(du) + dp / M at each point (alpha, M the
 * material's), then Kozeny-Carman permeability, mixture density and
 * Hashin-Shtrikman moduli of the new porosity, in vp. */
int orc_elmt18(const double *e, const int *iqm, const int *iqh, const double *x,
               const double *u, const double *dp, double *p, double *s, double *vp, int isw,
               const orc_env *env) {
  if (isw == 2 || isw == 4) return elmt_hexa20(e, iqm, iqh, x, u, dp, p, s, isw, env, vp);
  double hu[20], hux[20], huy[20], huz[20], hp[8], hpx[8], hpy[8], hpz[8];
  double xj[3][3], xji[3][3], det, epsi[6];
  const double scale = 1.0e-06;
  const double div43 = 0.133333333333333e+01, div23 = 0.666666666666667;
  double gl[3] = {env->gravity[0], env->gravity[1], env->gravity[2]};
  double pm_d = e[5] * scale, fluid_d = e[6] * scale;
  double fluid_sw = fluid_d * env->gravity_mod;
  double modK = e[0], mu = e[1];
  double perm0 = e[2], perm = perm0 / fluid_sw;
  double imod_biot = 1.0 / e[3], coef_biot = e[4];
  double a = modK + div43 * mu, b = (modK - div23 * mu) / a, c = mu / a;
  double poro0 = e[7];
  if (isw == 3) {
    double itx[64 * 6], itxb[64 * 6], iflux[64 * 3], etx[48], etxb[48], eflux[24];
    int inpi = 0;
    for (int lz = 0; lz < 4; lz++) {
      double ti = PG4[lz];
      for (int ly = 0; ly < 4; ly++) {
        double si = PG4[ly];
        for (int lx = 0; lx < 4; lx++) {
          inpi++;
          double ri = PG4[lx];
          prop_porosity(&a, &b, &c, &imod_biot, &coef_biot, &perm, vp + 7 * (inpi - 1),
                        fluid_sw);
          sfhexa8(hp, hpx, hpy, hpz, ri, si, ti, 1, 1);
          if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
          sfhexa20(hu, hux, huy, huz, ri, si, ti, 0, 1);
          jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
          deform3d(hux, huy, huz, u, epsi, 20);
          double *t = itx + 6 * (inpi - 1), *tb = itxb + 6 * (inpi - 1);
          stress3d(a, b, c, epsi, t);
          double dpm = hp[0] * dp[0] + hp[1] * dp[1] + hp[2] * dp[2] + hp[3] * dp[3] +
                       hp[4] * dp[4] + hp[5] * dp[5] + hp[6] * dp[6] + hp[7] * dp[7];
          double pi = hp[0] * U(61) + hp[1] * U(62) + hp[2] * U(63) + hp[3] * U(64) +
                      hp[4] * U(65) + hp[5] * U(66) + hp[6] * U(67) + hp[7] * U(68);
          dpm = coef_biot * dpm;
          double pm = coef_biot * pi;
          t[0] = t[0] - dpm;
          t[1] = t[1] - dpm;
          t[2] = t[2] - dpm;
          tb[0] = t[0] + pm;
          tb[1] = t[1] + pm;
          tb[2] = t[2] + pm;
          tb[3] = t[3];
          tb[4] = t[4];
          tb[5] = t[5];
          double *q = iflux + 3 * (inpi - 1);
          q[0] = fluid_d * gl[0];
          q[1] = fluid_d * gl[1];
          q[2] = fluid_d * gl[2];
          for (int k = 0; k < 8; k++) {
            q[0] = q[0] - hpx[k] * U(61 + k);
            q[1] = q[1] - hpy[k] * U(61 + k);
            q[2] = q[2] - hpz[k] * U(61 + k);
          }
          q[0] = perm * q[0];
          q[1] = perm * q[1];
          q[2] = perm * q[2];
        }
      }
    }
    orc_extrapol_hexa20(iflux, eflux, 3);
    orc_extrapol_hexa20(itx, etx, 6);
    orc_extrapol_hexa20(itxb, etxb, 6);
    for (int i = 1; i <= 8; i++) {
      int tp = 6 * i - 5, tp1 = 6 * i + 43, tq = 3 * i + 94;
      for (int k = 0; k < 6; k++) {
        P(tp + k) = etx[6 * (i - 1) + k];
        P(tp1 + k) = etxb[6 * (i - 1) + k];
      }
      for (int k = 0; k < 3; k++) P(tq + k) = eflux[3 * (i - 1) + k];
    }
    return 0;
  }
  if (isw == 7) {
    double ip[64], pe[8];
    for (int i = 0; i < 64; i++) ip[i] = vp[7 * i];
    orc_extrapol_hexa20(ip, pe, 1);
    for (int i = 0; i < 8; i++) p[i] = pe[i];
    return 0;
  }
  if (isw == 10) {
    int inpi = 0;
    for (int lz = 0; lz < 4; lz++) {
      double ti = PG4[lz];
      for (int ly = 0; ly < 4; ly++) {
        double si = PG4[ly];
        for (int lx = 0; lx < 4; lx++) {
          inpi++;
          double ri = PG4[lx];
          sfhexa8(hp, hpx, hpy, hpz, ri, si, ti, 1, 1);
          if (jacob3d(hpx, hpy, hpz, xj, xji, x, &det, 8, 1)) return -1;
          sfhexa20(hu, hux, huy, huz, ri, si, ti, 0, 1);
          jacob3d(hux, huy, huz, xj, xji, x, &det, 20, 0);
          deform3d(hux, huy, huz, u, epsi, 20);
          double pi = hp[0] * U(61) + hp[1] * U(62) + hp[2] * U(63) + hp[3] * U(64) +
                      hp[4] * U(65) + hp[5] * U(66) + hp[6] * U(67) + hp[7] * U(68);
          pi = imod_biot * pi;
          double tmp = coef_biot * (epsi[0] + epsi[1] + epsi[2]) + pi;
          double *v = vp + 7 * (inpi - 1);
          v[0] = v[0] + tmp;
          double poro = v[0];
          v[1] = orc_konzey_carman(perm0, poro, poro0);
          v[2] = (1.0 - poro) * pm_d + poro * fluid_d;
          orc_hashin_shtrikman(poro, modK, mu, &v[3], &v[4], &v[5], &v[6]);
        }
      }
    }
    return 0;
  }
  (void)iqm;
  (void)iqh;
  (void)s;
  return -3;
}
