/*
 * pm_elmt07.c -- oracle of the equal-order Q1/Q1 hexa8 Biot element,
 * element type 7 (elmt07_pm).
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 * The reference reserves this element -- "ELMT07_PM - hexaedros de 8 nos
 * para o problema poromecanico elastico estatico (NAO IMPLEMENTADO)",
 * Elmlib03.f:22-23 -- and its mesh reader already sizes a hexa8 mesh read
 * with maxno 8 for it (nst = nenv*ndf = 32, Rdata.f:650-660), but the
 * element itself does not exist there, so there is no reference routine to
 * restate.  This is the weak form elmt17_pm integrates (Elmlib03.f:2022-2256)
 * with trilinear displacements in place of the serendipity ones:
 *
 *   Kuu = int B^T D B,  Kpu = -alpha int N_p div N_u,  Kup = Kpu^T,
 *   Kpp = -(int N N / M + dt perm int grad N . grad N)
 *
 * isw 2: the matrix and p = s u; isw 4: int B^T (D eps(u) - alpha N dp)
 * - rho_s int N g on the displacement rows, dt rho_f perm int grad N . g -
 * dt perm int grad N . grad p on the pressure rows, then the face loads
 * (types 40, 41 tractions and 43 fluxes over the 4-vertex faces); isw 3:
 * vertex stresses and Darcy flux; isw 7: vertex porosity change.  Material
 * constants e(1..8) mean what they mean for elmt17_pm; e(9) is the element's
 * own: the pressure stabilization coefficient beta (isw 2 below).
 *
 * Local order (nst = 32): u, v, w of vertex i at 3i-2 .. 3i, p of vertex i
 * at 24 + i (i = 1..8, the reference's hexa8 vertex order).  Volume
 * integrals use the 2 x 2 x 2 Gauss rule, face integrals the 2 x 2 rule;
 * both are exact for the trilinear products on parallelepipeds.  Points are
 * visited lz, ly, lx (lx fastest), the order elmt17_pm uses.
 *
 * Pinned without a reference (tests/test_oracle_elmt07.py): brick
 * matrices against exact Kronecker products of 1-D integrals, rigid-body
 * and constant-pressure null spaces, rotation covariance, the patch test on
 * a distorted mesh, isw 4 against isw 2 on the same state, isw 3 and isw 7
 * against closed forms of linear fields, and the consolidation column
 * against Terzaghi's series.
 */
#include <math.h>
#include <string.h>
#include "pm_oracle.h"

#define NST7 32

static const double G2 = 0.57735026918962576451; /* 1 / sqrt(3) */
static const double RN[8] = {1, -1, -1, 1, 1, -1, -1, 1};
static const double SN[8] = {1, 1, -1, -1, 1, 1, -1, -1};
static const double TN[8] = {1, 1, 1, 1, -1, -1, -1, -1};
/* the 4 vertices of each face (the vertex part of the reference's hexa20
 * face table hexa_face_node20), 1-based */
static const int FACE8[6][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}, {1, 5, 6, 2},
                                {4, 3, 7, 8}, {1, 4, 8, 5}, {2, 6, 7, 3}};

/* trilinear shape functions and their natural derivatives */
static void shape8(double r, double s, double t, double *h, double *hr, double *hs,
                   double *ht) {
  for (int k = 0; k < 8; k++) {
    double a = 1.0 + RN[k] * r, b = 1.0 + SN[k] * s, c = 1.0 + TN[k] * t;
    h[k] = 0.125 * a * b * c;
    hr[k] = 0.125 * RN[k] * b * c;
    hs[k] = 0.125 * a * SN[k] * c;
    ht[k] = 0.125 * a * b * TN[k];
  }
}

/* global derivatives in place; returns det J (<= 1e-14: -1 via *bad) */
static double to_global(const double *x, double *hx, double *hy, double *hz, int *bad) {
  double j[3][3], ji[3][3];
  for (int d = 0; d < 3; d++) {
    j[0][d] = j[1][d] = j[2][d] = 0.0;
    for (int k = 0; k < 8; k++) {
      j[0][d] += hx[k] * x[3 * k + d];
      j[1][d] += hy[k] * x[3 * k + d];
      j[2][d] += hz[k] * x[3 * k + d];
    }
  }
  double det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
               j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
               j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
  if (det <= 1.0e-14) {
    *bad = 1;
    return det;
  }
  /* inverse: d(r,s,t)/d(x,y,z); ji[d][a] = d r_a / d x_d */
  ji[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) / det;
  ji[0][1] = -(j[0][1] * j[2][2] - j[0][2] * j[2][1]) / det;
  ji[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) / det;
  ji[1][0] = -(j[1][0] * j[2][2] - j[1][2] * j[2][0]) / det;
  ji[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) / det;
  ji[1][2] = -(j[0][0] * j[1][2] - j[0][2] * j[1][0]) / det;
  ji[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) / det;
  ji[2][1] = -(j[0][0] * j[2][1] - j[0][1] * j[2][0]) / det;
  ji[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) / det;
  for (int k = 0; k < 8; k++) {
    double a = hx[k], b = hy[k], c = hz[k];
    hx[k] = ji[0][0] * a + ji[0][1] * b + ji[0][2] * c;
    hy[k] = ji[1][0] * a + ji[1][1] * b + ji[1][2] * c;
    hz[k] = ji[2][0] * a + ji[2][1] * b + ji[2][2] * c;
  }
  return det;
}

/* strain (xx, yy, zz, xy, yz, xz engineering shears) of the trilinear field */
static void strain8(const double *hx, const double *hy, const double *hz, const double *u,
                    double *eps) {
  double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
  for (int k = 0; k < 8; k++) {
    ux += hx[k] * u[3 * k];     uy += hy[k] * u[3 * k];     uz += hz[k] * u[3 * k];
    vx += hx[k] * u[3 * k + 1]; vy += hy[k] * u[3 * k + 1]; vz += hz[k] * u[3 * k + 1];
    wx += hx[k] * u[3 * k + 2]; wy += hy[k] * u[3 * k + 2]; wz += hz[k] * u[3 * k + 2];
  }
  eps[0] = ux; eps[1] = vy; eps[2] = wz;
  eps[3] = uy + vx; eps[4] = vz + wy; eps[5] = uz + wx;
}

/* D eps with D written as a [1 b b; b 1 b; b b 1; c c c] (elmt17_pm's
 * a, b, c: a = E (1-nu) / ((1+nu)(1-2nu)), b = nu/(1-nu), c = (1-2nu)/(2(1-nu))) */
static void stress(double a, double b, double c, const double *e, double *t) {
  t[0] = (e[0] + b * e[1] + b * e[2]) * a;
  t[1] = (b * e[0] + e[1] + b * e[2]) * a;
  t[2] = (b * e[0] + b * e[1] + e[2]) * a;
  t[3] = e[3] * c * a;
  t[4] = e[4] * c * a;
  t[5] = e[5] * c * a;
}

/* the face's 4 vertices in a plane frame: origin vertex 1, x along edge 1-2,
 * z the right-hand unit normal of the vertex order (1-2) x (1-3); returns
 * the 2-D coordinates, counterclockwise seen from +z, and that normal */
static void face_frame(const double *xf, double *x2, double *nrm) {
  double e1[3] = {xf[3] - xf[0], xf[4] - xf[1], xf[5] - xf[2]};
  double e2[3] = {xf[6] - xf[0], xf[7] - xf[1], xf[8] - xf[2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                 e1[0] * e2[1] - e1[1] * e2[0]};
  double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  double ax[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
  double az[3] = {n[0] / ln, n[1] / ln, n[2] / ln};
  double ay[3] = {az[1] * ax[2] - az[2] * ax[1], az[2] * ax[0] - az[0] * ax[2],
                  az[0] * ax[1] - az[1] * ax[0]};
  for (int i = 0; i < 4; i++) {
    double d[3] = {xf[3 * i] - xf[0], xf[3 * i + 1] - xf[1], xf[3 * i + 2] - xf[2]};
    x2[2 * i] = d[0] * ax[0] + d[1] * ax[1] + d[2] * ax[2];
    x2[2 * i + 1] = d[0] * ay[0] + d[1] * ay[1] + d[2] * ay[2];
  }
  for (int d = 0; d < 3; d++) nrm[d] = az[d];
}

/* area element of the bilinear face at (r, s) from its plane coordinates */
static double face_det(const double *x2, double r, double s) {
  const double qr[4] = {1, -1, -1, 1}, qs[4] = {1, 1, -1, -1};
  double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
  for (int k = 0; k < 4; k++) {
    double dr = 0.25 * qr[k] * (1.0 + qs[k] * s), ds = 0.25 * (1.0 + qr[k] * r) * qs[k];
    j00 += dr * x2[2 * k];
    j01 += dr * x2[2 * k + 1];
    j10 += ds * x2[2 * k];
    j11 += ds * x2[2 * k + 1];
  }
  return j00 * j11 - j01 * j10;
}

#define S7(i, j) s[(size_t)((j) - 1) * NST7 + (i) - 1]
#define P7(i) p[(i) - 1]
#define U7(i) u[(i) - 1]

int orc_elmt07(const double *e, const int *iqm, const int *iqh, const double *x,
               const double *u, const double *dp, double *p, double *s, int isw,
               const orc_env *env) {
  double h[8], hx[8], hy[8], hz[8], eps[6], tx[6];
  const double scale = 1.0e-06;
  const double *gl = env->gravity;
  const double pm_d = e[5] * scale, fluid_d = e[6] * scale;
  const double fluid_sw = fluid_d * env->gravity_mod;
  const double ym = e[0], ps = e[1];
  const double perm = e[2] / fluid_sw;
  const double imod_biot = 1.0 / e[3], coef_biot = e[4];
  const double dt_perm = perm * env->dt, dt_fluid_perm = fluid_d * dt_perm;
  const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
  const double a = (ym * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
  const double pg[2] = {-G2, G2};
  int bad = 0;

  if (isw == 2) {
    memset(s, 0, sizeof(double) * NST7 * NST7);
    /* pressure stabilization (e(9) = beta, 0: none): tau = beta h^2 / (lambda
     * + 2 mu), h = V^(1/3), adds -tau int grad N . grad N to Kpp -- the
     * perturbation tau d/dt (div grad p) of the flow equation, acting on the
     * step's pressure change as the storage term does (Aguilar, Gaspar,
     * Lisbona, Rodrigo, IJNME 75 (2008) 1282-1300; beta = 1/4 is their 1-D
     * value for equal-order linear elements) */
    double tau = 0.0;
    if (e[8] != 0.0) {
      double vol = 0.0;
      for (int lz = 0; lz < 2; lz++)
        for (int ly = 0; ly < 2; ly++)
          for (int lx = 0; lx < 2; lx++) {
            shape8(pg[lx], pg[ly], pg[lz], h, hx, hy, hz);
            vol = vol + to_global(x, hx, hy, hz, &bad);
          }
      if (bad) return -1;
      tau = e[8] * cbrt(vol) * cbrt(vol) / a;
    }
    for (int lz = 0; lz < 2; lz++)
      for (int ly = 0; ly < 2; ly++)
        for (int lx = 0; lx < 2; lx++) {
          shape8(pg[lx], pg[ly], pg[lz], h, hx, hy, hz);
          double w = to_global(x, hx, hy, hz, &bad);
          if (bad) return -1;
          double wt1 = w * a;
          for (int i = 1; i <= 8; i++) {
            int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
            double xi = hx[i - 1], yi = hy[i - 1], zi = hz[i - 1];
            for (int j = 1; j <= 8; j++) {
              int k1 = 3 * j - 2, k2 = k1 + 1, k3 = k2 + 1;
              double xj = hx[j - 1], yj = hy[j - 1], zj = hz[j - 1];
              S7(l1, k1) += (xi * xj + c * (yi * yj + zi * zj)) * wt1;
              S7(l1, k2) += (b * xi * yj + c * yi * xj) * wt1;
              S7(l1, k3) += (b * xi * zj + c * zi * xj) * wt1;
              S7(l2, k1) += (b * yi * xj + c * xi * yj) * wt1;
              S7(l2, k2) += (yi * yj + c * (xi * xj + zi * zj)) * wt1;
              S7(l2, k3) += (b * yi * zj + c * zi * yj) * wt1;
              S7(l3, k1) += (b * zi * xj + c * xi * zj) * wt1;
              S7(l3, k2) += (b * zi * yj + c * yi * zj) * wt1;
              S7(l3, k3) += (zi * zj + c * (yi * yj + xi * xj)) * wt1;
            }
          }
          for (int i = 1; i <= 8; i++) {
            int l = 24 + i;
            double wq = h[i - 1] * (w * coef_biot);
            for (int j = 1; j <= 8; j++) {
              S7(l, 3 * j - 2) -= hx[j - 1] * wq;
              S7(l, 3 * j - 1) -= hy[j - 1] * wq;
              S7(l, 3 * j) -= hz[j - 1] * wq;
            }
          }
          double w1 = w * imod_biot, w2 = w * dt_perm;
          for (int i = 1; i <= 8; i++)
            for (int j = 1; j <= 8; j++)
              S7(24 + i, 24 + j) -= w1 * h[i - 1] * h[j - 1] +
                                    w2 * (hx[i - 1] * hx[j - 1] + hy[i - 1] * hy[j - 1] +
                                          hz[i - 1] * hz[j - 1]);
          if (tau != 0.0) {
            double w3 = w * tau;
            for (int i = 1; i <= 8; i++)
              for (int j = 1; j <= 8; j++)
                S7(24 + i, 24 + j) -= w3 * (hx[i - 1] * hx[j - 1] + hy[i - 1] * hy[j - 1] +
                                            hz[i - 1] * hz[j - 1]);
          }
        }
    for (int i = 25; i <= 32; i++)
      for (int j = 1; j <= 24; j++) S7(j, i) = S7(i, j);
    for (int i = 1; i <= NST7; i++) {
      double v = 0.0;
      for (int j = 1; j <= NST7; j++) v += S7(i, j) * U7(j);
      P7(i) = v;
    }
    return 0;
  }

  if (isw == 4) {
    for (int lz = 0; lz < 2; lz++)
      for (int ly = 0; ly < 2; ly++)
        for (int lx = 0; lx < 2; lx++) {
          shape8(pg[lx], pg[ly], pg[lz], h, hx, hy, hz);
          double w = to_global(x, hx, hy, hz, &bad);
          if (bad) return -1;
          strain8(hx, hy, hz, u, eps);
          stress(a, b, c, eps, tx);
          double dpm = 0.0;
          for (int k = 0; k < 8; k++) dpm += h[k] * dp[k];
          dpm = coef_biot * dpm;
          tx[0] -= dpm; tx[1] -= dpm; tx[2] -= dpm;
          double gpx = 0, gpy = 0, gpz = 0;
          for (int k = 0; k < 8; k++) {
            gpx += hx[k] * U7(25 + k);
            gpy += hy[k] * U7(25 + k);
            gpz += hz[k] * U7(25 + k);
          }
          double wg = w * pm_d;
          for (int i = 1; i <= 8; i++) {
            int l1 = 3 * i - 2, l2 = l1 + 1, l3 = l2 + 1;
            double xi = hx[i - 1], yi = hy[i - 1], zi = hz[i - 1];
            P7(l1) += w * (xi * tx[0] + yi * tx[3] + zi * tx[5]);
            P7(l2) += w * (xi * tx[3] + yi * tx[1] + zi * tx[4]);
            P7(l3) += w * (xi * tx[5] + yi * tx[4] + zi * tx[2]);
            double wt = h[i - 1] * wg;
            P7(l1) -= wt * gl[0];
            P7(l2) -= wt * gl[1];
            P7(l3) -= wt * gl[2];
            P7(24 + i) += w * dt_fluid_perm * (xi * gl[0] + yi * gl[1] + zi * gl[2]);
            P7(24 + i) -= w * dt_perm * (xi * gpx + yi * gpy + zi * gpz);
          }
        }
    /* face loads: each face names its own load (no carry-over between faces) */
    for (int j = 1; j <= 6; j++) {
      int nm = iqm[j - 1], nh = iqh[j - 1];
      if (nm <= 0 && nh <= 0) continue;
      double xf[12], x2[8], nrm[3];
      for (int i = 0; i < 4; i++)
        for (int d = 0; d < 3; d++) xf[3 * i + d] = x[3 * (FACE8[j - 1][i] - 1) + d];
      face_frame(xf, x2, nrm);
      double f[3] = {0, 0, 0}, qh = 0.0, mp[3] = {0, 0, 0};
      int tm = nm > 0 ? env->load[3 * (nm - 1)] : 0;
      int th = nh > 0 ? env->load[3 * (nh - 1)] : 0;
      if (nm > 0) {
        if (tm == 40) {
          orc_tload(nm, env->t, mp, 0.0, f, env);
        } else if (tm == 41) {
          /* a normal traction: the load value times the outward unit
           * normal, as elmt17_pm applies it (Elmlib03.f:2401-2424); the
           * right-hand normal of the vertex order points outward on faces
           * 1 and 3..6 and inward on face 2 */
          double q = orc_tload(nm, env->t, mp, 0.0, NULL, env);
          double sg = j == 2 ? 1.0 : -1.0;
          for (int d = 0; d < 3; d++) f[d] = -sg * nrm[d] * q;
        } else {
          return -2;
        }
      }
      if (nh > 0) {
        if (th != 43) return -2;
        qh = orc_tload(nh, env->t, mp, 0.0, NULL, env);
      }
      for (int ly = 0; ly < 2; ly++)
        for (int lx = 0; lx < 2; lx++) {
          double r = pg[lx], sv = pg[ly];
          double w = face_det(x2, r, sv);
          if (w <= 1.0e-14) return -1;
          const double qr[4] = {1, -1, -1, 1}, qs[4] = {1, 1, -1, -1};
          for (int i = 0; i < 4; i++) {
            double hi = 0.25 * (1.0 + qr[i] * r) * (1.0 + qs[i] * sv) * w;
            int no = FACE8[j - 1][i];
            if (nm > 0)
              for (int d = 0; d < 3; d++) P7(3 * no - 2 + d) -= f[d] * hi;
            if (nh > 0) P7(24 + no) -= env->dt * qh * hi;
          }
        }
    }
    return 0;
  }

  if (isw == 3) {
    /* p(6i-5 .. 6i): total stress at vertex i; p(48 + ..): + alpha p;
     * p(96 + 3i-2 ..): perm (rho_f g - grad p) */
    for (int i = 1; i <= 8; i++) {
      shape8(RN[i - 1], SN[i - 1], TN[i - 1], h, hx, hy, hz);
      to_global(x, hx, hy, hz, &bad);
      if (bad) return -1;
      strain8(hx, hy, hz, u, eps);
      double *t6 = &P7(6 * i - 5), *b6 = &P7(6 * i + 43), *q = &P7(3 * i + 94);
      stress(a, b, c, eps, t6);
      double dpm = coef_biot * dp[i - 1], pm = coef_biot * U7(24 + i);
      t6[0] -= dpm; t6[1] -= dpm; t6[2] -= dpm;
      for (int k = 0; k < 6; k++) b6[k] = t6[k] + (k < 3 ? pm : 0.0);
      q[0] = fluid_d * gl[0];
      q[1] = fluid_d * gl[1];
      q[2] = fluid_d * gl[2];
      for (int k = 0; k < 8; k++) {
        q[0] -= hx[k] * U7(25 + k);
        q[1] -= hy[k] * U7(25 + k);
        q[2] -= hz[k] * U7(25 + k);
      }
      q[0] *= perm; q[1] *= perm; q[2] *= perm;
    }
    return 0;
  }

  if (isw == 7) {
    for (int i = 1; i <= 8; i++) {
      shape8(RN[i - 1], SN[i - 1], TN[i - 1], h, hx, hy, hz);
      to_global(x, hx, hy, hz, &bad);
      if (bad) return -1;
      strain8(hx, hy, hz, u, eps);
      P7(i) = coef_biot * (eps[0] + eps[1] + eps[2]) + imod_biot * U7(24 + i);
    }
    return 0;
  }
  return -3;
}
