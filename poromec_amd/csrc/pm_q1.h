// pm_q1.h -- device helpers of the equal-order Q1/Q1 hexa8 Biot element
// (element type 7; the reference reserves it as ELMT07_PM, Elmlib03.f:22-23,
// without implementing it).  Trilinear shape functions at the reference's
// hexa8 vertex order (sfhexa8_m, Elmlib.f:1536-1544), the 2 x 2 x 2 Gauss
// rule (points +-1/sqrt(3), weights 1) and the Jacobian.  Every expression is
// grouped as in the element's CPU statement, so that with contraction off
// the kernels reproduce its element arrays bit for bit.
#pragma once
#include "pm_device.h"

#define Q1_NST 32  // u, v, w of the 8 vertices, then their 8 pressures

static __constant__ double c_q1_pg[2] = {-0.57735026918962576451, 0.57735026918962576451};
static __constant__ double c_q1_rn[8] = {1, -1, -1, 1, 1, -1, -1, 1};
static __constant__ double c_q1_sn[8] = {1, 1, -1, -1, 1, 1, -1, -1};
static __constant__ double c_q1_tn[8] = {1, 1, 1, 1, -1, -1, -1, -1};

// trilinear shape functions and natural derivatives of the 8 vertices
__device__ __forceinline__ void q1_shape(double r, double s, double t, double *h, double *hr,
                                         double *hs, double *ht) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double a = 1.0 + c_q1_rn[k] * r, b = 1.0 + c_q1_sn[k] * s, c = 1.0 + c_q1_tn[k] * t;
    h[k] = 0.125 * a * b * c;
    hr[k] = 0.125 * c_q1_rn[k] * b * c;
    hs[k] = 0.125 * a * c_q1_sn[k] * c;
    ht[k] = 0.125 * a * b * c_q1_tn[k];
  }
}

// Jacobian of the vertex geometry from the natural derivatives: returns det
// and ji[d][a] = d r_a / d x_d
__device__ __forceinline__ double q1_jacobian(const double (*xv)[3], const double *hr,
                                              const double *hs, const double *ht,
                                              double ji[3][3]) {
  double j[3][3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    j[0][d] = 0.0;
    j[1][d] = 0.0;
    j[2][d] = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      j[0][d] = j[0][d] + hr[k] * xv[k][d];
      j[1][d] = j[1][d] + hs[k] * xv[k][d];
      j[2][d] = j[2][d] + ht[k] * xv[k][d];
    }
  }
  const double det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                     j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                     j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
  ji[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) / det;
  ji[0][1] = -(j[0][1] * j[2][2] - j[0][2] * j[2][1]) / det;
  ji[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) / det;
  ji[1][0] = -(j[1][0] * j[2][2] - j[1][2] * j[2][0]) / det;
  ji[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) / det;
  ji[1][2] = -(j[0][0] * j[1][2] - j[0][2] * j[1][0]) / det;
  ji[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) / det;
  ji[2][1] = -(j[0][0] * j[2][1] - j[0][1] * j[2][0]) / det;
  ji[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) / det;
  return det;
}

// global derivatives of vertex k
__device__ __forceinline__ void q1_grad(const double ji[3][3], double a, double b, double c,
                                        double &gx, double &gy, double &gz) {
  gx = ji[0][0] * a + ji[0][1] * b + ji[0][2] * c;
  gy = ji[1][0] * a + ji[1][1] * b + ji[1][2] * c;
  gz = ji[2][0] * a + ji[2][1] * b + ji[2][2] * c;
}

// strain (xx, yy, zz, xy, yz, xz, engineering shears) of the trilinear
// field u (u, v, w of each vertex) from the global derivatives
__device__ __forceinline__ void q1_strain(const double *hx, const double *hy, const double *hz,
                                          const double *u, double *eps) {
  double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    ux = ux + hx[k] * u[3 * k];
    uy = uy + hy[k] * u[3 * k];
    uz = uz + hz[k] * u[3 * k];
    vx = vx + hx[k] * u[3 * k + 1];
    vy = vy + hy[k] * u[3 * k + 1];
    vz = vz + hz[k] * u[3 * k + 1];
    wx = wx + hx[k] * u[3 * k + 2];
    wy = wy + hy[k] * u[3 * k + 2];
    wz = wz + hz[k] * u[3 * k + 2];
  }
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = wz;
  eps[3] = uy + vx;
  eps[4] = vz + wy;
  eps[5] = uz + wx;
}

// D eps, D = a [1 b b; b 1 b; b b 1; c c c] (stress3d, Elmlib.f:2906-2936)
__device__ __forceinline__ void q1_stress(double a, double b, double c, const double *e,
                                          double *t) {
  t[0] = (e[0] + b * e[1] + b * e[2]) * a;
  t[1] = (b * e[0] + e[1] + b * e[2]) * a;
  t[2] = (b * e[0] + b * e[1] + e[2]) * a;
  t[3] = e[3] * c * a;
  t[4] = e[4] * c * a;
  t[5] = e[5] * c * a;
}
