/*
 * pm_solver.c -- oracle restatement of the CSRC SpMV, dot, Jacobi and the
 * Krylov solvers of the mef path.
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 *   orc_matvec_csrc  matvec_csrc_sym_pm   Matvec.f:1504-1595
 *   orc_dot          dot_par              Matvec.f:1714-1743 (one rank)
 *   orc_precond      cal_precond/pre_diag Precond.f:1601-1695, 22-44
 *   orc_pcg          pcg                  Iters.f:299-595
 *   orc_gmres2       gmres2 + sym_ortho2  Iters.f:1411-1775, 6718-6745
 *   orc_pbicgstab    pbicgstab            Iters.f:2817-3105
 *   orc_pminres      pminres + sym_ortho  Iters.f:3987-4302, 6595-6688
 *   orc_rpsqmr       rpsqmr               Iters.f:6100-6436
 *   orc_solv         solv_pm dispatch     Solv.f:27-780 (SQMR -> GMRES fallback :92-94,765)
 * Return value of each solver: iterations taken (> 0), or a negative code:
 *   -1 zero diagonal, -2 maxit reached, -3 breakdown.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "pm_oracle.h"

/* y = A x with A = diag(ad) + L + L^T, L the strict lower CSRC (al).  The
 * reference never clears y: row i sets y(i) after rows j < i added their
 * transposed terms into y(j), and y(j) of a later row is set before any
 * transposed term reaches it. */
void orc_matvec_csrc(int neq, const int64_t *ia, const int *ja, const double *ad,
                     const double *al, const double *x, double *y) {
  for (int i = 0; i < neq; i++) {
    double xi = x[i];
    double t = ad[i] * xi;
    for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) {
      int jak = ja[k] - 1;
      double s = al[k];
      t = t + s * x[jak];
      y[jak] = y[jak] + s * xi;
    }
    y[i] = t;
  }
}

double orc_dot(const double *a, const double *b, int n) {
  double tmp = 0.0;
  for (int i = 0; i < n; i++) tmp = tmp + a[i] * b[i];
  return tmp;
}

void orc_precond(int precond, int neq, const double *ad, double *m) {
  if (precond == 2)
    for (int i = 0; i < neq; i++) m[i] = 1.0 / ad[i];
  else if (precond == 5)
    for (int i = 0; i < neq; i++) m[i] = fabs(1.0 / ad[i]);
  else
    for (int i = 0; i < neq; i++) m[i] = 1.0;
}

static int zero_diag(int neq, const double *ad) {
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) return 1;
  return 0;
}

#define MATVEC(xv, yv) orc_matvec_csrc(neq, ia, ja, ad, al, xv, yv)

int orc_pcg(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
            const double *m, const double *b, double *x, double tol, int maxit, double *res) {
  if (zero_diag(neq, ad)) return -1;
  double *z = malloc(sizeof(double) * neq), *r = malloc(sizeof(double) * neq),
         *p = malloc(sizeof(double) * neq);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  for (int i = 0; i < neq; i++) z[i] = b[i] * m[i];
  double d = orc_dot(b, z, neq);
  double conv = tol * sqrt(fabs(d));
  MATVEC(x, z);
  for (int i = 0; i < neq; i++) {
    r[i] = b[i] - z[i];
    z[i] = r[i] * m[i];
    p[i] = z[i];
  }
  d = orc_dot(r, z, neq);
  int j, ret = -2;
  for (j = 1; j <= maxit; j++) {
    MATVEC(p, z);
    double alpha = d / orc_dot(z, p, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * z[i];
      z[i] = r[i] * m[i];
    }
    double di = orc_dot(r, z, neq);
    double beta = di / d;
    for (int i = 0; i < neq; i++) p[i] = z[i] + beta * p[i];
    d = di;
    if (sqrt(fabs(d)) < conv) { ret = j; break; }
  }
  if (res) *res = sqrt(fabs(d));
  free(z); free(r); free(p);
  return ret;
}

int orc_rpsqmr(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
               const double *m, const double *b, double *x, double tol, int maxit,
               double *res, int *fail) {
  *fail = 0;
  if (zero_diag(neq, ad)) return -1;
  double *t = malloc(sizeof(double) * neq), *r = malloc(sizeof(double) * neq),
         *q = malloc(sizeof(double) * neq), *d = malloc(sizeof(double) * neq);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  double norm_b = orc_dot(b, b, neq);
  double conv = tol * sqrt(fabs(norm_b));
  MATVEC(x, t);
  for (int i = 0; i < neq; i++) {
    r[i] = b[i] - t[i];
    q[i] = r[i] * m[i];
    d[i] = 0.0;
  }
  double tau = sqrt(orc_dot(r, r, neq));
  double ro = orc_dot(r, q, neq);
  double v0 = 0.0, norm_r = 0.0;
  int j, ret = -2;
  for (j = 1; j <= maxit; j++) {
    MATVEC(q, t);
    double sigma = orc_dot(q, t, neq);
    if (sigma == 0.0) { ret = -3; break; }
    double alpha = ro / sigma;
    for (int i = 0; i < neq; i++) r[i] = r[i] - alpha * t[i];
    double vn = sqrt(orc_dot(r, r, neq)) / tau;
    double cn = 1.0 / sqrt(1.0 + vn * vn);
    tau = tau * vn * cn;
    double tmp1 = cn * cn * v0 * v0;
    double tmp2 = cn * cn * alpha;
    for (int i = 0; i < neq; i++) {
      d[i] = tmp1 * d[i] + tmp2 * q[i];
      x[i] = x[i] + d[i];
      t[i] = r[i] * m[i];
    }
    v0 = vn;
    if (ro == 0.0) { ret = -3; break; }
    tmp1 = orc_dot(r, t, neq);
    double beta = tmp1 / ro;
    ro = tmp1;
    for (int i = 0; i < neq; i++) q[i] = t[i] + beta * q[i];
    norm_r = sqrt(orc_dot(r, r, neq));
    if (norm_r < conv) { ret = j; break; }
  }
  if (ret == -2) *fail = 1;
  if (res) *res = norm_r;
  free(t); free(r); free(q); free(d);
  return ret;
}

int orc_pbicgstab(int neq, const int64_t *ia, const int *ja, const double *ad,
                  const double *al, const double *m, const double *b, double *x, double tol,
                  int maxit, double *res) {
  const double btol = 1.0e-32;
  if (zero_diag(neq, ad)) return -1;
  double *r0 = malloc(sizeof(double) * neq), *p = malloc(sizeof(double) * neq),
         *r = malloc(sizeof(double) * neq), *z = malloc(sizeof(double) * neq),
         *v = malloc(sizeof(double) * neq), *t = malloc(sizeof(double) * neq);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  double d = orc_dot(b, b, neq);
  double conv = tol * sqrt(d);
  double breaktol = btol * sqrt(d);
  MATVEC(x, z);
  for (int i = 0; i < neq; i++) {
    r0[i] = b[i] - z[i];
    p[i] = r0[i];
    r[i] = r0[i];
    z[i] = p[i] * m[i];
  }
  int j, ret = -2;
  for (j = 1; j <= maxit; j++) {
    MATVEC(z, v);
    double rr0 = orc_dot(r, r0, neq);
    if (sqrt(fabs(rr0)) < breaktol) { ret = -3; break; }
    double alpha = rr0 / orc_dot(v, r0, neq);
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + alpha * z[i];
      r[i] = r[i] - alpha * v[i];
      z[i] = r[i] * m[i];
    }
    d = orc_dot(r, r, neq);
    if (sqrt(fabs(d)) < conv) { ret = j; break; }
    MATVEC(z, t);
    double w = orc_dot(t, r, neq) / orc_dot(t, t, neq);
    if (fabs(w) < breaktol) { ret = -3; break; }
    for (int i = 0; i < neq; i++) {
      x[i] = x[i] + w * z[i];
      r[i] = r[i] - w * t[i];
    }
    d = orc_dot(r, r, neq);
    if (sqrt(fabs(d)) < conv) { ret = j; break; }
    double beta = (orc_dot(r, r0, neq) / rr0) * (alpha / w);
    for (int i = 0; i < neq; i++) {
      p[i] = r[i] + beta * (p[i] - w * v[i]);
      z[i] = p[i] * m[i];
    }
  }
  if (res) *res = sqrt(fabs(d));
  free(r0); free(p); free(r); free(z); free(v); free(t);
  return ret;
}

static double sign1(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }

/* sym_ortho (Iters.f:6595): note the third branch compares |b| with a, not |a| */
static void sym_ortho(double a, double b, double *c, double *s, double *r) {
  double ma = fabs(a), mb = fabs(b), sa = sign1(a), sb = sign1(b);
  if (b == 0.0) {
    *s = 0.0;
    *r = ma;
    *c = (a == 0.0) ? 1.0 : sa;
  } else if (a == 0.0) {
    *c = 0.0;
    *s = sb;
    *r = mb;
  } else if (mb > a) {
    double t = a / b;
    *s = sb / sqrt(1.0 + t * t);
    *c = *s * t;
    *r = b / *s;
  } else if (ma > mb) {
    double t = b / a;
    *c = sa / sqrt(1.0 + t * t);
    *s = *c * t;
    *r = a / *c;
  }
}

int orc_pminres(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
                const double *m, const double *b, double *x, double tol, int maxit,
                int nrestart, double *res) {
  if (zero_diag(neq, ad)) return -1;
  size_t nb = sizeof(double) * neq;
  double *v0 = malloc(nb), *v = malloc(nb), *w = malloc(nb), *w0 = malloc(nb),
         *w00 = malloc(nb), *z = malloc(nb), *z0 = malloc(nb), *p = malloc(nb);
  for (int i = 0; i < neq; i++) x[i] = 0.0;
  for (int i = 0; i < neq; i++) z[i] = b[i] * sqrt(fabs(m[i]));
  double tmp1 = orc_dot(z, z, neq);
  double conv = tol * sqrt(fabs(tmp1));
  int it = 0, ret = -2;
  double normr = 0.0;
  for (;;) {
    MATVEC(x, z);
    for (int i = 0; i < neq; i++) {
      v[i] = b[i] - z[i];
      v0[i] = 0.0;
      w0[i] = 0.0;
      w00[i] = 0.0;
      z0[i] = v[i] * m[i];
    }
    double beta = sqrt(fabs(orc_dot(v, z0, neq)));
    normr = beta;
    double neta = beta, beta_old = beta, c_old = 1.0, c = 1.0, s_old = 0.0, s = 0.0;
    int done = 0;
    for (int j = 1; j <= maxit; j++) {
      it++;
      MATVEC(z0, p);
      tmp1 = 1.0 / (beta * beta);
      double alpha = fabs(orc_dot(z0, p, neq)) * tmp1;
      tmp1 = 1.0 / beta;
      double tmp2 = -alpha / beta, tmp3 = -beta / beta_old;
      for (int i = 0; i < neq; i++) {
        p[i] = tmp1 * p[i] + tmp2 * v[i] + tmp3 * v0[i];
        z[i] = p[i] * m[i];
      }
      beta_old = beta;
      tmp1 = orc_dot(z, p, neq);
      beta = sqrt(fabs(tmp1));
      double delta = c * alpha - c_old * s * beta_old;
      double ro2 = s * alpha + c_old * c * beta_old;
      double ro3 = s_old * beta_old;
      c_old = c;
      s_old = s;
      double ro1 = 0.0;
      sym_ortho(delta, beta, &c, &s, &ro1);
      tmp1 = 1.0 / (beta_old * ro1);
      tmp2 = -ro2 / ro1;
      tmp3 = -ro3 / ro1;
      double tmp4 = c * neta;
      for (int i = 0; i < neq; i++) {
        w[i] = tmp1 * z0[i] + tmp2 * w0[i] + tmp3 * w00[i];
        x[i] = x[i] + tmp4 * w[i];
        v0[i] = v[i];
        v[i] = p[i];
        z0[i] = z[i];
        w00[i] = w0[i];
        w0[i] = w[i];
      }
      normr = fabs(s) * normr;
      if (normr < conv) { done = 1; break; }
      neta = -s * neta;
    }
    if (done) { ret = it; break; }
    if (nrestart != 0) { nrestart--; continue; }
    break;
  }
  if (res) *res = normr;
  free(v0); free(v); free(w); free(w0); free(w00); free(z); free(z0); free(p);
  return ret;
}

/* sym_ortho2 (Iters.f:6718-6745) */
static void sym_ortho2(double a, double b, double *c, double *s, double *r) {
  double ma = fabs(a), mb = fabs(b);
  *c = 1.0; *s = 0.0; *r = a;
  if (b != 0.0) {
    if (mb > ma) {
      double t = a / b;
      *s = 1.0 / sqrt(1.0 + t * t);
      *c = *s * t;
      *r = b / *s;
    } else {
      double t = b / a;
      *c = 1.0 / sqrt(1.0 + t * t);
      *s = *c * t;
      *r = a / *c;
    }
  }
}

int orc_gmres2(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
               const double *m, const double *b, double *x, int k, double tol, int maxit,
               double *res) {
  const double kesp = 0.25;
  double *g = malloc(sizeof(double) * (size_t)neq * (k + 1));
  double *h = calloc((size_t)(k + 1) * k, sizeof(double));
  double *y = malloc(sizeof(double) * k), *c = malloc(sizeof(double) * k),
         *s = malloc(sizeof(double) * k), *e = malloc(sizeof(double) * (k + 1));
#define G(i, j) g[(size_t)(j) * neq + (i)]
#define H(i, j) h[(size_t)(j) * (k + 1) + (i)]
  for (int i = 0; i < neq; i++) {
    x[i] = 0.0;
    G(i, 0) = b[i] * m[i];
  }
  double norm = sqrt(orc_dot(&G(0, 0), &G(0, 0), neq));
  double econv = tol * norm;
  int nit = 0, ni = 0, ret = -2;
  for (int l = 1; l <= maxit; l++) {
    MATVEC(x, &G(0, 0));
    for (int i = 0; i < neq; i++) G(i, 0) = (b[i] - G(i, 0)) * m[i];
    e[0] = sqrt(orc_dot(&G(0, 0), &G(0, 0), neq));
    double inorm = 1.0 / e[0];
    for (int i = 0; i < neq; i++) G(i, 0) = G(i, 0) * inorm;
    ni = 0;
    for (int i = 0; i < k; i++) {
      nit++;
      ni++;
      MATVEC(&G(0, i), &G(0, i + 1));
      for (int j = 0; j < neq; j++) G(j, i + 1) = G(j, i + 1) * m[j];
      double tau = sqrt(orc_dot(&G(0, i + 1), &G(0, i + 1), neq));
      for (int j = 0; j <= i; j++) {
        double beta = orc_dot(&G(0, i + 1), &G(0, j), neq);
        for (int ic = 0; ic < neq; ic++) G(ic, i + 1) = G(ic, i + 1) - beta * G(ic, j);
        H(j, i) = beta;
      }
      norm = sqrt(orc_dot(&G(0, i + 1), &G(0, i + 1), neq));
      if (norm <= kesp * tau) {
        for (int j = 0; j <= i; j++) {
          double beta = orc_dot(&G(0, i + 1), &G(0, j), neq);
          for (int ic = 0; ic < neq; ic++) G(ic, i + 1) = G(ic, i + 1) - beta * G(ic, j);
          H(j, i) = H(j, i) + beta;
        }
      }
      norm = sqrt(orc_dot(&G(0, i + 1), &G(0, i + 1), neq));
      H(i + 1, i) = norm;
      inorm = 1.0 / norm;
      for (int ic = 0; ic < neq; ic++) G(ic, i + 1) = G(ic, i + 1) * inorm;
      for (int j = 0; j < i; j++) {
        double aux1 = c[j] * H(j, i) + s[j] * H(j + 1, i);
        double aux2 = -s[j] * H(j, i) + c[j] * H(j + 1, i);
        H(j, i) = aux1;
        H(j + 1, i) = aux2;
      }
      double rr;
      sym_ortho2(H(i, i), H(i + 1, i), &c[i], &s[i], &rr);
      H(i, i) = rr;
      H(i + 1, i) = 0.0;
      e[i + 1] = -s[i] * e[i];
      e[i] = c[i] * e[i];
      if (fabs(e[i + 1]) <= econv) break;
    }
    for (int j = 0; j < ni; j++) y[j] = e[j];
    for (int j = ni - 1; j >= 1; j--) {
      y[j] = y[j] / H(j, j);
      double rr = y[j];
      for (int i = 0; i < j; i++) y[i] = y[i] - H(i, j) * rr;
    }
    y[0] = y[0] / H(0, 0);
    for (int j = 0; j < ni; j++) {
      double rr = y[j];
      for (int i = 0; i < neq; i++) x[i] = x[i] + rr * G(i, j);
    }
    if (fabs(e[ni]) <= econv) { ret = nit; break; }
  }
  if (res) *res = fabs(e[ni]);
#undef G
#undef H
  free(g); free(h); free(y); free(c); free(s); free(e);
  return ret;
}

int orc_solv(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
             double *m, const double *b, double *x, const orc_cfg *cfg, int *used,
             double *res) {
  int solver = cfg->solver;
  *used = solver;
  orc_precond(cfg->precond, neq, ad, m);
  if (solver == 1) return orc_pcg(neq, ia, ja, ad, al, m, b, x, cfg->solvtol, cfg->maxit, res);
  if (solver == 4)
    return orc_pbicgstab(neq, ia, ja, ad, al, m, b, x, cfg->solvtol, cfg->maxit, res);
  if (solver == 7)
    return orc_pminres(neq, ia, ja, ad, al, m, b, x, cfg->solvtol, cfg->maxit, 10, res);
  if (solver == 11) {
    int fail = 0;
    int it = orc_rpsqmr(neq, ia, ja, ad, al, m, b, x, cfg->solvtol, cfg->maxit, res, &fail);
    if (!fail) return it;
    *used = 2;
    orc_precond(cfg->precond, neq, ad, m);
  }
  return orc_gmres2(neq, ia, ja, ad, al, m, b, x, cfg->ngram, cfg->solvtol, cfg->maxit, res);
}
