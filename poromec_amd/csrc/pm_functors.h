// pm_functors.h -- per-row bodies and last-workgroup epilogues of the PCG and
// RPSQMR iterations (Iters.f:299-595, 6100-6436), shared by the single-system
// and the batched solvers.
#pragma once
#include "pm_device.h"

struct NoEpi {
  __device__ void operator()(PmScalars *, const double *) const {}
};
struct NoBody {
  __device__ void operator()(int64_t, const PmScalars &, double *) const {}
};

// scalar slot names
enum {
  S_ALPHA = 0, S_BETA, S_RO, S_TAU, S_V0, S_T1, S_T2, S_T3, S_T4, S_D, S_NORMR, S_NETA,
  S_BETAOLD, S_COLD, S_C, S_SOLD, S_S, S_W, S_RR0, S_NB, S_E0
};

// ---------------------------------------------------------------------------
// PCG (Iters.f:299-595)
// ---------------------------------------------------------------------------
struct PcgInit {
  const double *b, *m;
  double *x, *r, *z, *p;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double bi = b[i], zi = bi * m[i];
    x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    acc[0] += bi * zi;
  }
};
struct PcgInitEpi {
  double tol;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->conv = tol * sqrt(fabs(r[0]));
    S->s[S_D] = r[0];
    S->it = 0;
  }
};
struct PcgAlpha {
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->it += 1;
    S->s[S_ALPHA] = S->s[S_D] / r[0];
  }
};
struct PcgUpd {
  double *x, *r, *z;
  const double *p, *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double al = s.s[S_ALPHA];
    x[i] = x[i] + al * p[i];
    double ri = r[i] - al * z[i];
    r[i] = ri;
    double zi = ri * m[i];
    z[i] = zi;
    acc[0] += ri * zi;
  }
};
struct PcgUpdEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    double di = r[0];
    S->s[S_BETA] = di / S->s[S_D];
    S->s[S_D] = di;
    S->resid = sqrt(fabs(di));
    if (sqrt(fabs(di)) < S->conv) S->pending = 1;
  }
};
struct PcgP {
  double *p;
  const double *z;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    p[i] = z[i] + s.s[S_BETA] * p[i];
  }
};
struct FinishEpi {  // turns a pending convergence into done after the update
  __device__ void operator()(PmScalars *S, const double *) const {
    if (S->pending) {
      S->done = 1;
      S->iter = S->it;
    }
  }
};

// ---------------------------------------------------------------------------
// RPSQMR (Iters.f:6100-6436)
// ---------------------------------------------------------------------------
struct SqmrInit {
  const double *b, *m;
  double *x, *r, *q, *d;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    double qi = bi * m[i];
    q[i] = qi;
    d[i] = 0.0;
    acc[0] += bi * bi;
    acc[1] += bi * qi;
  }
};
struct SqmrInitEpi {
  double tol;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->conv = tol * sqrt(fabs(r[0]));
    S->s[S_TAU] = sqrt(r[0]);
    S->s[S_RO] = r[1];
    S->s[S_V0] = 0.0;
    S->it = 0;
  }
};
struct SqmrSigma {
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->it += 1;
    double sigma = r[0];
    if (sigma == 0.0) {
      S->done = PM_ERR_BREAKDOWN;
      return;
    }
    S->s[S_ALPHA] = S->s[S_RO] / sigma;
  }
};
// After t = A q and alpha (SqmrSigma), an iteration is two vector passes.
// SqmrRT: r -= alpha t, then t = M r, with r.r and r.t in one reduction
// (SqmrRTEpi: tau, c and the coefficients of d, then rho, beta and the stop
// test).  SqmrDXQ: d = T1 d + T2 q, x += d, q = t + beta q.  Each value is
// the expression of the reference's loop (Iters.f:6100-6436) and each dot
// runs over the rows in order, as the reference's separate loops do.
// Convergence and breakdown are held in `pending` until SqmrDXQ has made the
// iteration's update of x; SqmrFinishEpi then sets `done`.
struct SqmrREpi {  // the part of SqmrRTEpi that needs only r.r
  __device__ void operator()(PmScalars *S, const double *r) const {
    double vn = sqrt(r[0]) / S->s[S_TAU];
    double cn = 1.0 / sqrt(1.0 + vn * vn);
    S->s[S_TAU] = S->s[S_TAU] * vn * cn;
    S->s[S_T1] = cn * cn * S->s[S_V0] * S->s[S_V0];
    S->s[S_T2] = cn * cn * S->s[S_ALPHA];
    S->s[S_V0] = vn;
    S->s[S_NORMR] = sqrt(r[0]);
  }
};
struct SqmrRT {
  double *r, *t;
  const double *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double ri = r[i] - s.s[S_ALPHA] * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
    double ti = ri * m[i];
    t[i] = ti;
    acc[1] += ri * ti;
  }
};
struct SqmrRTEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    SqmrREpi{}(S, r);
    if (S->s[S_RO] == 0.0) {
      S->pending = PM_ERR_BREAKDOWN;
      return;
    }
    S->s[S_BETA] = r[1] / S->s[S_RO];
    S->s[S_RO] = r[1];
    S->resid = S->s[S_NORMR];
    if (S->s[S_NORMR] < S->conv) S->pending = 1;
  }
};
struct SqmrDXQ {
  double *d, *x, *q;
  const double *t;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    double qi = q[i];
    double di = s.s[S_T1] * d[i] + s.s[S_T2] * qi;
    d[i] = di;
    x[i] = x[i] + di;
    q[i] = t[i] + s.s[S_BETA] * qi;
  }
};
struct SqmrFinishEpi {  // a pending convergence or breakdown takes effect
  __device__ void operator()(PmScalars *S, const double *) const {
    if (S->pending == 1) {
      S->done = 1;
      S->iter = S->it;
    } else if (S->pending < 0) {
      S->done = S->pending;
    }
  }
};

