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
struct SqmrR {
  double *r;
  const double *t;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double ri = r[i] - s.s[S_ALPHA] * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
  }
};
struct SqmrREpi {
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
struct SqmrD {
  double *d, *x, *t;
  const double *q, *r, *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double di = s.s[S_T1] * d[i] + s.s[S_T2] * q[i];
    d[i] = di;
    x[i] = x[i] + di;
    double ri = r[i], ti = ri * m[i];
    t[i] = ti;
    acc[0] += ri * ti;
  }
};
struct SqmrDEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    if (S->s[S_RO] == 0.0) {
      S->done = PM_ERR_BREAKDOWN;
      return;
    }
    S->s[S_BETA] = r[0] / S->s[S_RO];
    S->s[S_RO] = r[0];
    S->resid = S->s[S_NORMR];
    if (S->s[S_NORMR] < S->conv) {
      S->done = 1;
      S->iter = S->it;
    }
  }
};
struct SqmrQ {
  double *q;
  const double *t;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    q[i] = t[i] + s.s[S_BETA] * q[i];
  }
};

// The same iteration in two vector passes instead of three (the single-system
// loop; the batched loop keeps SqmrR / SqmrD / SqmrQ for its per-box
// epilogues).  The first pass does SqmrR's update and the part of SqmrD that
// needs only the new r (t = r m and r.t), so one reduction carries both
// dots; the second pass does the rest of SqmrD (d, x) and SqmrQ.  Every
// value is formed by the same expression and every sum over the same rows in
// the same order as in the three-pass form, so the results are the same to
// the bit.  Convergence and breakdown are held in `pending` until the second
// pass has made this iteration's update of x, as the three-pass form does.
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

