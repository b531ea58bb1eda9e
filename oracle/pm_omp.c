/*
 * pm_omp.c -- the reference's threaded CPU path, restated with OpenMP: what
 * bench.py times on the host cores as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 *   omp_colour        element colouring: a colour never holds two elements
 *                     that share a vertex (greedy, as coloredmesh/setcolor,
 *                     Colormsh.f:9-308)
 *   omp_pform_sample  the coloured element loop of pform_pm with omp_elmt
 *                     (Pform.f:172-303): gather, elmt17_pm, csr scatter, one
 *                     `omp do` per colour; here over a bounded sample of one
 *                     colour's elements, for timing
 *   omp_rpsqmr        rpsqmr_omp (Iters_omp.f:349-700): one parallel region
 *                     around the whole solve, with
 *     - matvec_csrc_sym_pm_omp (Matvec_omp.f:461-587): rows split by
 *       nonzeros (partition_csrc_bynonzeros, Matrix_partition.f:130-162), the
 *       transposed terms of each thread's rows summed into its own buffer
 *       thread_y from its "height" (compute_effective_work,
 *       Matrix_partition.f:222-236) up, then every buffer added into y;
 *     - dot_par_omp (Matvec_omp.f:845-873): `omp do reduction(+)` into a
 *       shared sum.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "pm_oracle.h"

#define NEN1 (ORC_NEN + 1)
#define NST ORC_NST

int omp_colour(const orc_model *md, int *elcolor, int maxcolors) {
  const int numel = md->numel, nnode = md->nnode;
  /* vertex -> elements */
  int64_t *ptr = calloc((size_t)nnode + 1, sizeof(int64_t));
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < ORC_NENV; k++) ptr[md->ix[(size_t)e * NEN1 + k]]++;
  for (int n = 0; n < nnode; n++) ptr[n + 1] += ptr[n];
  int *adj = malloc(sizeof(int) * (size_t)ptr[nnode]);
  int64_t *fill = malloc(sizeof(int64_t) * (size_t)nnode);
  for (int n = 0; n < nnode; n++) fill[n] = ptr[n];
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < ORC_NENV; k++) {
      int n = md->ix[(size_t)e * NEN1 + k] - 1;
      adj[fill[n]++] = e;
    }
  int *mark = malloc(sizeof(int) * (size_t)(maxcolors + 1));
  for (int c = 0; c <= maxcolors; c++) mark[c] = -1;
  int ncolors = 0;
  for (int e = 0; e < numel; e++) elcolor[e] = 0;
  for (int e = 0; e < numel; e++) {
    for (int k = 0; k < ORC_NENV; k++) {
      int n = md->ix[(size_t)e * NEN1 + k] - 1;
      for (int64_t j = ptr[n]; j < ptr[n + 1]; j++) {
        int o = adj[j];
        if (o < e) mark[elcolor[o]] = e;
      }
    }
    int c = 1;
    while (c <= maxcolors && mark[c] == e) c++;
    if (c > maxcolors) {
      ncolors = -1;
      break;
    }
    elcolor[e] = c;
    if (c > ncolors) ncolors = c;
  }
  free(mark);
  free(fill);
  free(adj);
  free(ptr);
  return ncolors;
}

double omp_pform_sample(orc_model *md, int isw, int lhs, int rhs, double *bvec,
                        const orc_env *env, const int *elcolor, int colour, int64_t max_elems,
                        int nthreads, int64_t *nel_done) {
  int64_t n = 0;
  int *list = malloc(sizeof(int) * (size_t)md->numel);
  for (int e = 0; e < md->numel && n < max_elems; e++)
    if (elcolor[e] == colour) list[n++] = e;
  double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
  {
    int ld[NST];
    double ul[NST], pl[NST], dpl[ORC_NENV], xl[3 * ORC_NENV];
    double *sl = malloc(sizeof(double) * NST * NST);
#pragma omp for schedule(static)
    for (int64_t j = 0; j < n; j++) {
      const int nel = list[j];
      orc_gather(md, nel, ld, ul, dpl, xl, pl);
      const int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
      const double *el = md->e + (size_t)(ma - 1) * ORC_PROP;
      if (md->nen == 8)
        orc_elmt07(el, md->eload + 7 * (size_t)nel, md->eloadp + 7 * (size_t)nel, xl, ul, dpl,
                   pl, sl, isw, env);
      else
        orc_elmt17(el, md->eload + 7 * (size_t)nel, md->eloadp + 7 * (size_t)nel, xl, ul, dpl,
                   pl, sl, isw, env);
      orc_assemble_csr(sl, pl, ld, md->ia, md->ja, md->al, md->ad, bvec, md->neq, lhs, rhs,
                       orc_nst(md));
    }
    free(sl);
  }
  double t = omp_get_wtime() - t0;
  free(list);
  if (nel_done) *nel_done = n;
  return t;
}

/* thread row ranges: partition_csrc_bynonzeros + compute_effective_work
 * (0-based rows; height[t] is the lowest row thread t's transposed terms
 * reach) */
static void partition(int neq, const int64_t *ia, const int *ja, int nth, int *begin, int *end,
                      int *height) {
  const int64_t nad = ia[neq] - 1;
  const int64_t mean = (2 * nad + neq) / nth + 1;
  int row = 0; /* next row to hand out */
  begin[0] = 0;
  for (int t = 0; t < nth - 1; t++) {
    int64_t size = 0;
    end[t] = begin[t] - 1;
    while (row < neq) {
      const int64_t tam = 2 * (ia[row + 1] - ia[row]) + 1;
      if (size + tam > mean) break;
      size += tam;
      end[t] = row;
      row++;
    }
    begin[t + 1] = row;
  }
  end[nth - 1] = neq - 1;
  for (int t = 0; t < nth; t++) {
    int h = begin[t];
    for (int i = begin[t]; i <= end[t]; i++)
      if (ia[i + 1] > ia[i] && ja[ia[i] - 1] - 1 < h) h = ja[ia[i] - 1] - 1;
    height[t] = h;
  }
}

/* shared state of the solver's parallel region */
static double g_dot;

static double dot_omp(const double *a, const double *b, int n) {
#pragma omp single
  g_dot = 0.0;
#pragma omp for reduction(+ : g_dot) schedule(static)
  for (int i = 0; i < n; i++) g_dot = g_dot + a[i] * b[i];
  const double v = g_dot; /* complete after the implicit barrier of the `for` */
#pragma omp barrier
  return v; /* every thread has read it before the next dot clears it */
}

static void matvec_omp(int neq, const int64_t *ia, const int *ja, const double *ad,
                       const double *al, const double *x, double *y, double *thread_y,
                       const int *begin, const int *end, const int *height, int nth) {
  const int id = omp_get_thread_num();
  for (int t = 0; t < nth; t++) {
    double *ty = thread_y + (size_t)t * neq;
#pragma omp for schedule(static)
    for (int j = height[t]; j < begin[t]; j++) ty[j] = 0.0;
  }
  double *ty = thread_y + (size_t)id * neq;
  for (int i = begin[id]; i <= end[id]; i++) {
    y[i] = 0.0;
    const double xi = x[i];
    double t = ad[i] * xi;
    for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) {
      const int jak = ja[k] - 1;
      const double s = al[k];
      t = t + s * x[jak];
      ty[jak] = ty[jak] + s * xi;
    }
    ty[i] = t;
  }
#pragma omp barrier
  for (int t = 0; t < nth; t++) {
    const double *tyt = thread_y + (size_t)t * neq;
#pragma omp for schedule(static)
    for (int j = height[t]; j <= end[t]; j++) y[j] = y[j] + tyt[j];
  }
}

int omp_rpsqmr(int neq, const int64_t *ia, const int *ja, const double *ad, const double *al,
               const double *m, const double *b, double *x, double tol, int maxit,
               int nthreads, double *res, int *fail) {
  for (int i = 0; i < neq; i++)
    if (ad[i] == 0.0) return -1;
  int *begin = malloc(sizeof(int) * nthreads), *end = malloc(sizeof(int) * nthreads),
      *height = malloc(sizeof(int) * nthreads);
  double *thread_y = malloc(sizeof(double) * (size_t)neq * nthreads);
  int nth = nthreads;
  double *t = malloc(sizeof(double) * neq), *r = malloc(sizeof(double) * neq),
         *q = malloc(sizeof(double) * neq), *d = malloc(sizeof(double) * neq);
  int ret = -2;
  double norm_r_out = 0.0;
  *fail = 0;
#pragma omp parallel num_threads(nthreads)
  {
    /* the rows are split over the team the runtime actually gave */
#pragma omp single
    {
      nth = omp_get_num_threads();
      partition(neq, ia, ja, nth, begin, end, height);
    }
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) x[i] = 0.0;
    const double norm_b = dot_omp(b, b, neq);
    const double conv = tol * sqrt(fabs(norm_b));
    matvec_omp(neq, ia, ja, ad, al, x, t, thread_y, begin, end, height, nth);
#pragma omp for schedule(static)
    for (int i = 0; i < neq; i++) {
      r[i] = b[i] - t[i];
      q[i] = r[i] * m[i];
      d[i] = 0.0;
    }
    double tau = sqrt(dot_omp(r, r, neq));
    double ro = dot_omp(r, q, neq);
    double v0 = 0.0, norm_r = 0.0;
    int j;
    for (j = 1; j <= maxit; j++) {
      matvec_omp(neq, ia, ja, ad, al, q, t, thread_y, begin, end, height, nth);
      const double sigma = dot_omp(q, t, neq);
      if (sigma == 0.0) break;
      const double alpha = ro / sigma;
#pragma omp for schedule(static)
      for (int i = 0; i < neq; i++) r[i] = r[i] - alpha * t[i];
      const double vn = sqrt(dot_omp(r, r, neq)) / tau;
      const double cn = 1.0 / sqrt(1.0 + vn * vn);
      tau = tau * vn * cn;
      const double tmp1 = cn * cn * v0 * v0, tmp2 = cn * cn * alpha;
#pragma omp for schedule(static)
      for (int i = 0; i < neq; i++) {
        d[i] = tmp1 * d[i] + tmp2 * q[i];
        x[i] = x[i] + d[i];
        t[i] = r[i] * m[i];
      }
      v0 = vn;
      if (ro == 0.0) break;
      const double ru = dot_omp(r, t, neq);
      const double beta = ru / ro;
      ro = ru;
#pragma omp for schedule(static)
      for (int i = 0; i < neq; i++) q[i] = t[i] + beta * q[i];
      norm_r = sqrt(dot_omp(r, r, neq));
      if (norm_r < conv) break;
    }
#pragma omp single
    {
      norm_r_out = norm_r;
      if (j <= maxit && norm_r < conv) ret = j;
      else if (j > maxit) *fail = 1;
      else ret = -3;
    }
  }
  if (res) *res = norm_r_out;
  free(t); free(r); free(q); free(d);
  free(thread_y);
  free(begin); free(end); free(height);
  return ret;
}
