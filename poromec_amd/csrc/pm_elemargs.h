// pm_elemargs.h -- arguments of the element kernels (pform_pm's element
// loop, Pform.f:1-429) and the load-table lookup they share.
#pragma once
#include "pm_device.h"

struct ElemArgs {
  const int *ix;       // (numel, nen): 20 hexa20/8 nodes, or the 8 of a Q1/Q1 element
  const double *xv;    // (nnodev, 3)
  const int *eq;       // (nnode, 4)
  const int *mat;      // (numel)
  const double *e;     // (numat, 15)
  const int *eload;    // (numel, 14)
  const int *load;     // (200, 3)
  const double *fload; // (200, 5, 1000)
  const double *u;     // (nnode, 4)
  const double *dp;    // (nnode)
  double *b;           // right-hand side
  int *err;            // element error flag
  double dt, t, gmod;
  double grav[3];
  // matrix
  const int *eq_block, *fr;
  const int64_t *colptr, *valptr;
  const int *ncol;
  double *val;
  int lhs, rhs;
  const int *elems;    // elements of this colour
  int nelem;
  // reference summation order: element arrays are stored, not scattered
  double *ke;          // (numel, 68, 68) element matrices, or null
  double *pe;          // (numel, 68) element vectors, or null
  // variable properties: element type per material (null: all 17) and the
  // 7 properties of each element's 64 integration points (elmt18_pm)
  const int *ie;
  const double *vp;    // (numel, 64, 7)
};

__device__ inline double dev_interpol(const double *xs, const double *ys, double xi, int n) {
  if (xi <= xs[0]) return ys[0];
  if (xi >= xs[n - 1]) return ys[n - 1];
  for (int i = 0; i < n - 1; i++)
    if (xi >= xs[i] && xi <= xs[i + 1]) return ys[i] + (ys[i + 1] - ys[i]) * (xi - xs[i]) / (xs[i + 1] - xs[i]);
  return 0.0;
}
// fload(i, term, nc) of load.fi, read through the kernel argument A
#define DFLOAD(i, term, nc) \
  (A.fload + (size_t)((nc) - 1) * PM_NTERMS * PM_NPAR + (size_t)((term) - 1) * PM_NPAR + (i) - 1)


// Q1/Q1 element (element type 7, pm_elmt07.hip): one colour's elements
// through isw 2 or 4, and the vertex porosity change of isw 7
int pm_q1_elements(const ElemArgs &A, int isw, hipStream_t stream);
int pm_q1_porosity(const ElemArgs &A, double *dporo, hipStream_t stream);
