// pm_spmv.h -- the node-block SpMV body shared by the single-system and the
// batched kernels, with its tunable shape.
//
// A node block is the <= 4 rows of one free node; they share one sorted
// column list (int32) and store their values row by row.  LANES lanes of a
// wavefront (64, 32 or 16) work on one block, so a wavefront has 64/LANES
// blocks in flight; each lane takes columns lane, lane + LANES, ... and
// issues UNR of them back to back before it uses any, which raises the number
// of loads in flight per wave.  Each lane carries the partial sums of every
// row of its block; a shuffle butterfly over the LANES lanes finishes them.
#pragma once
#include "pm_device.h"

// (LANES, UNR) of each SpMV shape the tuner may pick
#define PM_SPMV_NCFG 6
static constexpr int pm_spmv_lanes[PM_SPMV_NCFG] = {64, 64, 64, 32, 32, 16};
static constexpr int pm_spmv_unr[PM_SPMV_NCFG] = {1, 2, 4, 2, 4, 4};

struct PmBlockMeta {
  int64_t cp, vp;
  int c, n;
};

// metadata of node block k (an empty block past the end k1)
__device__ __forceinline__ PmBlockMeta pm_block_meta(int64_t k, int64_t k1,
                                                     const int64_t *__restrict__ colptr,
                                                     const int64_t *__restrict__ valptr,
                                                     const int *__restrict__ nr) {
  PmBlockMeta m{0, 0, 0, 0};
  if (k < k1) {
    m.cp = colptr[k];
    m.c = (int)(colptr[k + 1] - m.cp);
    m.vp = valptr[k];
    m.n = nr[k];
  }
  return m;
}

template <int LANES>
__device__ __forceinline__ double pm_group_sum(double v) {
#pragma unroll
  for (int off = LANES / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// partial row sums of node block (cp, c, vp, n) held by this lane
template <typename VT, int LANES, int UNR>
__device__ __forceinline__ void pm_spmv_block(int lane, int64_t cp, int c, int64_t vp, int n,
                                              const int *__restrict__ ncol,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ x, double a[4]) {
  const int *cols = ncol + cp;
  const VT *v0 = val + vp;
  int j = lane;
  for (; j + (UNR - 1) * LANES < c; j += UNR * LANES) {
    int col[UNR];
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) col[u] = __builtin_nontemporal_load(cols + j + u * LANES);
#pragma unroll
    for (int u = 0; u < UNR; u++) xv[u] = x[col[u]];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const VT *v = v0 + j + u * LANES;
      a[0] += (double)__builtin_nontemporal_load(v) * xv[u];
      if (n > 1) a[1] += (double)__builtin_nontemporal_load(v + c) * xv[u];
      if (n > 2) a[2] += (double)__builtin_nontemporal_load(v + 2 * c) * xv[u];
      if (n > 3) a[3] += (double)__builtin_nontemporal_load(v + 3 * c) * xv[u];
    }
  }
  for (; j < c; j += LANES) {
    const double xv = x[__builtin_nontemporal_load(cols + j)];
    const VT *v = v0 + j;
    a[0] += (double)__builtin_nontemporal_load(v) * xv;
    if (n > 1) a[1] += (double)__builtin_nontemporal_load(v + c) * xv;
    if (n > 2) a[2] += (double)__builtin_nontemporal_load(v + 2 * c) * xv;
    if (n > 3) a[3] += (double)__builtin_nontemporal_load(v + 3 * c) * xv;
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<LANES>(a[q]);
}
