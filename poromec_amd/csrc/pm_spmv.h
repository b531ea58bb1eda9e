// pm_spmv.h -- the node-block SpMV body shared by the single-system and the
// batched kernels, with its tunable shape.
//
// A node block is the <= 4 rows of one free node; they share one sorted
// column list (int32) and store their values row by row.  LANES lanes of a
// wavefront (64 or 32) work on one block, so a wavefront has 64/LANES blocks
// in flight; each lane takes columns lane, lane + LANES, ... and issues UNR of
// them back to back before it uses any, which raises the number of bytes in
// flight per wave (the SpMV is bound by them, not by instruction issue).
// Each lane carries the partial sums of every row of its block; a shuffle
// butterfly over the LANES lanes finishes them.
#pragma once
#include "pm_device.h"

// (LANES, UNR) of each SpMV shape the tuner may pick
#define PM_SPMV_NCFG 6
static constexpr int pm_spmv_lanes[PM_SPMV_NCFG] = {64, 64, 64, 64, 32, 32};
static constexpr int pm_spmv_unr[PM_SPMV_NCFG] = {1, 2, 3, 5, 4, 8};

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

// partial row sums of node block (cp, c, vp, n) held by this lane.  Each
// pass covers UNR * LANES columns: every lane issues its UNR index loads and
// all value loads of the pass (masked past the end of the block) before the
// x gathers and the products, so a wave keeps a whole block's streams in
// flight instead of one 64-column slice at a time.
template <typename VT, int LANES, int UNR>
__device__ __forceinline__ void pm_spmv_block(int lane, int64_t cp, int c, int64_t vp, int n,
                                              const int *__restrict__ ncol,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ x, double a[4]) {
  const int *cols = ncol + cp;
  const VT *v0 = val + vp;
  for (int j0 = lane; j0 < c; j0 += UNR * LANES) {
    int col[UNR];
    VT w[4][UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int j = j0 + u * LANES;
      const bool ok = j < c;
      col[u] = ok ? __builtin_nontemporal_load(cols + j) : 0;
      const VT *v = v0 + j;
      w[0][u] = ok ? __builtin_nontemporal_load(v) : VT(0);
      w[1][u] = (ok && n > 1) ? __builtin_nontemporal_load(v + c) : VT(0);
      w[2][u] = (ok && n > 2) ? __builtin_nontemporal_load(v + 2 * c) : VT(0);
      w[3][u] = (ok && n > 3) ? __builtin_nontemporal_load(v + 3 * c) : VT(0);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) xv[u] = (j0 + u * LANES < c) ? x[col[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w[0][u] * xv[u];
      a[1] += (double)w[1][u] * xv[u];
      a[2] += (double)w[2][u] * xv[u];
      a[3] += (double)w[3][u] * xv[u];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<LANES>(a[q]);
}
