// pm_spmv.h -- the node-block SpMV body shared by the single-system and the
// batched kernels, with its tunable shape.
//
// A node block is the <= 4 rows of one free node; they share one sorted
// column list (int32) and store their values row by row.  LANES lanes of a
// wavefront (64 or 32) work on one block, so a wavefront has 64/LANES blocks
// in flight; each lane takes columns lane, lane + LANES, ... and issues UNR of
// them back to back before it uses any, which raises the number of bytes in
// flight per wave (the SpMV is bound by them, not by instruction issue).
// Each lane carries the partial sums of every row of its block; DPP moves
// over the LANES lanes finish them.  Column indices come as runs of
// consecutive equations, expanded into LDS per block.
#pragma once
#include "pm_device.h"

// (LANES, UNR) of each SpMV shape the tuner may pick
#define PM_SPMV_NCFG 6
static constexpr int pm_spmv_lanes[PM_SPMV_NCFG] = {64, 64, 64, 64, 32, 32};
static constexpr int pm_spmv_unr[PM_SPMV_NCFG] = {1, 2, 3, 5, 4, 8};

// columns of one node block held in LDS when it comes as runs
#define PM_SPMV_MAXC 512

struct PmBlockMeta {
  int64_t cp, vp, rp;
  int c, n, nrun;
};

// metadata of node block k (an empty block past the end k1)
__device__ __forceinline__ PmBlockMeta pm_block_meta(int64_t k, int64_t k1,
                                                     const int64_t *__restrict__ colptr,
                                                     const int64_t *__restrict__ valptr,
                                                     const int *__restrict__ nr,
                                                     const int64_t *__restrict__ runptr) {
  PmBlockMeta m{0, 0, 0, 0, 0, 0};
  if (k < k1) {
    m.cp = colptr[k];
    m.c = (int)(colptr[k + 1] - m.cp);
    m.vp = valptr[k];
    m.n = nr[k];
    if (runptr) {
      m.rp = runptr[k];
      m.nrun = (int)(runptr[k + 1] - m.rp);
    }
  }
  return m;
}

__device__ __forceinline__ void pm_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Expands the block's runs of consecutive columns (first << 6 | length - 1)
// into scol[0 .. c): each lane takes one run, an inclusive scan over the
// group gives its start.  Reading 4 bytes per run instead of per column cuts
// the index stream about ninefold on these meshes (runs of ~9 columns).
template <int LANES>
__device__ __forceinline__ void pm_expand_runs(int lane, const unsigned int *__restrict__ rr,
                                               int nrun, int *scol) {
  int carry = 0;
  for (int r0 = 0; r0 < nrun; r0 += LANES) {
    const int r = r0 + lane;
    const unsigned int w = r < nrun ? __builtin_nontemporal_load(rr + r) : 0u;
    const int len = r < nrun ? (int)(w & 63u) + 1 : 0;
    const int first = (int)(w >> 6);
    int incl = len;
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1) {
      const int t = __shfl_up(incl, off, LANES);
      if (lane >= off) incl += t;
    }
    const int start = carry + incl - len;
    for (int t = 0; t < len; t++) scol[start + t] = first + t;
    carry += __shfl(incl, LANES - 1, LANES);
  }
  pm_wave_sync();
}

// one DPP move of a double (two 32-bit halves); lanes outside row_mask get 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double pm_dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over each group of LANES (32 or 64) lanes with DPP moves instead of
// LDS permutes; the sum is valid in the group's last lane (31 / 63)
template <int LANES>
__device__ __forceinline__ double pm_group_sum(double v) {
  v += pm_dpp<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += pm_dpp<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += pm_dpp<0x124, 0xf>(v);  // row_ror:4
  v += pm_dpp<0x128, 0xf>(v);  // row_ror:8  -> every lane holds its row of 16
  v += pm_dpp<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3 hold rows 0+1, 2+3
  if (LANES == 64) v += pm_dpp<0x143, 0xc>(v);  // row_bcast:31 -> row 3 holds all
  return v;
}

// partial row sums of node block (cp, c, vp, n), complete in the group's
// last lane.  Each pass covers UNR * LANES columns: every lane issues its UNR
// index loads and all value loads of the pass before the x gathers and the
// products, so a wave keeps a whole block's streams in flight instead of one
// 64-column slice at a time.  Past the end of the block (or of its rows) a
// lane loads the block's last valid element again and multiplies by zero:
// straight-line code, no exec-mask branches between the loads.
template <typename VT, int LANES, int UNR>
__device__ __forceinline__ void pm_spmv_block(int lane, const int *cols, int c, int64_t vp, int n,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ x, double a[4]) {
  const VT *v0 = val + vp;
  const int64_t rs1 = n > 1 ? c : 0, rs2 = n > 2 ? 2 * (int64_t)c : 0, rs3 = n > 3 ? 3 * (int64_t)c : 0;
  const double m1 = n > 1 ? 1.0 : 0.0, m2 = n > 2 ? 1.0 : 0.0, m3 = n > 3 ? 1.0 : 0.0;
  for (int base = 0; base < c; base += UNR * LANES) {  // same trip count in the group
    const int j0 = base + lane;
    int col[UNR];
    VT w[4][UNR];
    double keep[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int j = j0 + u * LANES;
      const int jj = j < c ? j : c - 1;
      keep[u] = j < c ? 1.0 : 0.0;
      col[u] = cols[jj];
      const VT *v = v0 + jj;
      w[0][u] = __builtin_nontemporal_load(v);
      w[1][u] = __builtin_nontemporal_load(v + rs1);
      w[2][u] = __builtin_nontemporal_load(v + rs2);
      w[3][u] = __builtin_nontemporal_load(v + rs3);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) xv[u] = x[col[u]] * keep[u];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w[0][u] * xv[u];
      a[1] += (double)w[1][u] * (xv[u] * m1);
      a[2] += (double)w[2][u] * (xv[u] * m2);
      a[3] += (double)w[3][u] * (xv[u] * m3);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<LANES>(a[q]);
}
