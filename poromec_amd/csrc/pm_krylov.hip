// pm_krylov.hip -- sparse matrix-vector product over node blocks and the
// Krylov solvers of solv_pm (Solv.f:27-780) with device-resident scalars.
//
// Each solver follows the reference's recurrences and stop tests:
//   pcg        Iters.f:299-595    stop sqrt|r.z| < tol sqrt|b.M^-1 b|
//   rpsqmr     Iters.f:6100-6436  stop ||r|| < tol ||b||, GMRES on maxit
//   pbicgstab  Iters.f:2817-3105  stop ||r|| < tol ||b||, breakdown 1e-32
//   pminres    Iters.f:3987-4302  stop recurrence ||r|| < tol ||M^-1/2 b||
//   gmres2     Iters.f:1411-1775  MGS with one refinement, Givens (sym_ortho2)
// The vector updates of an iteration are fused into few passes; every dot
// product is a deterministic two-level reduction whose last workgroup also
// computes the next scalars, so the host only polls a done flag.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "pm_device.h"
#include "pm_functors.h"
#include "pm_spmv.h"
#include <vector>

#define PM_SPMV_MAX_GRID 8192

// ---------------------------------------------------------------------------
// SpMV over node blocks (shape (LANES, UNR) from the tuner, pm_spmv.h);
// optional fused Jacobi scaling and dots y.w, y.y into workgroup partials.
// ---------------------------------------------------------------------------
template <typename VT, int LANES, int UNR, int NV, bool SCALE_M>
__global__ __launch_bounds__(PM_WG) void k_spmv(
    int nrn, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int *__restrict__ ncol, const VT *__restrict__ val, const double *__restrict__ x,
    double *__restrict__ y, const double *__restrict__ w, const double *__restrict__ m,
    double *partials, const PmScalars *S) {
  if (S != nullptr && S->done) return;
  constexpr int GROUPS = PM_WG / LANES;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t k0, k1;
  pm_chunk(nrn, lb, gridDim.x, k0, k1);
  double d0 = 0.0, d1 = 0.0;
  // the next block's metadata is loaded while this block streams
  PmBlockMeta nx = pm_block_meta(k0 + grp, k1, colptr, valptr, nr);
  for (int64_t k = k0 + grp; k < k1; k += GROUPS) {
    const PmBlockMeta cur = pm_uniform<LANES>(nx);
    nx = pm_block_meta(k + GROUPS, k1, colptr, valptr, nr);
    const int n = cur.n;
    // the row data the block's epilogue needs, loaded before the column
    // stream so that its latency hides behind it
    const int f = fr[k];
    double wq[4] = {0.0, 0.0, 0.0, 0.0}, mq[4] = {1.0, 1.0, 1.0, 1.0};
    if (lane == LANES - 1) {
      for (int q = 0; q < n; q++) {
        if (NV >= 1) wq[q] = w[f + q];
        if (SCALE_M) mq[q] = m[f + q];
      }
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (LANES == 64 && sizeof(VT) == 4)
      pm_spmv_block64<VT, UNR>(lane, cur, ncol, val, xr, a);
    else
      pm_spmv_block<VT, LANES, UNR>(lane, cur, ncol, val, x, a);
    if (lane == LANES - 1) {
      for (int q = 0; q < n; q++) {
        double yq = a[q];
        if (SCALE_M) yq *= mq[q];
        y[f + q] = yq;
        if (NV >= 1) d0 += yq * wq[q];
        if (NV >= 2) d1 += yq * yq;
      }
    }
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, partials);
  }
}

// The same product with each block's first-pass column indices loaded one
// block ahead (pm_spmv_block64_pf).  A wavefront's blocks k, k + GROUPS, ...
// are wave-uniform (grp via readfirstlane), so their metadata loads are
// scalar loads, two blocks ahead.
template <typename VT, int UNR, int NV, bool SCALE_M, bool SWEEP = false>
__global__ __launch_bounds__(PM_WG) void k_spmv_pf(
    int nrn, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int *__restrict__ ncol, const VT *__restrict__ val, const double *__restrict__ x,
    double *__restrict__ y, const double *__restrict__ w, const double *__restrict__ m,
    double *partials, const PmScalars *S) {
  if (S != nullptr && S->done) return;
  constexpr int GROUPS = PM_WG / 64;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % 64;
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t k0, k1, kfirst, kstep;
  if (SWEEP && (gridDim.x & 7) == 0) {  // as in k_spmv_ch
    const int W = (int)(gridDim.x >> 3);
    pm_chunk(nrn, lb / W, 8, k0, k1);
    kfirst = k0 + (int64_t)(lb % W) * GROUPS + grp;
    kstep = (int64_t)W * GROUPS;
  } else {
    pm_chunk(nrn, lb, gridDim.x, k0, k1);
    kfirst = k0 + grp;
    kstep = GROUPS;
  }
  double d0 = 0.0, d1 = 0.0;
  PmBlockMeta cur = pm_uniform<64>(pm_block_meta(kfirst, k1, colptr, valptr, nr));
  PmBlockMeta nx = pm_uniform<64>(pm_block_meta(kfirst + kstep, k1, colptr, valptr, nr));
  unsigned int col[UNR];
  {
    const __amdgpu_buffer_rsrc_t c0 = pm_rsrc(ncol + cur.cp, (unsigned int)cur.c * 4u);
#pragma unroll
    for (int u = 0; u < UNR; u++)
      col[u] = __builtin_amdgcn_raw_buffer_load_b32(c0, (unsigned int)(lane + u * 64) * 4u, 0, 2);
  }
  for (int64_t k = kfirst; k < k1; k += kstep) {
    const PmBlockMeta nn = pm_uniform<64>(pm_block_meta(k + 2 * kstep, k1, colptr, valptr, nr));
    const int n = cur.n;
    const int f = fr[k];
    double wq[4] = {0.0, 0.0, 0.0, 0.0}, mq[4] = {1.0, 1.0, 1.0, 1.0};
    if (lane == 63) {
      for (int q = 0; q < n; q++) {
        if (NV >= 1) wq[q] = w[f + q];
        if (SCALE_M) mq[q] = m[f + q];
      }
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    pm_spmv_block64_pf<VT, UNR>(lane, cur, nx, col, ncol, val, xr, a);
    if (lane == 63) {
      for (int q = 0; q < n; q++) {
        double yq = a[q];
        if (SCALE_M) yq *= mq[q];
        y[f + q] = yq;
        if (NV >= 1) d0 += yq * wq[q];
        if (NV >= 2) d1 += yq * yq;
      }
    }
    cur = nx;
    nx = nn;
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, partials);
  }
}

// The product with one run of <= 4 consecutive columns per lane (chunk[],
// chptr[]): a lane loads its run's 32-bit descriptor, the run's 4 values of
// each row as one 16-byte load (two for FP64), and x at the run's columns.
// That is ~3.1 columns per lane for one index word instead of one, and half
// the load instructions per column.  A lane's position in the block's
// column list is the sum of the runs before it, taken from two ballots of
// the length bits (mbcnt), with no LDS.  The descriptors of a block's first
// pass are loaded one block ahead, as k_spmv_pf does with column indices.
// Loads past a run read the next columns' values (or the 4 values of padding
// after the array) and are multiplied by zero; rows past the block's count
// accumulate into sums that are not stored.
// FP32 values: at least 6 waves per SIMD (two runs per lane then fit in 80
// VGPRs without spills, against 101 and 4 waves unbounded)
// SWEEP: the workgroups of an XCD sweep its eighth of the blocks together,
// wave slot s taking blocks s, s + S, s + 2S, ... (S slots in the XCD), so
// the x the XCD gathers at a time spans about the matrix band instead of its
// whole eighth of x; otherwise each workgroup streams a contiguous chunk.
template <typename VT, int UNR, int NV, bool SCALE_M, bool SWEEP = false>
__global__ __launch_bounds__(PM_WG, sizeof(VT) == 4 ? 6 : 1) void k_spmv_ch(
    int nrn, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int64_t *__restrict__ chptr, const unsigned int *__restrict__ chunk,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    const double *__restrict__ w, const double *__restrict__ m, double *partials,
    const PmScalars *S) {
  if (S != nullptr && S->done) return;
  constexpr int GROUPS = PM_WG / 64;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % 64;
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t k0, k1, kfirst, kstep;
  if (SWEEP && (gridDim.x & 7) == 0) {
    const int W = (int)(gridDim.x >> 3);
    pm_chunk(nrn, lb / W, 8, k0, k1);
    kfirst = k0 + (int64_t)(lb % W) * GROUPS + grp;
    kstep = (int64_t)W * GROUPS;
  } else {
    pm_chunk(nrn, lb, gridDim.x, k0, k1);
    kfirst = k0 + grp;
    kstep = GROUPS;
  }
  double d0 = 0.0, d1 = 0.0;
  PmRunMeta cur = pm_run_meta(kfirst, k1, colptr, valptr, nr, chptr);
  PmRunMeta nx = pm_run_meta(kfirst + kstep, k1, colptr, valptr, nr, chptr);
  unsigned int dsc[UNR];
  {
    const __amdgpu_buffer_rsrc_t h0 = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
#pragma unroll
    for (int u = 0; u < UNR; u++)
      dsc[u] = __builtin_amdgcn_raw_buffer_load_b32(h0, (unsigned int)(lane + u * 64) * 4u, 0, 2);
  }
  for (int64_t k = kfirst; k < k1; k += kstep) {
    const PmRunMeta nn = pm_run_meta(k + 2 * kstep, k1, colptr, valptr, nr, chptr);
    const int n = cur.n;
    const int f = fr[k];
    double wq[4] = {0.0, 0.0, 0.0, 0.0}, mq[4] = {1.0, 1.0, 1.0, 1.0};
    if (lane == 63) {
      for (int q = 0; q < n; q++) {
        if (NV >= 1) wq[q] = w[f + q];
        if (SCALE_M) mq[q] = m[f + q];
      }
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const __amdgpu_buffer_rsrc_t vr = pm_rsrc(
        val + cur.vp, ((unsigned int)n * (unsigned int)cur.c + 3u) * (unsigned int)sizeof(VT));
    const unsigned int rowb = (unsigned int)cur.c * (unsigned int)sizeof(VT);
    const __amdgpu_buffer_rsrc_t hr = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
    const __amdgpu_buffer_rsrc_t hrn = pm_rsrc(chunk + nx.hp, (unsigned int)nx.nh * 4u);
    unsigned int dn[UNR];
    unsigned int base = pm_run_pass<VT, UNR, true>(lane, 0, cur, 0u, dsc, vr, rowb, xr, hrn, dn, a);
    for (int h0 = UNR * 64; h0 < cur.nh; h0 += UNR * 64) {
      unsigned int dp[UNR], unused[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++)
        dp[u] = __builtin_amdgcn_raw_buffer_load_b32(hr, (unsigned int)(h0 + lane + u * 64) * 4u, 0, 2);
      base = pm_run_pass<VT, UNR, false>(lane, h0, cur, base, dp, vr, rowb, xr, hrn, unused, a);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) dsc[u] = dn[u];
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = pm_group_sum<64>(a[q]);
    if (lane == 63) {
      for (int q = 0; q < n; q++) {
        double yq = a[q];
        if (SCALE_M) yq *= mq[q];
        y[f + q] = yq;
        if (NV >= 1) d0 += yq * wq[q];
        if (NV >= 2) d1 += yq * yq;
      }
    }
    cur = nx;
    nx = nn;
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, partials);
  }
}

// k_spmv_ch with the pass width chosen per block: blocks of <= 64 runs (the
// mid-edge nodes') take one run per lane, wider blocks (the vertices', ~77
// runs at 87^3) two, so neither needs a second round trip; two units of the
// next block's descriptors are always loaded ahead.  XCD sweep mapping.
template <typename VT, int NV, bool SCALE_M>
__global__ __launch_bounds__(PM_WG, sizeof(VT) == 4 ? 6 : 1) void k_spmv_cha(
    int nrn, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int64_t *__restrict__ chptr, const unsigned int *__restrict__ chunk,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    const double *__restrict__ w, const double *__restrict__ m, double *partials,
    const PmScalars *S) {
  if (S != nullptr && S->done) return;
  constexpr int GROUPS = PM_WG / 64;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % 64;
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t k0, k1, kfirst, kstep;
  if ((gridDim.x & 7) == 0) {
    const int W = (int)(gridDim.x >> 3);
    pm_chunk(nrn, lb / W, 8, k0, k1);
    kfirst = k0 + (int64_t)(lb % W) * GROUPS + grp;
    kstep = (int64_t)W * GROUPS;
  } else {
    pm_chunk(nrn, lb, gridDim.x, k0, k1);
    kfirst = k0 + grp;
    kstep = GROUPS;
  }
  double d0 = 0.0, d1 = 0.0;
  PmRunMeta cur = pm_run_meta(kfirst, k1, colptr, valptr, nr, chptr);
  PmRunMeta nx = pm_run_meta(kfirst + kstep, k1, colptr, valptr, nr, chptr);
  unsigned int dsc[2];
  {
    const __amdgpu_buffer_rsrc_t h0 = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
#pragma unroll
    for (int u = 0; u < 2; u++)
      dsc[u] = __builtin_amdgcn_raw_buffer_load_b32(h0, (unsigned int)(lane + u * 64) * 4u, 0, 2);
  }
  for (int64_t k = kfirst; k < k1; k += kstep) {
    const PmRunMeta nn = pm_run_meta(k + 2 * kstep, k1, colptr, valptr, nr, chptr);
    const int n = cur.n;
    const int f = fr[k];
    // the rows' w and m at wave-uniform addresses (scalar loads)
    double wq[4] = {0.0, 0.0, 0.0, 0.0}, mq[4] = {1.0, 1.0, 1.0, 1.0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (NV >= 1 && q < n) wq[q] = w[f + q];
      if (SCALE_M && q < n) mq[q] = m[f + q];
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const __amdgpu_buffer_rsrc_t vr = pm_rsrc(
        val + cur.vp, ((unsigned int)n * (unsigned int)cur.c + 3u) * (unsigned int)sizeof(VT));
    const unsigned int rowb = (unsigned int)cur.c * (unsigned int)sizeof(VT);
    const __amdgpu_buffer_rsrc_t hr = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
    const __amdgpu_buffer_rsrc_t hrn = pm_rsrc(chunk + nx.hp, (unsigned int)nx.nh * 4u);
    unsigned int dn[2];
    unsigned int base;
    if (cur.nh > 64) {
      base = pm_run_pass<VT, 2, true, 2>(lane, 0, cur, 0u, dsc, vr, rowb, xr, hrn, dn, a);
      for (int h0 = 128; h0 < cur.nh; h0 += 128) {
        unsigned int dp[2], unused[2];
#pragma unroll
        for (int u = 0; u < 2; u++)
          dp[u] = __builtin_amdgcn_raw_buffer_load_b32(hr, (unsigned int)(h0 + lane + u * 64) * 4u, 0, 2);
        base = pm_run_pass<VT, 2, false>(lane, h0, cur, base, dp, vr, rowb, xr, hrn, unused, a);
      }
    } else {
      const unsigned int d1[1] = {dsc[0]};
      base = pm_run_pass<VT, 1, true, 2>(lane, 0, cur, 0u, d1, vr, rowb, xr, hrn, dn, a);
    }
    (void)base;
    dsc[0] = dn[0];
    dsc[1] = dn[1];
    // the four rows' sums together; lanes 0-3 hold rows 0, 2, 1, 3
    const double sr = pm_rows4_sum<64>(lane, a);
    const int qr = pm_row_of(lane);
    if (lane < 4 && qr < n) {
      double wl = wq[0], ml = mq[0];
#pragma unroll
      for (int q = 1; q < 4; q++)
        if (qr == q) {
          wl = wq[q];
          ml = mq[q];
        }
      double yq = sr;
      if (SCALE_M) yq *= ml;
      y[f + qr] = yq;
      if (NV >= 1) d0 += yq * wl;
      if (NV >= 2) d1 += yq * yq;
    }
    cur = nx;
    nx = nn;
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, partials);
  }
}

// Runs of <= 4 consecutive columns with LPB lanes per block: a wavefront
// takes G = 64 / LPB consecutive blocks, one per group of LPB lanes, and
// each lane U runs per pass.  Shape 15 is (LPB, U) = (32, 1); (16, 2), four
// blocks per wavefront, measured 1.8x slower on the Q1/Q1 128^3 box and is
// not offered.  Blocks of <= LPB * U runs -- the
// Q1/Q1 element's vertices have 27 -- take one pass and keep 27 of 32 run
// slots busy instead of 27 of 64; wider blocks take more passes.  The G
// blocks' values and descriptors lie one after the other, so one resource
// based at the first block's covers them all and each group adds its
// block's offset; rows past a block's count read the next block's values
// (or zero past the last) and are never written.  The next group of
// blocks' descriptors are loaded during the first pass.  XCD sweep mapping
// over the groups of blocks.
template <int G>
struct PmBlocksMeta {
  PmRunMeta b[G];
};

template <int G>
__device__ __forceinline__ PmBlocksMeta<G> pm_blocks_meta(int64_t p, int64_t p1, int nrn,
                                                          const int64_t *__restrict__ colptr,
                                                          const int64_t *__restrict__ valptr,
                                                          const int *__restrict__ nr,
                                                          const int64_t *__restrict__ chptr) {
  PmBlocksMeta<G> m;
  const int64_t k = p < p1 ? (int64_t)G * p : (int64_t)nrn;  // past the range: empty blocks
#pragma unroll
  for (int j = 0; j < G; j++) m.b[j] = pm_run_meta(k + j, nrn, colptr, valptr, nr, chptr);
  return m;
}

// the block of lane group gi among the G scalar metadata sets
template <int G>
__device__ __forceinline__ PmRunMeta pm_pick(const PmBlocksMeta<G> &m, int gi) {
  PmRunMeta r = m.b[0];
#pragma unroll
  for (int j = 1; j < G; j++)
    if (gi == j) r = m.b[j];
  return r;
}

template <typename VT, int LPB, int U, int NV, bool SCALE_M>
__global__ __launch_bounds__(PM_WG, sizeof(VT) == 4 ? 6 : 1) void k_spmv_chg(
    int nrn, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int64_t *__restrict__ chptr, const unsigned int *__restrict__ chunk,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    const double *__restrict__ w, const double *__restrict__ m, double *partials,
    const PmScalars *S) {
  if (S != nullptr && S->done) return;
  constexpr int G = 64 / LPB;  // blocks per wavefront
  constexpr int GROUPS = PM_WG / 64;
  // FP32 values stream with the non-temporal hint; FP64 values (two 16-byte
  // loads per row and run, into the same cache lines) without it: at 87^3 the
  // hint cost the FP64 product 13 % (3.71 against 3.23 ms) and left FP32 within 1 %
  constexpr int NT = sizeof(VT) == 4 ? 2 : 0;
  constexpr int VPR = sizeof(VT) == 4 ? 1 : 2;  // 16-byte loads per row and run
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % 64;
  const int gi = lane / LPB, gl = lane % LPB;
  const unsigned long long gmask = (LPB == 64 ? ~0ull : ((1ull << LPB) - 1ull)) << (gi * LPB);
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  const int64_t nitem = ((int64_t)nrn + G - 1) / G;
  int64_t p0, p1, pfirst, pstep;
  if ((gridDim.x & 7) == 0) {
    const int W = (int)(gridDim.x >> 3);
    pm_chunk(nitem, lb / W, 8, p0, p1);
    pfirst = p0 + (int64_t)(lb % W) * GROUPS + grp;
    pstep = (int64_t)W * GROUPS;
  } else {
    pm_chunk(nitem, lb, gridDim.x, p0, p1);
    pfirst = p0 + grp;
    pstep = GROUPS;
  }
  double d0 = 0.0, d1 = 0.0;
  PmBlocksMeta<G> cur = pm_blocks_meta<G>(pfirst, p1, nrn, colptr, valptr, nr, chptr);
  PmBlocksMeta<G> nx = pm_blocks_meta<G>(pfirst + pstep, p1, nrn, colptr, valptr, nr, chptr);
  auto desc_bytes = [](const PmBlocksMeta<G> &mm) {
    int nh = 0;
#pragma unroll
    for (int j = 0; j < G; j++) nh += mm.b[j].nh;
    return (unsigned int)nh * 4u;
  };
  unsigned int dsc[U];
  {
    const __amdgpu_buffer_rsrc_t h0 = pm_rsrc(chunk + cur.b[0].hp, desc_bytes(cur));
    const unsigned int ho = (unsigned int)(pm_pick<G>(cur, gi).hp - cur.b[0].hp);
#pragma unroll
    for (int u = 0; u < U; u++)
      dsc[u] = __builtin_amdgcn_raw_buffer_load_b32(h0, (ho + (unsigned int)(u * LPB + gl)) * 4u,
                                                    0, NT);
  }
  for (int64_t p = pfirst; p < p1; p += pstep) {
    const PmBlocksMeta<G> nn = pm_blocks_meta<G>(p + 2 * pstep, p1, nrn, colptr, valptr, nr, chptr);
    const int64_t k = (int64_t)G * p + gi;
    const PmRunMeta me = pm_pick<G>(cur, gi);
    const int n = me.n, nh = me.nh;
    const int f = n > 0 ? fr[k] : 0;
    // w and m of the rows this lane writes: one row per lane 0-3 of the group
    // (transposed reduction, 32 or 64 lanes), all rows in its last lane
    // otherwise
    double wq[4] = {0.0, 0.0, 0.0, 0.0}, mq[4] = {1.0, 1.0, 1.0, 1.0};
    if constexpr (LPB == 32 || LPB == 64) {
      const int qr = pm_row_of(gl);
      if (gl < 4 && qr < n) {
        if (NV >= 1) wq[0] = w[f + qr];
        if (SCALE_M) mq[0] = m[f + qr];
      }
    } else if (gl == LPB - 1) {
      for (int q = 0; q < n; q++) {
        if (NV >= 1) wq[q] = w[f + q];
        if (SCALE_M) mq[q] = m[f + q];
      }
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    unsigned int nvals = 3u;
    int passes = 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
      nvals += (unsigned int)(cur.b[j].n * cur.b[j].c);
      passes = max(passes, cur.b[j].nh);  // wave-uniform trip count
    }
    const __amdgpu_buffer_rsrc_t vr = pm_rsrc(val + cur.b[0].vp, nvals * (unsigned int)sizeof(VT));
    const __amdgpu_buffer_rsrc_t hr = pm_rsrc(chunk + cur.b[0].hp, desc_bytes(cur));
    const __amdgpu_buffer_rsrc_t hrn = pm_rsrc(chunk + nx.b[0].hp, desc_bytes(nx));
    const unsigned int voff = (unsigned int)(me.vp - cur.b[0].vp);
    const unsigned int hoff = (unsigned int)(me.hp - cur.b[0].hp);
    const unsigned int hoffn = (unsigned int)(pm_pick<G>(nx, gi).hp - nx.b[0].hp);
    const unsigned int rowb = (unsigned int)me.c * (unsigned int)sizeof(VT);
    unsigned int base = 0, dnext[U];
#pragma unroll
    for (int u = 0; u < U; u++) dnext[u] = 0;
    for (int h0 = 0; h0 < passes; h0 += LPB * U) {
      unsigned int d[U], off[U];
      int len[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const unsigned int r = (unsigned int)(h0 + u * LPB + gl);
        d[u] = h0 == 0 ? dsc[u] : __builtin_amdgcn_raw_buffer_load_b32(hr, (hoff + r) * 4u, 0, NT);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool valid = h0 + u * LPB + gl < nh;
        const unsigned int l1 = d[u] & 3u;
        const unsigned long long bv = __ballot(valid) & gmask;
        const unsigned long long b0 = __ballot(valid && (l1 & 1u)) & gmask;
        const unsigned long long b1 = __ballot(valid && (l1 & 2u)) & gmask;
        off[u] = base + pm_lanes_below(bv) + pm_lanes_below(b0) + 2u * pm_lanes_below(b1);
        len[u] = valid ? (int)l1 + 1 : 0;
        base += (unsigned int)(__popcll(bv) + __popcll(b0) + 2 * __popcll(b1));
      }
      typedef unsigned int u4 __attribute__((ext_vector_type(4)));
      u4 wv[U][4][VPR];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const unsigned int o = (voff + off[u]) * (unsigned int)sizeof(VT);
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int hh = 0; hh < VPR; hh++)
            wv[u][q][hh] = __builtin_bit_cast(
                u4, __builtin_amdgcn_raw_buffer_load_b128(vr, o + (unsigned int)q * rowb + 16u * hh,
                                                          0, NT));
      }
      double xv[U][4];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const unsigned int col = d[u] >> 2;
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
          const double xc = pm_bload<double, 0>(xr, (col + (unsigned int)kk) * 8u);
          xv[u][kk] = kk < len[u] ? xc : 0.0;
        }
      }
      if (h0 == 0)  // the next blocks' first descriptors, behind this pass's loads
#pragma unroll
        for (int u = 0; u < U; u++)
          dnext[u] = __builtin_amdgcn_raw_buffer_load_b32(
              hrn, (hoffn + (unsigned int)(u * LPB + gl)) * 4u, 0, NT);
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          VT v[4];
          __builtin_memcpy(v, &wv[u][q][0], sizeof(v));
#pragma unroll
          for (int kk = 0; kk < 4; kk++) a[q] += (double)v[kk] * xv[u][kk];
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) dsc[u] = dnext[u];
    if constexpr (LPB == 32 || LPB == 64) {
      // the four rows' sums together; the group's lanes 0-3 hold rows 0, 2, 1, 3
      const double sr = pm_rows4_sum<LPB>(gl, a);
      const int qr = pm_row_of(gl);
      if (gl < 4 && qr < n) {
        double yq = sr;
        if (SCALE_M) yq *= mq[0];
        y[f + qr] = yq;
        if (NV >= 1) d0 += yq * wq[0];
        if (NV >= 2) d1 += yq * yq;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) a[q] = pm_group_sum<LPB>(a[q]);
      if (gl == LPB - 1) {
        for (int q = 0; q < n; q++) {
          double yq = a[q];
          if (SCALE_M) yq *= mq[q];
          y[f + q] = yq;
          if (NV >= 1) d0 += yq * wq[q];
          if (NV >= 2) d1 += yq * yq;
        }
      }
    }
    cur = nx;
    nx = nn;
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, partials);
  }
}

// Sums the workgroup partials of the kernel before it (fixed order) into
// S->red and runs the epilogue that computes the next solver scalars.
template <int NV, class Epi>
__global__ __launch_bounds__(PM_WG) void k_finish(int nblk, const double *partials, PmScalars *S,
                                                  Epi epi, int ignore_done) {
  if (!ignore_done && S->done) return;
  double out[NV > 0 ? NV : 1];
  pm_sum_partials<NV>(nblk, partials, out);
  if (threadIdx.x == 0) {
    for (int i = 0; i < NV; i++) S->red[i] = out[i];
    epi(S, out);
  }
}

template <int NV, class Epi>
static void launch_finish(pm_ctx *c, int nblk, PmScalars *S, Epi epi, int ignore_done) {
  if (NV == 0 && std::is_same<Epi, NoEpi>::value) return;
  hipLaunchKernelGGL((k_finish<NV, Epi>), dim3(1), dim3(PM_WG), 0, c->stream, nblk,
                     c->d_partials, S, epi, ignore_done);
}

static int spmv_grid(int nrn, int lanes) {
  const int groups = PM_WG / lanes;
  int g = (nrn + groups - 1) / groups;
  if (g < 8) g = 8;
  if (g > PM_SPMV_MAX_GRID) g = PM_SPMV_MAX_GRID;
  return (g + 7) & ~7;
}

template <typename VT, int NV, bool SCALE_M>
static int launch_spmv_shape(pm_ctx *c, int cfg, const VT *val, const double *x, double *y,
                             const double *w, const double *m, const PmScalars *S) {
  const PmDevMatrix &A = c->A;
  const int g = spmv_grid(A.nrn, pm_spmv_lanes[cfg]);
#define PM_SPMV_CASE(id, L, U)                                                                  \
  case id:                                                                                      \
    hipLaunchKernelGGL((k_spmv<VT, L, U, NV, SCALE_M>), dim3(g), dim3(PM_WG), 0, c->stream,   \
                       A.nrn, A.neq, A.fr, A.nr, A.colptr, A.valptr, A.ncol, val, x, y, w, m,   \
                       c->d_partials, S);                                                       \
    break;
#define PM_SPMV_PF_CASE(id, U, SW)                                                              \
  case id:                                                                                      \
    hipLaunchKernelGGL((k_spmv_pf<VT, U, NV, SCALE_M, SW>), dim3(g), dim3(PM_WG), 0, c->stream,\
                       A.nrn, A.neq, A.fr, A.nr, A.colptr, A.valptr, A.ncol, val, x, y, w, m,   \
                       c->d_partials, S);                                                       \
    break;
#define PM_SPMV_CH_CASE(id, U, SW)                                                              \
  case id:                                                                                      \
    hipLaunchKernelGGL((k_spmv_ch<VT, U, NV, SCALE_M, SW>), dim3(g), dim3(PM_WG), 0, c->stream,\
                       A.nrn, A.neq, A.fr, A.nr, A.colptr, A.valptr, A.chptr, A.chunk, val, x,  \
                       y, w, m, c->d_partials, S);                                              \
    break;
  switch (cfg) {
    PM_SPMV_CASE(0, 64, 1)
    PM_SPMV_CASE(1, 64, 2)
    PM_SPMV_CASE(2, 64, 3)
    PM_SPMV_CASE(3, 64, 4)
    PM_SPMV_CASE(4, 64, 5)
    PM_SPMV_PF_CASE(6, 3, false)
    PM_SPMV_PF_CASE(7, 4, false)
    PM_SPMV_PF_CASE(8, 5, false)
    PM_SPMV_PF_CASE(13, 5, true)
    case 14:
      hipLaunchKernelGGL((k_spmv_cha<VT, NV, SCALE_M>), dim3(g), dim3(PM_WG), 0, c->stream, A.nrn,
                         A.neq, A.fr, A.nr, A.colptr, A.valptr, A.chptr, A.chunk, val, x, y, w, m,
                         c->d_partials, S);
      break;
    PM_SPMV_CH_CASE(9, 1, false)
    PM_SPMV_CH_CASE(10, 2, false)
    PM_SPMV_CH_CASE(11, 1, true)
    PM_SPMV_CH_CASE(12, 2, true)
    case 15:
      hipLaunchKernelGGL((k_spmv_chg<VT, 32, 1, NV, SCALE_M>), dim3(g), dim3(PM_WG), 0, c->stream,
                         A.nrn, A.neq, A.fr, A.nr, A.colptr, A.valptr, A.chptr, A.chunk, val, x, y,
                         w, m, c->d_partials, S);
      break;
    default: PM_SPMV_CASE(5, 32, 4)
  }
#undef PM_SPMV_CASE
#undef PM_SPMV_PF_CASE
#undef PM_SPMV_CH_CASE
  return g;
}


// ---------------------------------------------------------------------------
// generic fused vector kernel: body(i, acc) for each row, NV reductions
// ---------------------------------------------------------------------------
// With seq != 0 (reference-order mode, one workgroup) thread 0 alone walks
// the rows in order, so every sum is the sequential left-to-right sum of the
// reference's loops (the other lanes contribute exact zeros).
template <int NV, class Body>
__global__ __launch_bounds__(PM_WG) void k_vec(int64_t n, const PmScalars *S, double *partials,
                                               Body body, int ignore_done, int seq) {
  if (!ignore_done && S->done) return;
  const int lb = pm_xcd_block(blockIdx.x, gridDim.x);
  int64_t a, b;
  pm_chunk(n, lb, gridDim.x, a, b);
  double acc[NV > 0 ? NV : 1];
  for (int i = 0; i < (NV > 0 ? NV : 1); i++) acc[i] = 0.0;
  const PmScalars &s = *S;
  if (seq) {
    if (threadIdx.x == 0)
      for (int64_t i = a; i < b; i++) body(i, s, acc);
  } else {
    for (int64_t i = a + threadIdx.x; i < b; i += PM_WG) body(i, s, acc);
  }
  pm_block_partials<NV>(acc, partials);
}

template <int NV, class Body, class Epi>
static void launch_vec(pm_ctx *c, int64_t n, PmScalars *S, Body body, Epi epi,
                       int ignore_done = 0) {
  const int g = c->ordered ? 1 : pm_grid_for(n);
  hipLaunchKernelGGL((k_vec<NV, Body>), dim3(g), dim3(PM_WG), 0, c->stream, n, S, c->d_partials,
                     body, ignore_done, c->ordered);
  launch_finish<NV>(c, g, S, epi, ignore_done);
}

// ---------------------------------------------------------------------------
// SpMV in the reference's summation order (matvec_csrc_sym_pm, Matvec.f:
// 1504-1595, with the lower CSRC): row i is ad(i) x(i), then its lower
// columns in ascending order, then the transposed terms of the rows below
// it in ascending order -- i.e. diagonal first, then the stored row left to
// right skipping the diagonal.  One thread per row.
// ---------------------------------------------------------------------------
template <typename VT, bool SCALE_M>
__global__ __launch_bounds__(PM_WG) void k_spmv_ref(
    int neq, const int *__restrict__ eq_block, const int *__restrict__ fr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int *__restrict__ ncol, const VT *__restrict__ val, const double *__restrict__ x,
    double *__restrict__ y, const double *__restrict__ m, const PmScalars *S) {
  if (S != nullptr && S->done) return;
  const int row = blockIdx.x * PM_WG + threadIdx.x;
  if (row >= neq) return;
  const int k = eq_block[row];
  const int64_t cp = colptr[k];
  const int c = (int)(colptr[k + 1] - cp);
  const VT *v = val + valptr[k] + (int64_t)(row - fr[k]) * c;
  const int *cols = ncol + cp;
  int jd = 0;
  while (jd < c && cols[jd] != row) jd++;
  double t = jd < c ? (double)v[jd] * x[row] : 0.0;
  for (int j = 0; j < c; j++)
    if (j != jd) t = t + (double)v[j] * x[cols[j]];
  if (SCALE_M) t = t * m[row];
  y[row] = t;
}

template <int NV>
struct SpmvDots {  // the dots fused into the fast SpMV: y.w and y.y
  const double *y, *w;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    const double yi = y[i];
    if (NV >= 1) acc[0] += yi * w[i];
    if (NV >= 2) acc[1] += yi * yi;
  }
};

static void spmv_time_begin(pm_ctx *c) {
  c->tev_open = false;
  if (!c->time_spmv || (c->spmv_seq++ % c->time_spmv) != 0) return;  // every time_spmv-th
  if ((size_t)(2 * c->tev_used + 2) > c->tev.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      c->time_spmv = 0;  // timing is best effort; the solve goes on
      return;
    }
    c->tev.push_back(a);
    c->tev.push_back(b);
    c->tev_kind.push_back(0);
    c->tev_iter.push_back(0);
  }
  (void)hipEventRecord(c->tev[2 * c->tev_used], c->stream);
  c->tev_open = true;
}

static void spmv_time_end(pm_ctx *c, int use32) {
  if (!c->tev_open) return;
  c->tev_open = false;
  (void)hipEventRecord(c->tev[2 * c->tev_used + 1], c->stream);
  c->tev_kind[c->tev_used] = use32 ? 1 : 0;
  c->tev_iter[c->tev_used] = c->cur_iter;
  c->tev_used++;
}

// after a host sync: add the durations of the recorded pairs; with done != 0
// pairs tagged past `iter` ran on a finished solve and are dropped
void pm_spmv_harvest(pm_ctx *c, int done, int iter) {
  for (int i = 0; i < c->tev_used; i++) {
    if (done && c->tev_iter[i] > iter) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->tev[2 * i], c->tev[2 * i + 1]) == hipSuccess) {
      c->spmv_ms[c->tev_kind[i]] += ms;
      c->spmv_n[c->tev_kind[i]]++;
    }
  }
  c->tev_used = 0;
}

template <int NV, bool SCALE_M, class Epi>
static void launch_spmv_raw(pm_ctx *c, int use32, const double *x, double *y, const double *w,
                            const double *m, PmScalars *S, Epi epi) {
  const PmDevMatrix &A = c->A;
  if (c->ordered) {
    const int g = (A.neq + PM_WG - 1) / PM_WG;
    if (use32)
      hipLaunchKernelGGL((k_spmv_ref<float, SCALE_M>), dim3(g), dim3(PM_WG), 0, c->stream,
                         A.neq, A.eq_block, A.fr, A.colptr, A.valptr, A.ncol, A.val32, x, y, m,
                         S);
    else
      hipLaunchKernelGGL((k_spmv_ref<double, SCALE_M>), dim3(g), dim3(PM_WG), 0, c->stream,
                         A.neq, A.eq_block, A.fr, A.colptr, A.valptr, A.ncol, A.val, x, y, m, S);
    if (NV > 0) launch_vec<NV>(c, A.neq, S, SpmvDots<NV>{y, w}, epi, 0);
    return;
  }
  const int cfg = c->spmv_cfg[use32 ? 1 : 0];
  const int g = use32 ? launch_spmv_shape<float, NV, SCALE_M>(c, cfg, A.val32, x, y, w, m, S)
                      : launch_spmv_shape<double, NV, SCALE_M>(c, cfg, A.val, x, y, w, m, S);
  if (NV > 0) launch_finish<NV>(c, g, S, epi, 0);
}

template <int NV, bool SCALE_M, class Epi>
static void launch_spmv(pm_ctx *c, int use32, const double *x, double *y, const double *w,
                        const double *m, PmScalars *S, Epi epi) {
  spmv_time_begin(c);
  launch_spmv_raw<NV, SCALE_M>(c, use32, x, y, w, m, S, epi);
  spmv_time_end(c, use32);
}

int pm_spmv_dev(pm_ctx *c, const double *x, double *y, int use32) {
  launch_spmv<0, false>(c, use32, x, y, nullptr, nullptr, nullptr, NoEpi{});
  PM_CHECK_HIP(hipGetLastError());
  return PM_OK;
}

// the product as the Krylov loops launch it: y = A x with the dot y.x fused
// and reduced by a finisher (what the shape tuner times)
int pm_spmv_fused_dev(pm_ctx *c, const double *x, double *y, int use32) {
  launch_spmv<1, false>(c, use32, x, y, x, nullptr, c->d_scal, NoEpi{});
  PM_CHECK_HIP(hipGetLastError());
  return PM_OK;
}

static int poll_done(pm_ctx *c, int *done, int *iter) {
  PM_CHECK_HIP(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(PmScalars), hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  *done = c->h_scal->done;
  *iter = c->h_scal->iter;
  // PM_PROGRESS_SECONDS=s (opt-in): the solve's iteration and recurrence
  // residual on stderr every s seconds (long solves under a watchdog)
  static const double progress_s = [] {
    const char *e = std::getenv("PM_PROGRESS_SECONDS");
    return e ? std::atof(e) : 0.0;
  }();
  if (progress_s > 0.0) {
    static thread_local auto last = std::chrono::steady_clock::now();
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last).count() >= progress_s) {
      std::fprintf(stderr, "[pm_solve] %d iterations launched, residual %.3e\n", c->cur_iter,
                   c->h_scal->resid);
      std::fflush(stderr);
      last = now;
    }
  }
  if (c->tev_used) pm_spmv_harvest(c, *done, *iter);
  return PM_OK;
}

static int reset_scalars(pm_ctx *c) {
  PmScalars z;
  memset(&z, 0, sizeof(z));
  PM_CHECK_HIP(hipMemcpyAsync(c->d_scal, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return PM_OK;
}

// iterations launched between two polls of the done flag: grows so that the
// host syncs rarely on long solves
static int next_chunk(int launched) {
  if (launched < 32) return 8;
  if (launched < 256) return 32;
  return 64;
}

// ---------------------------------------------------------------------------
// PBiCGSTAB (Iters.f:2817-3105), right Jacobi preconditioning
// ---------------------------------------------------------------------------
struct BcgInit {
  const double *b, *m;
  double *x, *r0, *p, *r, *z;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double bi = b[i];
    x[i] = 0.0;
    r0[i] = bi;
    p[i] = bi;
    r[i] = bi;
    z[i] = bi * m[i];
    acc[0] += bi * bi;
  }
};
struct BcgInitEpi {
  double tol;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->conv = tol * sqrt(r[0]);
    S->breaktol = 1.0e-32 * sqrt(r[0]);
    S->s[S_RR0] = r[0];  // (r, r0) with r = r0 = b
    S->it = 0;
  }
};
struct BcgAlpha {  // after v = A z, red = v . r0
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->it += 1;
    double rr0 = S->s[S_RR0];
    if (sqrt(fabs(rr0)) < S->breaktol) {
      S->done = PM_ERR_BREAKDOWN;
      return;
    }
    S->s[S_ALPHA] = rr0 / r[0];
  }
};
struct BcgS {
  double *x, *r, *z;
  const double *v, *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double al = s.s[S_ALPHA];
    x[i] = x[i] + al * z[i];
    double ri = r[i] - al * v[i];
    r[i] = ri;
    z[i] = ri * m[i];
    acc[0] += ri * ri;
  }
};
struct BcgSEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->resid = sqrt(fabs(r[0]));
    if (sqrt(fabs(r[0])) < S->conv) {
      S->done = 1;
      S->iter = S->it;
    }
  }
};
struct BcgW {  // after t = A z: red = (t . r, t . t)
  __device__ void operator()(PmScalars *S, const double *r) const {
    double w = r[0] / r[1];
    if (fabs(w) < S->breaktol) {
      S->done = PM_ERR_BREAKDOWN;
      return;
    }
    S->s[S_W] = w;
  }
};
struct BcgR {
  double *x, *r;
  const double *z, *t, *r0;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double w = s.s[S_W];
    x[i] = x[i] + w * z[i];
    double ri = r[i] - w * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += ri * r0[i];
  }
};
struct BcgREpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->resid = sqrt(fabs(r[0]));
    if (sqrt(fabs(r[0])) < S->conv) {
      S->done = 1;
      S->iter = S->it;
      return;
    }
    S->s[S_BETA] = (r[1] / S->s[S_RR0]) * (S->s[S_ALPHA] / S->s[S_W]);
    S->s[S_RR0] = r[1];
  }
};
struct BcgP {
  double *p, *z;
  const double *r, *v, *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    double pi = r[i] + s.s[S_BETA] * (p[i] - s.s[S_W] * v[i]);
    p[i] = pi;
    z[i] = pi * m[i];
  }
};

// ---------------------------------------------------------------------------
// PMINRES (Iters.f:3987-4302)
// ---------------------------------------------------------------------------
struct MrConv {
  const double *b, *m;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double zi = b[i] * sqrt(fabs(m[i]));
    acc[0] += zi * zi;
  }
};
struct MrConvEpi {
  double tol;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->conv = tol * sqrt(fabs(r[0]));
    S->it = 0;
  }
};
struct MrStart {  // z = A x already in z
  const double *b, *m, *z;
  double *v, *v0, *w0, *w00, *z0;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double vi = b[i] - z[i];
    v[i] = vi;
    v0[i] = 0.0;
    w0[i] = 0.0;
    w00[i] = 0.0;
    double zi = vi * m[i];
    z0[i] = zi;
    acc[0] += vi * zi;
  }
};
struct MrStartEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    double beta = sqrt(fabs(r[0]));
    S->s[S_NORMR] = beta;
    S->s[S_NETA] = beta;
    S->s[S_BETA] = beta;
    S->s[S_BETAOLD] = beta;
    S->s[S_COLD] = 1.0;
    S->s[S_C] = 1.0;
    S->s[S_SOLD] = 0.0;
    S->s[S_S] = 0.0;
  }
};
struct MrAlpha {  // after p = A z0, red = z0 . p
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->it += 1;
    double beta = S->s[S_BETA];
    double alpha = fabs(r[0]) * (1.0 / (beta * beta));
    S->s[S_ALPHA] = alpha;
    S->s[S_T1] = 1.0 / beta;
    S->s[S_T2] = -alpha / beta;
    S->s[S_T3] = -beta / S->s[S_BETAOLD];
  }
};
struct MrP {
  double *p, *z;
  const double *v, *v0, *m;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double pi = s.s[S_T1] * p[i] + s.s[S_T2] * v[i] + s.s[S_T3] * v0[i];
    p[i] = pi;
    double zi = pi * m[i];
    z[i] = zi;
    acc[0] += zi * pi;
  }
};
__device__ __forceinline__ double pm_sign1(double a) {
  return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0);
}
// sym_ortho (Iters.f:6595-6658), including its (|b| > a) comparison
__device__ __forceinline__ void pm_sym_ortho(double a, double b, double &c, double &s, double &r) {
  double ma = fabs(a), mb = fabs(b), sa = pm_sign1(a), sb = pm_sign1(b);
  if (b == 0.0) {
    s = 0.0;
    r = ma;
    c = (a == 0.0) ? 1.0 : sa;
  } else if (a == 0.0) {
    c = 0.0;
    s = sb;
    r = mb;
  } else if (mb > a) {
    double t = a / b;
    s = sb / sqrt(1.0 + t * t);
    c = s * t;
    r = b / s;
  } else if (ma > mb) {
    double t = b / a;
    c = sa / sqrt(1.0 + t * t);
    s = c * t;
    r = a / c;
  }
}
struct MrQr {
  __device__ void operator()(PmScalars *S, const double *r) const {
    double *s = S->s;
    double beta_old = s[S_BETA];
    double beta = sqrt(fabs(r[0]));
    double alpha = s[S_ALPHA];
    double delta = s[S_C] * alpha - s[S_COLD] * s[S_S] * beta_old;
    double ro2 = s[S_S] * alpha + s[S_COLD] * s[S_C] * beta_old;
    double ro3 = s[S_SOLD] * beta_old;
    s[S_COLD] = s[S_C];
    s[S_SOLD] = s[S_S];
    double c = s[S_C], sn = s[S_S], ro1 = 0.0;
    pm_sym_ortho(delta, beta, c, sn, ro1);
    s[S_C] = c;
    s[S_S] = sn;
    s[S_BETAOLD] = beta_old;
    s[S_BETA] = beta;
    s[S_T1] = 1.0 / (beta_old * ro1);
    s[S_T2] = -ro2 / ro1;
    s[S_T3] = -ro3 / ro1;
    s[S_T4] = c * s[S_NETA];
    s[S_NORMR] = fabs(sn) * s[S_NORMR];
    S->resid = s[S_NORMR];
    if (s[S_NORMR] < S->conv) S->pending = 1;
    else s[S_NETA] = -sn * s[S_NETA];
  }
};
struct MrW {  // w = t1 z0 + t2 w0 + t3 w00 (into w00's buffer), x += t4 w
  double *x, *wout;
  const double *z0, *w0;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    double w = s.s[S_T1] * z0[i] + s.s[S_T2] * w0[i] + s.s[S_T3] * wout[i];
    wout[i] = w;
    x[i] = x[i] + s.s[S_T4] * w;
  }
};

// ---------------------------------------------------------------------------
// GMRES(k) with MGS + one refinement (Iters.f:1411-1775)
// ---------------------------------------------------------------------------
// h (k+1) x k column-major, c, s, e, y live in one device array after PmScalars
struct GmInitNorm {
  const double *b, *m;
  double *x, *g0;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    x[i] = 0.0;
    double gi = b[i] * m[i];
    g0[i] = gi;
    acc[0] += gi * gi;
  }
};
struct GmInitNormEpi {
  double tol;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->conv = tol * sqrt(r[0]);
    S->it = 0;
  }
};
struct GmRes0 {  // g0 = (b - g0) * m, red = g0 . g0
  const double *b, *m;
  double *g0;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double gi = (b[i] - g0[i]) * m[i];
    g0[i] = gi;
    acc[0] += gi * gi;
  }
};
struct GmE0 {
  double *e;
  __device__ void operator()(PmScalars *S, const double *r) const {
    e[0] = sqrt(r[0]);
    S->s[S_E0] = 1.0 / e[0];
    S->ni = 0;
    S->flag = 0;
  }
};
struct GmScale {  // g *= s[S_E0]  (1 / norm)
  double *g;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    g[i] = g[i] * s.s[S_E0];
  }
};
struct GmTau {  // after g_{i+1} = M A g_i: red = (g.g) ; store tau
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->it += 1;
    S->ni += 1;
    S->s[S_TAU] = sqrt(r[0]);
  }
};
struct GmDot {  // red = g . gj
  const double *g, *gj;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    acc[0] += g[i] * gj[i];
  }
};
struct GmSetH {
  double *h;
  int idx, add;
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->s[S_BETA] = r[0];
    if (add) h[idx] = h[idx] + r[0];
    else h[idx] = r[0];
  }
};
struct GmAxpyDot {  // g -= beta gj ; red = g . gnext (gnext may be g itself)
  double *g;
  const double *gj, *gnext;
  __device__ void operator()(int64_t i, const PmScalars &s, double *acc) const {
    double gi = g[i] - s.s[S_BETA] * gj[i];
    g[i] = gi;
    acc[0] += gi * (gnext ? gnext[i] : gi);
  }
};
// Givens update of column i and the convergence test of the cycle
struct GmGivens {
  double *h, *cs, *sn, *e;
  int i, k, last_refine_check;
  __device__ void operator()(PmScalars *S, const double *r) const {
    (void)last_refine_check;
    double norm = sqrt(r[0]);
    const int ld = k + 1;
    h[(size_t)i * ld + i + 1] = norm;
    S->s[S_E0] = 1.0 / norm;
    for (int j = 0; j < i; j++) {
      double a1 = cs[j] * h[(size_t)i * ld + j] + sn[j] * h[(size_t)i * ld + j + 1];
      double a2 = -sn[j] * h[(size_t)i * ld + j] + cs[j] * h[(size_t)i * ld + j + 1];
      h[(size_t)i * ld + j] = a1;
      h[(size_t)i * ld + j + 1] = a2;
    }
    double a = h[(size_t)i * ld + i], b = h[(size_t)i * ld + i + 1];
    double c = 1.0, s = 0.0, rr = a;
    if (b != 0.0) {  // sym_ortho2 (Iters.f:6718-6745)
      if (fabs(b) > fabs(a)) {
        double t = a / b;
        s = 1.0 / sqrt(1.0 + t * t);
        c = s * t;
        rr = b / s;
      } else {
        double t = b / a;
        c = 1.0 / sqrt(1.0 + t * t);
        s = c * t;
        rr = a / c;
      }
    }
    cs[i] = c;
    sn[i] = s;
    h[(size_t)i * ld + i] = rr;
    h[(size_t)i * ld + i + 1] = 0.0;
    e[i + 1] = -s * e[i];
    e[i] = c * e[i];
    S->resid = fabs(e[i + 1]);
    if (fabs(e[i + 1]) <= S->conv) S->flag = 2;  // cycle ends here
  }
};
struct GmRefineCheck {  // red = g.g after the first MGS sweep
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->flag = (sqrt(r[0]) <= 0.25 * S->s[S_TAU]) ? 1 : 0;
  }
};
struct GmBack {  // back substitution by one thread
  double *h, *e, *y;
  int k;
  __device__ void operator()(PmScalars *S, const double *) const {
    const int ni = S->ni, ld = k + 1;
    for (int j = 0; j < ni; j++) y[j] = e[j];
    for (int j = ni - 1; j >= 1; j--) {
      y[j] = y[j] / h[(size_t)j * ld + j];
      double r = y[j];
      for (int i = 0; i < j; i++) y[i] = y[i] - h[(size_t)j * ld + i] * r;
    }
    y[0] = y[0] / h[0];
  }
};
struct GmX {  // x += sum_j y_j g_j, in j order
  double *x;
  const double *g, *y;
  int64_t ldg;
  int ni_max;
  __device__ void operator()(int64_t i, const PmScalars &s, double *) const {
    double xi = x[i];
    for (int j = 0; j < s.ni; j++) xi = xi + y[j] * g[(size_t)j * ldg + i];
    x[i] = xi;
  }
};
struct GmEnd {
  const double *e;
  __device__ void operator()(PmScalars *S, const double *) const {
    S->resid = fabs(e[S->ni]);
    if (fabs(e[S->ni]) <= S->conv) {
      S->done = 1;
      S->iter = S->it;
    }
  }
};

// ---------------------------------------------------------------------------
// drivers
// ---------------------------------------------------------------------------
struct Vecs {
  double *v[12];
};

static int zero_diag_check(pm_ctx *c);

static int run_pcg(pm_ctx *c, int n, const double *b, double *x, const double *m, double tol,
                   int maxit, int use32, pm_solve_info *info) {
  double *r = c->d_wk[0], *z = c->d_wk[1], *p = c->d_wk[2];
  PmScalars *S = c->d_scal;
  PM_TRY(reset_scalars(c));
  launch_vec<1>(c, n, S, PcgInit{b, m, x, r, z, p}, PcgInitEpi{tol});
  int launched = 0, done = 0, iter = 0;
  while (launched < maxit) {
    int chunk = std::min(next_chunk(launched), maxit - launched);
    for (int it = 0; it < chunk; it++) {
      c->cur_iter = launched + it + 1;
      launch_spmv<1, false>(c, use32, p, z, p, nullptr, S, PcgAlpha{});
      (use32 ? info->spmv32 : info->spmv64)++;
      launch_vec<1>(c, n, S, PcgUpd{x, r, z, p, m}, PcgUpdEpi{});
      launch_vec<0>(c, n, S, PcgP{p, z}, FinishEpi{});
    }
    launched += chunk;
    PM_CHECK_HIP(hipGetLastError());
    PM_TRY(poll_done(c, &done, &iter));
    if (done) break;
  }
  if (done < 0) return done;
  if (!done) return PM_ERR_MAXIT;
  info->iters += iter;
  info->resid = c->h_scal->resid;
  return PM_OK;
}

static int run_sqmr(pm_ctx *c, int n, const double *b, double *x, const double *m, double tol,
                    int maxit, int use32, pm_solve_info *info, int *failed) {
  double *t = c->d_wk[0], *r = c->d_wk[1], *q = c->d_wk[2], *d = c->d_wk[3];
  PmScalars *S = c->d_scal;
  *failed = 0;
  PM_TRY(reset_scalars(c));
  launch_vec<2>(c, n, S, SqmrInit{b, m, x, r, q, d}, SqmrInitEpi{tol});
  int launched = 0, done = 0, iter = 0;
  while (launched < maxit) {
    int chunk = std::min(next_chunk(launched), maxit - launched);
    for (int it = 0; it < chunk; it++) {
      c->cur_iter = launched + it + 1;
      launch_spmv<1, false>(c, use32, q, t, q, nullptr, S, SqmrSigma{});
      (use32 ? info->spmv32 : info->spmv64)++;
      launch_vec<2>(c, n, S, SqmrRT{r, t, m}, SqmrRTEpi{});
      launch_vec<0>(c, n, S, SqmrDXQ{d, x, q, t}, SqmrFinishEpi{});
    }
    launched += chunk;
    PM_CHECK_HIP(hipGetLastError());
    PM_TRY(poll_done(c, &done, &iter));
    if (done) break;
  }
  if (done < 0) return done;
  if (!done) {
    *failed = 1;  // the reference sets fsqmr and falls back to GMRES
    info->iters += maxit;
    return PM_OK;
  }
  info->iters += iter;
  info->resid = c->h_scal->resid;
  return PM_OK;
}

static int run_bicgstab(pm_ctx *c, int n, const double *b, double *x, const double *m,
                        double tol, int maxit, int use32, pm_solve_info *info) {
  double *r0 = c->d_wk[0], *p = c->d_wk[1], *r = c->d_wk[2], *z = c->d_wk[3], *v = c->d_wk[4],
         *t = c->d_wk[5];
  PmScalars *S = c->d_scal;
  PM_TRY(reset_scalars(c));
  launch_vec<1>(c, n, S, BcgInit{b, m, x, r0, p, r, z}, BcgInitEpi{tol});
  int launched = 0, done = 0, iter = 0;
  while (launched < maxit) {
    int chunk = std::min(next_chunk(launched), maxit - launched);
    for (int it = 0; it < chunk; it++) {
      c->cur_iter = launched + it + 1;
      launch_spmv<1, false>(c, use32, z, v, r0, nullptr, S, BcgAlpha{});
      launch_vec<1>(c, n, S, BcgS{x, r, z, v, m}, BcgSEpi{});
      launch_spmv<2, false>(c, use32, z, t, r, nullptr, S, BcgW{});
      launch_vec<2>(c, n, S, BcgR{x, r, z, t, r0}, BcgREpi{});
      launch_vec<0>(c, n, S, BcgP{p, z, r, v, m}, NoEpi{});
      (use32 ? info->spmv32 : info->spmv64) += 2;
    }
    launched += chunk;
    PM_CHECK_HIP(hipGetLastError());
    PM_TRY(poll_done(c, &done, &iter));
    if (done) break;
  }
  if (done < 0) return done;
  if (!done) return PM_ERR_MAXIT;
  info->iters += iter;
  info->resid = c->h_scal->resid;
  return PM_OK;
}

static int run_minres(pm_ctx *c, int n, const double *b, double *x, const double *m,
                      double tol, int maxit, int nrestart, int use32, pm_solve_info *info) {
  // rotating buffers: v0, v, p ; z0, z ; w00, w0
  double *buf[7];
  for (int i = 0; i < 7; i++) buf[i] = c->d_wk[i];
  PmScalars *S = c->d_scal;
  PM_TRY(reset_scalars(c));
  hipMemsetAsync(x, 0, sizeof(double) * n, c->stream);
  launch_vec<1>(c, n, S, MrConv{b, m}, MrConvEpi{tol});
  int done = 0, iter = 0, total = 0;
  for (int cycle = 0; cycle <= nrestart && !done; cycle++) {
    double *v0 = buf[0], *v = buf[1], *p = buf[2], *z0 = buf[3], *z = buf[4], *w00 = buf[5],
           *w0 = buf[6];
    launch_spmv<0, false>(c, 0, x, z, nullptr, nullptr, S, NoEpi{});
    launch_vec<1>(c, n, S, MrStart{b, m, z, v, v0, w0, w00, z0}, MrStartEpi{});
    int launched = 0;
    while (launched < maxit) {
      int chunk = std::min(next_chunk(launched), maxit - launched);
      for (int it = 0; it < chunk; it++) {
        c->cur_iter = total + launched + it + 1;
        launch_spmv<1, false>(c, use32, z0, p, z0, nullptr, S, MrAlpha{});
        (use32 ? info->spmv32 : info->spmv64)++;
        launch_vec<1>(c, n, S, MrP{p, z, v, v0, m}, MrQr{});
        launch_vec<0>(c, n, S, MrW{x, w00, z0, w0}, FinishEpi{});
        // rotate: v0 <- v, v <- p, p <- old v0; z0 <- z, z <- old z0;
        //         w00 <- w0, w0 <- w (held in w00's buffer)
        double *t0 = v0;
        v0 = v;
        v = p;
        p = t0;
        double *t1 = z0;
        z0 = z;
        z = t1;
        double *t2 = w00;
        w00 = w0;
        w0 = t2;
      }
      launched += chunk;
      PM_CHECK_HIP(hipGetLastError());
      PM_TRY(poll_done(c, &done, &iter));
      if (done) break;
    }
    total += launched;
  }
  if (done < 0) return done;
  if (!done) return PM_ERR_MAXIT;
  info->iters += iter;
  info->resid = c->h_scal->resid;
  return PM_OK;
}

static int run_gmres(pm_ctx *c, int n, const double *b, double *x, const double *m, int k,
                     double tol, int maxit, int use32, pm_solve_info *info) {
  if (k < 1) return PM_ERR_ARG;
  if (c->gm_k < k || c->wk_n < n) {
    if (c->d_gm) hipFree(c->d_gm);
    c->d_gm = nullptr;
    size_t bytes = sizeof(double) * ((size_t)n * (k + 1) + (size_t)(k + 1) * (k + 4) + 3 * k + 8);
    PM_CHECK_HIP(hipMalloc(&c->d_gm, bytes));
    c->gm_k = k;
  }
  double *G = c->d_gm;
  double *h = G + (size_t)n * (k + 1);
  double *cs = h + (size_t)(k + 1) * k, *sn = cs + k, *e = sn + k, *y = e + k + 2;
  auto g = [&](int j) { return G + (size_t)j * n; };
  PmScalars *S = c->d_scal;
  PM_TRY(reset_scalars(c));
  launch_vec<1>(c, n, S, GmInitNorm{b, m, x, g(0)}, GmInitNormEpi{tol});
  int done = 0, iter = 0;
  for (int l = 0; l < maxit && !done; l++) {
    launch_spmv<0, false>(c, 0, x, g(0), nullptr, nullptr, S, NoEpi{});
    launch_vec<1>(c, n, S, GmRes0{b, m, g(0)}, GmE0{e});
    launch_vec<0>(c, n, S, GmScale{g(0)}, NoEpi{});
    for (int i = 0; i < k; i++) {
      // the cycle-end flag (2) makes the remaining launches of the cycle no-ops
      // through the done check of a shadow scalar block: poll every 8 columns
      if (i > 0 && (i % 8) == 0) {
        PM_TRY(poll_done(c, &done, &iter));
        if (c->h_scal->flag == 2 || done) break;
      }
      double *gi = g(i), *gn = g(i + 1);
      launch_spmv<1, true>(c, use32, gi, gn, gn, m, S, GmTau{});
      (use32 ? info->spmv32 : info->spmv64)++;
      for (int sweep = 0; sweep < 2; sweep++) {
        // first dot with g0
        launch_vec<1>(c, n, S, GmDot{gn, g(0)}, GmSetH{h, i * (k + 1) + 0, sweep});
        for (int j = 0; j <= i; j++) {
          const double *nxt = (j < i) ? g(j + 1) : nullptr;
          if (j < i)
            launch_vec<1>(c, n, S, GmAxpyDot{gn, g(j), nxt},
                          GmSetH{h, i * (k + 1) + j + 1, sweep});
          else if (sweep == 0)
            launch_vec<1>(c, n, S, GmAxpyDot{gn, g(j), nullptr}, GmRefineCheck{});
          else
            launch_vec<1>(c, n, S, GmAxpyDot{gn, g(j), nullptr}, NoEpi{});
        }
        if (sweep == 0) {
          PM_TRY(poll_done(c, &done, &iter));
          if (c->h_scal->flag != 1) break;
        }
      }
      launch_vec<1>(c, n, S, GmDot{gn, gn}, GmGivens{h, cs, sn, e, i, k, 0});
      launch_vec<0>(c, n, S, GmScale{gn}, NoEpi{});
      PM_TRY(poll_done(c, &done, &iter));
      if (c->h_scal->flag == 2) break;
    }
    launch_vec<0>(c, 8, S, NoBody{}, GmBack{h, e, y, k}, 1);
    launch_vec<0>(c, n, S, GmX{x, G, y, (int64_t)n, k}, GmEnd{e});
    PM_CHECK_HIP(hipGetLastError());
    PM_TRY(poll_done(c, &done, &iter));
  }
  if (done < 0) return done;
  if (!done) return PM_ERR_MAXIT;
  info->iters += iter;
  info->resid = c->h_scal->resid;
  return PM_OK;
}

// ---------------------------------------------------------------------------
// Jacobi preconditioner (cal_precond/pre_diag, Precond.f:1601-1695, 22-44)
// ---------------------------------------------------------------------------
struct JacobiBody {
  const double *ad;
  double *m;
  int absval;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double a = ad[i];
    if (a == 0.0) acc[0] += 1.0;
    double mi = 1.0 / a;
    m[i] = absval ? fabs(mi) : mi;
  }
};
struct CountEpi {
  __device__ void operator()(PmScalars *S, const double *r) const { S->red[0] = r[0]; }
};
// relative residual ||b - A x|| / ||b|| (r holds A x on entry)
struct TrueRes {
  const double *b;
  double *r;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    double bi = b[i], ri = bi - r[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += bi * bi;
  }
};
struct RedEpi {
  __device__ void operator()(PmScalars *S, const double *r) const {
    S->red[0] = r[0];
    S->red[1] = r[1];
  }
};
struct AxpyBody {  // x += dx
  double *x;
  const double *dx;
  __device__ void operator()(int64_t i, const PmScalars &, double *) const { x[i] = x[i] + dx[i]; }
};
struct CopyBody {
  double *dst;
  const double *src;
  __device__ void operator()(int64_t i, const PmScalars &, double *) const { dst[i] = src[i]; }
};
struct RoundM32 {  // m rounded to FP32 for the refinement inner solves
  const double *m;
  double *m32;
  __device__ void operator()(int64_t i, const PmScalars &, double *) const {
    m32[i] = (double)(float)m[i];
  }
};

int pm_ensure_workspace(pm_ctx *c, int n) {
  if (c->wk_n >= n && c->d_wk[0]) return PM_OK;
  for (int i = 0; i < 16; i++) {
    if (c->d_wk[i]) hipFree(c->d_wk[i]);
    c->d_wk[i] = nullptr;
  }
  for (int i = 0; i < 16; i++) PM_CHECK_HIP(hipMalloc(&c->d_wk[i], sizeof(double) * (size_t)n));
  c->wk_n = n;
  return PM_OK;
}

// scalar read of the last reduction (synchronous)
static int read_red(pm_ctx *c, double *r0, double *r1) {
  PM_CHECK_HIP(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(PmScalars), hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (r0) *r0 = c->h_scal->red[0];
  if (r1) *r1 = c->h_scal->red[1];
  return PM_OK;
}

static int krylov(pm_ctx *c, const pm_solver_cfg *cfg, int n, const double *b, double *x,
                  const double *m, double tol, int use32, pm_solve_info *info) {
  int used = cfg->solver;
  switch (cfg->solver) {
    case PM_SOLVER_PCG:
      PM_TRY(run_pcg(c, n, b, x, m, tol, cfg->maxit, use32, info));
      break;
    case PM_SOLVER_BICGSTAB:
      PM_TRY(run_bicgstab(c, n, b, x, m, tol, cfg->maxit, use32, info));
      break;
    case PM_SOLVER_MINRES:
      PM_TRY(run_minres(c, n, b, x, m, tol, cfg->maxit, 10, use32, info));
      break;
    case PM_SOLVER_GMRES:
      PM_TRY(run_gmres(c, n, b, x, m, cfg->nkrylov, tol, cfg->maxit, use32, info));
      break;
    case PM_SOLVER_SQMR: {
      int failed = 0;
      PM_TRY(run_sqmr(c, n, b, x, m, tol, cfg->maxit, use32, info, &failed));
      if (failed && c->no_alloc) return PM_ERR_MAXIT;
      if (failed) {
        used = PM_SOLVER_GMRES;
        PM_TRY(run_gmres(c, n, b, x, m, cfg->nkrylov, tol, cfg->maxit, use32, info));
      }
      break;
    }
    default:
      return PM_ERR_UNSUPPORTED;
  }
  info->used_solver = used;
  return PM_OK;
}

int pm_solve_dev(pm_ctx *c, const pm_solver_cfg *cfg, const double *b, double *x,
                 pm_solve_info *info) {
  const int n = c->A.neq;
  memset(info, 0, sizeof(*info));
  if (cfg->precond != PM_PRECOND_DIAG && cfg->precond != PM_PRECOND_ABSDIAG)
    return PM_ERR_UNSUPPORTED;
  PM_TRY(pm_ensure_workspace(c, n));
  PM_CHECK_HIP(hipEventRecord(c->ev0, c->stream));
  // m = 1/ad (or |1/ad|) with a zero-diagonal count
  PM_TRY(reset_scalars(c));
  launch_vec<1>(c, n, c->d_scal, JacobiBody{c->A.diag, c->d_m, cfg->precond == PM_PRECOND_ABSDIAG},
                CountEpi{});
  double nzero = 0.0;
  PM_TRY(read_red(c, &nzero, nullptr));
  if (nzero > 0.0) return PM_ERR_ZERO_DIAG;
  double *res = c->d_wk[12];
  if (c->batch) {  // every box its own system, all boxes in the same launches
    if (cfg->precision == PM_PREC_FP64) {
      PM_TRY(pm_batch_krylov(c, cfg->solver, cfg->maxit, cfg->tol, b, x, c->d_m, 0, info));
    } else {
      if (!c->A.have32) PM_TRY(pm_make_val32(c));
      double *rr = c->d_wk[13], *dx = c->d_wk[14], *m32 = c->d_wk[15];
      launch_vec<0>(c, n, c->d_scal, RoundM32{c->d_m, m32}, NoEpi{}, 1);
      launch_vec<0>(c, n, c->d_scal, CopyBody{rr, b}, NoEpi{}, 1);
      PM_CHECK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, c->stream));
      const int nbox = pm_batch_nbox(c);
      std::vector<double> r2, b2;
      std::vector<int> active(nbox, 1);
      const int maxouter = cfg->ir_maxouter > 0 ? cfg->ir_maxouter : 10;
      const double inner = cfg->ir_inner_tol > 0 ? cfg->ir_inner_tol : 1e-5;
      bool all = false;
      for (int o = 0; o < maxouter && !all; o++) {
        pm_solve_info sub;
        memset(&sub, 0, sizeof(sub));
        const int rc32 = pm_batch_krylov(c, cfg->solver, cfg->maxit, inner, rr, dx, m32, 1, &sub,
                                         /*cap_stragglers=*/1);
        if (rc32 != PM_OK && rc32 != PM_ERR_MAXIT) return rc32;
        info->iters += sub.iters;
        info->spmv32 += sub.spmv32;
        info->outer++;
        if (rc32 == PM_ERR_MAXIT) {
          // boxes whose FP32 inner solve did not converge: their correction
          // is solved again with FP64 values, the other boxes' kept
          std::vector<int> open, keep(nbox);
          PM_TRY(pm_batch_unconverged(c, open));
          for (int k = 0; k < nbox; k++) keep[k] = !open[k];
          double *rr64 = c->d_wk[10], *dx64 = c->d_wk[11];
          launch_vec<0>(c, n, c->d_scal, CopyBody{rr64, rr}, NoEpi{}, 1);
          PM_TRY(pm_batch_mask_rhs(c, rr64, open));
          PM_CHECK_HIP(hipMemsetAsync(dx64, 0, sizeof(double) * n, c->stream));
          pm_solve_info sub64;
          memset(&sub64, 0, sizeof(sub64));
          PM_TRY(pm_batch_krylov(c, cfg->solver, cfg->maxit, inner, rr64, dx64, c->d_m, 0, &sub64));
          info->iters += sub64.iters;
          info->spmv64 += sub64.spmv64;
          PM_TRY(pm_batch_mask_rhs(c, dx, keep));
          launch_vec<0>(c, n, c->d_scal, AxpyBody{dx, dx64}, NoEpi{}, 1);
        }
        launch_vec<0>(c, n, c->d_scal, AxpyBody{x, dx}, NoEpi{}, 1);
        PM_TRY(pm_batch_true_res(c, b, x, res, r2, b2));
        info->spmv64++;
        all = true;
        double worst = 0.0;
        for (int k = 0; k < nbox; k++) {
          const double rel = b2[k] > 0 ? sqrt(r2[k] / b2[k]) : 0.0;
          active[k] = rel >= cfg->tol;
          all = all && !active[k];
          worst = std::max(worst, rel);
        }
        info->resid = worst;
        if (!all) {
          launch_vec<0>(c, n, c->d_scal, CopyBody{rr, res}, NoEpi{}, 1);
          PM_TRY(pm_batch_mask_rhs(c, rr, active));
        }
      }
      if (!all) return PM_ERR_MAXIT;
      info->used_solver = cfg->solver;
    }
    std::vector<double> r2, b2;
    PM_TRY(pm_batch_true_res(c, b, x, res, r2, b2));
    double worst = 0.0;
    for (size_t k = 0; k < r2.size(); k++)
      if (b2[k] > 0) worst = std::max(worst, sqrt(r2[k] / b2[k]));
    info->true_rel = worst;
    PM_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
    PM_CHECK_HIP(hipEventSynchronize(c->ev1));
    float bms = 0.f;
    PM_CHECK_HIP(hipEventElapsedTime(&bms, c->ev0, c->ev1));
    info->ms = bms;
    return PM_OK;
  }
  if (cfg->precision == PM_PREC_FP64) {
    PM_TRY(krylov(c, cfg, n, b, x, c->d_m, cfg->tol, 0, info));
  } else if (cfg->precision == PM_PREC_FP32IR) {
    if (!c->A.have32) PM_TRY(pm_make_val32(c));
    double *rr = c->d_wk[13], *dx = c->d_wk[14], *m32 = c->d_wk[15];
    launch_vec<0>(c, n, c->d_scal, RoundM32{c->d_m, m32}, NoEpi{}, 1);
    launch_vec<0>(c, n, c->d_scal, CopyBody{rr, b}, NoEpi{}, 1);
    PM_CHECK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, c->stream));
    double bnorm2 = 0.0;
    {
      PM_TRY(reset_scalars(c));
      launch_vec<2>(c, n, c->d_scal, TrueRes{b, res}, RedEpi{});  // res holds garbage*0
    }
    // ||b|| from a clean pass: res = b - 0
    PM_CHECK_HIP(hipMemsetAsync(res, 0, sizeof(double) * n, c->stream));
    PM_TRY(reset_scalars(c));
    launch_vec<2>(c, n, c->d_scal, TrueRes{b, res}, RedEpi{});
    PM_TRY(read_red(c, nullptr, &bnorm2));
    double rel = 1.0;
    int maxouter = cfg->ir_maxouter > 0 ? cfg->ir_maxouter : 10;
    double inner = cfg->ir_inner_tol > 0 ? cfg->ir_inner_tol : 1e-5;
    for (int o = 0; o < maxouter; o++) {
      pm_solve_info sub;
      memset(&sub, 0, sizeof(sub));
      // stop the inner solve once it meets the outer target as well
      double itol = std::max(inner, 0.5 * cfg->tol / std::max(rel, 1e-300));
      int rc32 = krylov(c, cfg, n, rr, dx, m32, std::min(itol, 0.5), 1, &sub);
      if (rc32 == PM_ERR_MAXIT || rc32 == PM_ERR_BREAKDOWN) {
        // the FP32 inner solve failed: this correction is solved with FP64 values
        info->iters += sub.iters;
        info->spmv32 += sub.spmv32;
        memset(&sub, 0, sizeof(sub));
        rc32 = krylov(c, cfg, n, rr, dx, c->d_m, std::min(itol, 0.5), 0, &sub);
      }
      PM_TRY(rc32);
      info->iters += sub.iters;
      info->spmv32 += sub.spmv32;
      info->spmv64 += sub.spmv64;
      info->used_solver = sub.used_solver;
      info->outer++;
      launch_vec<0>(c, n, c->d_scal, AxpyBody{x, dx}, NoEpi{}, 1);
      // FP64 residual of the outer iterate
      c->cur_iter = 0;
      launch_spmv<0, false>(c, 0, x, res, nullptr, nullptr, nullptr, NoEpi{});
      info->spmv64++;
      PM_TRY(reset_scalars(c));
      launch_vec<2>(c, n, c->d_scal, TrueRes{b, res}, RedEpi{});
      double r2 = 0.0;
      PM_TRY(read_red(c, &r2, nullptr));
      rel = sqrt(r2 / bnorm2);
      info->resid = sqrt(r2);
      if (rel < cfg->tol) break;
      launch_vec<0>(c, n, c->d_scal, CopyBody{rr, res}, NoEpi{}, 1);
    }
    if (!(rel < cfg->tol)) return PM_ERR_MAXIT;
  } else {
    return PM_ERR_ARG;
  }
  // true FP64 residual of the result
  c->cur_iter = 0;
  launch_spmv<0, false>(c, 0, x, res, nullptr, nullptr, nullptr, NoEpi{});
  PM_TRY(reset_scalars(c));
  launch_vec<2>(c, n, c->d_scal, TrueRes{b, res}, RedEpi{});
  double r2 = 0.0, b2 = 0.0;
  PM_TRY(read_red(c, &r2, &b2));
  if (c->tev_used) pm_spmv_harvest(c, 0, 0);
  info->true_rel = b2 > 0 ? sqrt(r2 / b2) : 0.0;
  PM_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
  PM_CHECK_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  PM_CHECK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  info->ms = ms;
  return PM_OK;
}

// dot product of two device vectors (deterministic), host result
double pm_dot_host(pm_ctx *c, const double *a, const double *b, int n, int *rc) {
  struct DotBody {
    const double *a, *b;
    __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
      acc[0] += a[i] * b[i];
    }
  };
  *rc = reset_scalars(c);
  if (*rc) return 0.0;
  launch_vec<1>(c, n, c->d_scal, DotBody{a, b}, CountEpi{});
  double r = 0.0;
  *rc = read_red(c, &r, nullptr);
  return r;
}
