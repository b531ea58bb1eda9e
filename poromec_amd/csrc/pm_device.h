// pm_device.h -- device helpers shared by the kernels of libpmhip.
#pragma once
#include "pm_internal.h"

#define PM_WG 256          // threads per workgroup
#define PM_WAVES (PM_WG / 64)
#define PM_MAX_GRID 2048   // fixed grid of the reduction kernels (multiple of 8)

// Device-side scalars of one Krylov solve.  Kernels read the scalars they
// need at entry; the last workgroup of a reduction writes the next ones.
struct PmScalars {
  double conv;        // stop threshold of the recurrence residual
  double breaktol;    // BiCGSTAB breakdown threshold
  double s[24];       // solver scalars (named per solver in pm_krylov.hip)
  double red[4];      // results of the last reduction
  double resid;       // recurrence residual at exit
  int done;           // 0 running, 1 converged, <0 PM_ERR_* code
  int pending;        // converged, finishing this iteration's update
  int it;             // iterations started
  int iter;           // iterations at convergence
  int flag;           // solver-specific switch (GMRES refinement / cycle end)
  int ni;             // GMRES inner iterations of the current cycle
  int pad[2];
};

__device__ __forceinline__ double pm_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Workgroup i runs on XCD (i mod 8); give each XCD one contiguous eighth of
// the index space so that its L2 sees neighbouring rows.
__device__ __forceinline__ int pm_xcd_block(int b, int g) {
  if ((g & 7) != 0) return b;
  return (b & 7) * (g >> 3) + (b >> 3);
}

// Workgroup partial sums of NV values, one slot per workgroup (fixed order
// inside the workgroup).  No fence and no atomic: the partials are summed by
// a later kernel on the same stream, after the kernel boundary has made them
// visible on every XCD (a device-scope fence per workgroup would write back
// the XCD's L2 each time).
template <int NV>
__device__ __forceinline__ void pm_block_partials(const double *v, double *partials) {
  if constexpr (NV > 0) {
    __shared__ double sh[NV][PM_WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double w = pm_wave_sum(v[i]);
      if (lane == 0) sh[i][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double acc = sh[i][0];
        for (int w = 1; w < PM_WAVES; w++) acc += sh[i][w];
        partials[(size_t)blockIdx.x * NV + i] = acc;
      }
    }
  }
}

// Sum of the partials of nblk workgroups in a fixed order, by one workgroup;
// every thread receives the sums.
template <int NV>
__device__ __forceinline__ void pm_sum_partials(int nblk, const double *partials, double *out) {
  if constexpr (NV > 0) {
    __shared__ double sh[NV][PM_WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double acc = 0.0;
      for (int b = threadIdx.x; b < nblk; b += PM_WG) acc += partials[(size_t)b * NV + i];
      acc = pm_wave_sum(acc);
      if (lane == 0) sh[i][wave] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double acc = sh[i][0];
      for (int w = 1; w < PM_WAVES; w++) acc += sh[i][w];
      out[i] = acc;
    }
  }
}

inline int pm_grid_for(int64_t n) {
  int64_t g = (n + PM_WG * 4 - 1) / (PM_WG * 4);
  if (g < 8) g = 8;
  if (g > PM_MAX_GRID) g = PM_MAX_GRID;
  return (int)((g + 7) & ~7);
}

// contiguous chunk of [0, n) owned by logical workgroup lb of g
__device__ __forceinline__ void pm_chunk(int64_t n, int lb, int g, int64_t &a, int64_t &b) {
  int64_t per = (n + g - 1) / g;
  a = (int64_t)lb * per;
  b = a + per;
  if (a > n) a = n;
  if (b > n) b = n;
}
