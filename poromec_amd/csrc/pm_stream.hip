// pm_stream.hip -- streaming kernels that measure what HBM delivers on this
// device: the reference points for the SpMV's bandwidth figures (measurement
// hooks only; no counterpart in the reference).
//
// kind 0: copy, one 32-byte element per lane and grid-stride loop (8 workgroups
//         per CU): the first reference point used by the traffic counters
// kind 1: read-only, 4 x 16 B in flight per lane, one-shot grid (16 KB per
//         workgroup), each workgroup's sum written once
// kind 2: kind 1 with nontemporal loads
// kind 3: copy, 4 x 16 B per lane, one-shot grid
// kind 4: write-only (fill), 4 x 16 B per lane, one-shot grid
// kind 5: read-only, 8 x 16 B in flight per lane, one-shot grid
#include <hip/hip_runtime.h>

#include "pm_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy_gs(int64_t n, const double4 *__restrict__ a,
                                                 double4 *__restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    b[i] = a[i];
}

typedef float f4v __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_read(int64_t n, const f4v *__restrict__ a,
                                              float *__restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x;
  f4v v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t i = base + (int64_t)u * 256;
    if (i < n) {
      if (NT) v[u] = __builtin_nontemporal_load(a + i);
      else v[u] = a[i];
    } else {
      v[u] = f4v{0.f, 0.f, 0.f, 0.f};
    }
  }
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < U; u++) s += v[u].x + v[u].y + v[u].z + v[u].w;
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <int U>
__global__ __launch_bounds__(256) void k_copy4(int64_t n, const float4 *__restrict__ a,
                                               float4 *__restrict__ b) {
  const int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x;
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t i = base + (int64_t)u * 256;
    if (i < n) v[u] = a[i];
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t i = base + (int64_t)u * 256;
    if (i < n) b[i] = v[u];
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_fill4(int64_t n, float4 *__restrict__ b, float val) {
  const int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x;
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t i = base + (int64_t)u * 256;
    if (i < n) b[i] = make_float4(val, val, val, val);
  }
}

}  // namespace

extern "C" int pm_bench_stream(pm_ctx *c, int kind, int64_t bytes, int reps,
                               double *ms_per_launch) {
  if (!c || bytes < 4096 || reps <= 0 || !ms_per_launch || kind < 0 || kind > 5)
    return PM_ERR_ARG;
  PM_CHECK_HIP(hipSetDevice(c->device));
  bytes -= bytes % 4096;
  const int64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  PM_CHECK_HIP(hipMalloc(&a, bytes));
  hipError_t eb = hipMalloc(&b, bytes);
  if (eb != hipSuccess) {
    (void)hipFree(a);
    return PM_ERR_HIP;
  }
  PM_CHECK_HIP(hipMemsetAsync(a, 0, bytes, c->stream));
  auto launch = [&]() {
    switch (kind) {
      case 0:
        hipLaunchKernelGGL(k_copy_gs, dim3(8 * c->cu_count), dim3(256), 0, c->stream, bytes / 32,
                           (const double4 *)a, (double4 *)b);
        break;
      case 1:
        hipLaunchKernelGGL((k_read<4, false>), dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0,
                           c->stream, n16, (const f4v *)a, (float *)b);
        break;
      case 2:
        hipLaunchKernelGGL((k_read<4, true>), dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0,
                           c->stream, n16, (const f4v *)a, (float *)b);
        break;
      case 3:
        hipLaunchKernelGGL((k_copy4<4>), dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0,
                           c->stream, n16, (const float4 *)a, (float4 *)b);
        break;
      case 4:
        hipLaunchKernelGGL((k_fill4<4>), dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0,
                           c->stream, n16, (float4 *)b, 1.0f);
        break;
      default:
        hipLaunchKernelGGL((k_read<8, false>), dim3((unsigned)((n16 + 2047) / 2048)), dim3(256), 0,
                           c->stream, n16, (const f4v *)a, (float *)b);
        break;
    }
  };
  launch();
  PM_CHECK_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; r++) launch();
  PM_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
  PM_CHECK_HIP(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  PM_CHECK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_per_launch = ms / reps;
  PM_CHECK_HIP(hipGetLastError());
  (void)hipFree(a);
  (void)hipFree(b);
  return PM_OK;
}
