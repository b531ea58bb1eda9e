// pm_batch.hip -- the batching form: many independent Biot boxes held as one
// block-diagonal matrix and solved together.
//
// The boxes arrive as one disconnected mesh.  RCM numbers each connected
// component in turn (Reord.f genrcm_v1), so every box owns a contiguous
// range of equations and of node blocks.  Each launch covers all boxes:
// workgroups are assigned to boxes in proportion to their rows, every box
// has its own Krylov scalars, reductions and stop test (the reference's
// recurrences of pcg / rpsqmr, Iters.f:299-595, 6100-6436), and workgroups of
// a converged box return at entry, so finished boxes cost nothing.  Chunks of
// iterations are captured once in a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pm_functors.h"
#include "pm_spmv.h"

struct PmBatch {
  int nbox = 0;
  std::vector<int> eq0, blk0, wg0;  // nbox + 1
  int G = 0;
  int *d_wg_box = nullptr, *d_wg0 = nullptr, *d_eq0 = nullptr, *d_blk0 = nullptr;
  unsigned int *d_counters = nullptr;
  double *d_partials = nullptr;
  PmScalars *d_S = nullptr, *h_S = nullptr;
  int *d_ndone = nullptr, *h_ndone = nullptr;
  std::vector<int> last_iters;
  // hipGraph cache of one chunk of iterations
  hipGraphExec_t graph = nullptr;
  static constexpr int kGraphKey = 20;
  uintptr_t graph_key[kGraphKey] = {};
  int use_graph = 1;
  int rebalance = 1;  // hand converged boxes' workgroups to the open ones
  ~PmBatch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    (void)hipFree(d_wg_box);
    (void)hipFree(d_wg0);
    (void)hipFree(d_eq0);
    (void)hipFree(d_blk0);
    (void)hipFree(d_counters);
    (void)hipFree(d_partials);
    (void)hipFree(d_S);
    (void)hipFree(d_ndone);
    if (h_S) (void)hipHostFree(h_S);
    if (h_ndone) (void)hipHostFree(h_ndone);
  }
};

void pm_batch_free(PmBatch *b) { delete b; }

struct BatchArgs {
  const int *wg_box, *wg0, *eq0, *blk0;
  unsigned int *counters;
  double *partials;
  int *ndone;
};

template <class Epi>
struct CountDone {  // epilogue wrapper: counts boxes that finish
  Epi e;
  int *ndone;
  __device__ void operator()(PmScalars *S, const double *r) const {
    const int was = S->done;
    e(S, r);
    if (!was && S->done) atomicAdd(ndone, 1);
  }
};

template <typename VT, int LANES, int UNR, int NV>
__global__ __launch_bounds__(PM_WG) void k_spmv_b(BatchArgs B, int neq, const int *__restrict__ fr,
                                                  const int *__restrict__ nr,
                                                  const int64_t *__restrict__ colptr,
                                                  const int64_t *__restrict__ valptr,
                                                  const int *__restrict__ ncol,
                                                  const VT *__restrict__ val,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y,
                                                  const double *__restrict__ w,
                                                  const PmScalars *S) {
  const int box = B.wg_box[blockIdx.x];
  const PmScalars *Sb = S + box;
  if (Sb->done) return;
  constexpr int GROUPS = PM_WG / LANES;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int gb = B.wg0[box], gc = B.wg0[box + 1] - gb;
  const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
  const int64_t nb0 = B.blk0[box], nblk = B.blk0[box + 1] - nb0;
  int64_t k0, k1;
  pm_chunk(nblk, blockIdx.x - gb, gc, k0, k1);
  double d0 = 0.0, d1 = 0.0;
  PmBlockMeta nx = pm_block_meta(nb0 + k0 + grp, nb0 + k1, colptr, valptr, nr);
  for (int64_t kk = k0 + grp; kk < k1; kk += GROUPS) {
    const int64_t k = nb0 + kk;
    const PmBlockMeta cur = pm_uniform<LANES>(nx);
    nx = pm_block_meta(k + GROUPS, nb0 + k1, colptr, valptr, nr);
    const int n = cur.n;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (LANES == 64 && sizeof(VT) == 4)
      pm_spmv_block64<VT, UNR>(lane, cur, ncol, val, xr, a);
    else
      pm_spmv_block<VT, LANES, UNR>(lane, cur, ncol, val, x, a);
    if (lane == LANES - 1) {
      const int f = fr[k];
      for (int q = 0; q < n; q++) {
        y[f + q] = a[q];
        if (NV >= 1) d0 += a[q] * w[f + q];
        if (NV >= 2) d1 += a[q] * a[q];
      }
    }
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, B.partials);
  }
}

// The batched product with one run of <= 4 consecutive columns per lane, as
// k_spmv_ch: each workgroup streams its part of its box's node blocks.
template <typename VT, int UNR, int NV, bool ADAPT = false>
__global__ __launch_bounds__(PM_WG, sizeof(VT) == 4 ? 6 : 1) void k_spmv_b_ch(
    BatchArgs B, int neq, const int *__restrict__ fr, const int *__restrict__ nr,
    const int64_t *__restrict__ colptr, const int64_t *__restrict__ valptr,
    const int64_t *__restrict__ chptr, const unsigned int *__restrict__ chunk,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    const double *__restrict__ w, const PmScalars *S) {
  const int box = B.wg_box[blockIdx.x];
  if (S[box].done) return;
  constexpr int GROUPS = PM_WG / 64;
  const __amdgpu_buffer_rsrc_t xr = pm_rsrc(x, (unsigned int)neq * 8u);
  const int lane = threadIdx.x % 64;
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int gb = B.wg0[box], gc = B.wg0[box + 1] - gb;
  const int64_t nb0 = B.blk0[box], nblk = B.blk0[box + 1] - nb0;
  int64_t c0, c1;
  pm_chunk(nblk, blockIdx.x - gb, gc, c0, c1);
  const int64_t k0 = nb0 + c0, k1 = nb0 + c1;
  double d0 = 0.0, d1 = 0.0;
  PmRunMeta cur = pm_run_meta(k0 + grp, k1, colptr, valptr, nr, chptr);
  PmRunMeta nx = pm_run_meta(k0 + grp + GROUPS, k1, colptr, valptr, nr, chptr);
  unsigned int dsc[UNR];
  {
    const __amdgpu_buffer_rsrc_t h0 = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
#pragma unroll
    for (int u = 0; u < UNR; u++)
      dsc[u] = __builtin_amdgcn_raw_buffer_load_b32(h0, (unsigned int)(lane + u * 64) * 4u, 0, 2);
  }
  for (int64_t k = k0 + grp; k < k1; k += GROUPS) {  // ADAPT: blocks of <= 64 runs in one pass of one run per lane
    const PmRunMeta nn = pm_run_meta(k + 2 * GROUPS, k1, colptr, valptr, nr, chptr);
    const int n = cur.n;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const __amdgpu_buffer_rsrc_t vr = pm_rsrc(
        val + cur.vp, ((unsigned int)n * (unsigned int)cur.c + 3u) * (unsigned int)sizeof(VT));
    const unsigned int rowb = (unsigned int)cur.c * (unsigned int)sizeof(VT);
    const __amdgpu_buffer_rsrc_t hr = pm_rsrc(chunk + cur.hp, (unsigned int)cur.nh * 4u);
    const __amdgpu_buffer_rsrc_t hrn = pm_rsrc(chunk + nx.hp, (unsigned int)nx.nh * 4u);
    unsigned int dn[UNR];
    if (ADAPT && cur.nh <= 64) {
      const unsigned int d1[1] = {dsc[0]};
      (void)pm_run_pass<VT, 1, true, UNR>(lane, 0, cur, 0u, d1, vr, rowb, xr, hrn, dn, a);
    } else {
      unsigned int base = pm_run_pass<VT, UNR, true>(lane, 0, cur, 0u, dsc, vr, rowb, xr, hrn, dn, a);
      for (int h0 = UNR * 64; h0 < cur.nh; h0 += UNR * 64) {
        unsigned int dp[UNR], unused[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++)
          dp[u] = __builtin_amdgcn_raw_buffer_load_b32(hr, (unsigned int)(h0 + lane + u * 64) * 4u, 0, 2);
        base = pm_run_pass<VT, UNR, false>(lane, h0, cur, base, dp, vr, rowb, xr, hrn, unused, a);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) dsc[u] = dn[u];
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = pm_group_sum<64>(a[q]);
    if (lane == 63) {
      const int f = fr[k];
      for (int q = 0; q < n; q++) {
        y[f + q] = a[q];
        if (NV >= 1) d0 += a[q] * w[f + q];
        if (NV >= 2) d1 += a[q] * a[q];
      }
    }
    cur = nx;
    nx = nn;
  }
  if (NV > 0) {
    double v[2] = {d0, d1};
    pm_block_partials<NV>(v, B.partials);
  }
}

template <int NV, class Body>
__global__ __launch_bounds__(PM_WG) void k_vec_b(BatchArgs B, const PmScalars *S, Body body,
                                                 int ignore_done) {
  const int box = B.wg_box[blockIdx.x];
  const PmScalars *Sb = S + box;
  if (!ignore_done && Sb->done) return;
  const int gb = B.wg0[box], gc = B.wg0[box + 1] - gb;
  const int64_t r0 = B.eq0[box], nrow = B.eq0[box + 1] - r0;
  int64_t a, e;
  pm_chunk(nrow, blockIdx.x - gb, gc, a, e);
  double acc[NV > 0 ? NV : 1];
  for (int i = 0; i < (NV > 0 ? NV : 1); i++) acc[i] = 0.0;
  const PmScalars &s = *Sb;
  for (int64_t i = a + threadIdx.x; i < e; i += PM_WG) body(r0 + i, s, acc);
  pm_block_partials<NV>(acc, B.partials);
}

// one workgroup per box: sums the box's workgroup partials of the kernel
// before it (fixed order) and runs the box's epilogue
template <int NV, class Epi>
__global__ __launch_bounds__(PM_WG) void k_finish_b(BatchArgs B, PmScalars *S, Epi epi,
                                                    int ignore_done) {
  const int box = blockIdx.x;
  PmScalars *Sb = S + box;
  if (!ignore_done && Sb->done) return;
  const int gb = B.wg0[box], gc = B.wg0[box + 1] - gb;
  double out[NV > 0 ? NV : 1];
  pm_sum_partials<NV>(gc, B.partials + (size_t)gb * NV, out);
  if (threadIdx.x == 0) {
    for (int i = 0; i < NV; i++) Sb->red[i] = out[i];
    epi(Sb, out);
  }
}

// a box whose right-hand side is zero has x = 0 and is done at once
template <class Epi>
struct ZeroRhsDone {
  Epi e;
  int *ndone;
  __device__ void operator()(PmScalars *S, const double *r) const {
    e(S, r);
    if (r[0] == 0.0) {
      S->done = 1;
      S->iter = 0;
      atomicAdd(ndone, 1);
    }
  }
};

static BatchArgs batch_args(PmBatch *B) {
  return BatchArgs{B->d_wg_box, B->d_wg0, B->d_eq0, B->d_blk0, B->d_counters, B->d_partials,
                   B->d_ndone};
}

template <int NV, class Epi>
static void bfinish(pm_ctx *c, Epi epi, int ignore_done) {
  if (NV == 0 && std::is_same<Epi, NoEpi>::value) return;
  PmBatch *B = c->batch;
  hipLaunchKernelGGL((k_finish_b<NV, Epi>), dim3(B->nbox), dim3(PM_WG), 0, c->stream,
                     batch_args(B), B->d_S, epi, ignore_done);
}

template <typename VT, int NV>
static void bspmv_shape(pm_ctx *c, int cfg, const VT *val, const double *x, double *y,
                        const double *w) {
  PmBatch *B = c->batch;
  const PmDevMatrix &A = c->A;
#define PM_BSPMV_CASE(id, L, U)                                                                 \
  case id:                                                                                      \
    hipLaunchKernelGGL((k_spmv_b<VT, L, U, NV>), dim3(B->G), dim3(PM_WG), 0, c->stream,        \
                       batch_args(B), A.neq, A.fr, A.nr, A.colptr, A.valptr, A.ncol, val, x,   \
                       y, w,                                                                    \
                       B->d_S);                                                                 \
    break;
#define PM_BSPMV_CH_CASE(id, U)                                                                 \
  case id:                                                                                      \
    hipLaunchKernelGGL((k_spmv_b_ch<VT, U, NV>), dim3(B->G), dim3(PM_WG), 0, c->stream,        \
                       batch_args(B), A.neq, A.fr, A.nr, A.colptr, A.valptr, A.chptr, A.chunk,  \
                       val, x, y, w, B->d_S);                                                   \
    break;
  switch (cfg) {  // no column-prefetch nor sweep form here: 6-8 and 13 run as 2-4, 11-12 as 9-10
    PM_BSPMV_CASE(0, 64, 1)
    PM_BSPMV_CASE(1, 64, 2)
    PM_BSPMV_CASE(2, 64, 3)
    PM_BSPMV_CASE(3, 64, 4)
    PM_BSPMV_CASE(4, 64, 5)
    PM_BSPMV_CASE(6, 64, 3)
    PM_BSPMV_CASE(7, 64, 4)
    PM_BSPMV_CASE(8, 64, 5)
    PM_BSPMV_CH_CASE(9, 1)
    PM_BSPMV_CH_CASE(10, 2)
    PM_BSPMV_CH_CASE(11, 1)
    PM_BSPMV_CH_CASE(12, 2)
    PM_BSPMV_CASE(13, 64, 5)
    case 14:
      hipLaunchKernelGGL((k_spmv_b_ch<VT, 2, NV, true>), dim3(B->G), dim3(PM_WG), 0, c->stream,
                         batch_args(B), A.neq, A.fr, A.nr, A.colptr, A.valptr, A.chptr, A.chunk,
                         val, x, y, w, B->d_S);
      break;
    default: PM_BSPMV_CASE(5, 32, 4)  // shape 15 has no batched form: shape 5's kernel
  }
#undef PM_BSPMV_CASE
#undef PM_BSPMV_CH_CASE
}

template <int NV, class Epi>
static void bspmv(pm_ctx *c, int use32, const double *x, double *y, const double *w, Epi epi) {
  const int cfg = c->spmv_cfg[use32 ? 1 : 0];
  if (use32) bspmv_shape<float, NV>(c, cfg, c->A.val32, x, y, w);
  else bspmv_shape<double, NV>(c, cfg, c->A.val, x, y, w);
  if (NV > 0) bfinish<NV>(c, epi, 0);
}

// plain batched product y = A x (every box), for the shape tuner
int pm_batch_spmv_dev(pm_ctx *c, const double *x, double *y, int use32) {
  PmBatch *B = c->batch;
  PM_CHECK_HIP(hipMemsetAsync(B->d_S, 0, sizeof(PmScalars) * B->nbox, c->stream));
  bspmv<0>(c, use32, x, y, nullptr, NoEpi{});
  PM_CHECK_HIP(hipGetLastError());
  return PM_OK;
}

template <int NV, class Body, class Epi>
static void bvec(pm_ctx *c, Body body, Epi epi, int ignore_done = 0) {
  PmBatch *B = c->batch;
  hipLaunchKernelGGL((k_vec_b<NV, Body>), dim3(B->G), dim3(PM_WG), 0, c->stream, batch_args(B),
                     B->d_S, body, ignore_done);
  bfinish<NV>(c, epi, ignore_done);
}

static int breset(pm_ctx *c) {
  PmBatch *B = c->batch;
  PM_CHECK_HIP(hipMemsetAsync(B->d_S, 0, sizeof(PmScalars) * B->nbox, c->stream));
  PM_CHECK_HIP(hipMemsetAsync(B->d_ndone, 0, sizeof(int), c->stream));
  return PM_OK;
}

static int bpoll(pm_ctx *c, int *ndone) {
  PmBatch *B = c->batch;
  PM_CHECK_HIP(hipMemcpyAsync(B->h_ndone, B->d_ndone, sizeof(int), hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  *ndone = *B->h_ndone;
  return PM_OK;
}

// Hands the launch's G workgroups to the boxes in proportion to their rows:
// all boxes (open == nullptr, the definition's assignment) or the open ones
// only (at least one each, the remainder to the largest), so boxes that are
// still iterating get the workgroups of the converged ones.  The kernels
// read the assignment from device memory at every launch, captured graphs
// included.
static int bassign(pm_ctx *c, const std::vector<int> *open) {
  PmBatch *B = c->batch;
  const int nbox = B->nbox, G = B->G;
  std::vector<int> wg0(nbox + 1, 0);
  if (open == nullptr) {
    wg0 = B->wg0;
  } else {
    int64_t tot = 0;
    int nopen = 0, big = -1;
    for (int b = 0; b < nbox; b++)
      if ((*open)[b]) {
        const int rows = B->eq0[b + 1] - B->eq0[b];
        tot += rows;
        nopen++;
        if (big < 0 || rows > B->eq0[big + 1] - B->eq0[big]) big = b;
      }
    if (nopen == 0 || nopen > G) return PM_OK;
    std::vector<int> a(nbox, 0);
    int used = 0;
    for (int b = 0; b < nbox; b++)
      if ((*open)[b]) {
        a[b] = 1 + (int)((double)(B->eq0[b + 1] - B->eq0[b]) / (double)tot * (G - nopen));
        used += a[b];
      }
    a[big] += G - used;
    for (int b = 0; b < nbox; b++) wg0[b + 1] = wg0[b] + a[b];
  }
  std::vector<int> wg_box(G);
  for (int b = 0; b < nbox; b++)
    for (int g = wg0[b]; g < wg0[b + 1]; g++) wg_box[g] = b;
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  PM_CHECK_HIP(hipMemcpy(B->d_wg0, wg0.data(), sizeof(int) * (nbox + 1), hipMemcpyHostToDevice));
  PM_CHECK_HIP(hipMemcpy(B->d_wg_box, wg_box.data(), sizeof(int) * G, hipMemcpyHostToDevice));
  return PM_OK;
}

// one chunk of iterations of the batched solver
static void bchunk(pm_ctx *c, int solver, int use32, int chunk, const double *m, double *x) {
  PmBatch *B = c->batch;
  double **w = c->d_wk;
  for (int it = 0; it < chunk; it++) {
    if (solver == PM_SOLVER_PCG) {
      double *r = w[0], *z = w[1], *p = w[2];
      bspmv<1>(c, use32, p, z, p, PcgAlpha{});
      bvec<1>(c, PcgUpd{x, r, z, p, m}, PcgUpdEpi{});
      bvec<0>(c, PcgP{p, z}, CountDone<FinishEpi>{FinishEpi{}, B->d_ndone});
    } else {
      double *t = w[0], *r = w[1], *q = w[2], *d = w[3];
      bspmv<1>(c, use32, q, t, q, CountDone<SqmrSigma>{SqmrSigma{}, B->d_ndone});
      bvec<2>(c, SqmrRT{r, t, m}, SqmrRTEpi{});
      bvec<0>(c, SqmrDXQ{d, x, q, t}, CountDone<SqmrFinishEpi>{SqmrFinishEpi{}, B->d_ndone});
    }
  }
}

// batched solve of every box: b, x, m device vectors of neq
// cap_stragglers != 0: once half the boxes have converged at iteration L,
// stop at 4 L; the boxes still open are left unconverged (PM_ERR_MAXIT)
int pm_batch_krylov(pm_ctx *c, int solver, int maxit, double tol, const double *b, double *x,
                    const double *m, int use32, pm_solve_info *info, int cap_stragglers) {
  PmBatch *B = c->batch;
  if (solver != PM_SOLVER_SQMR && solver != PM_SOLVER_PCG) return PM_ERR_UNSUPPORTED;
  // the box ranges must describe the matrix the kernels are about to read
  if (B->eq0[B->nbox] != c->A.neq || B->blk0[B->nbox] != c->A.nrn) return PM_ERR_STATE;
  PM_TRY(breset(c));
  if (solver == PM_SOLVER_PCG) {
    double *r = c->d_wk[0], *z = c->d_wk[1], *p = c->d_wk[2];
    bvec<1>(c, PcgInit{b, m, x, r, z, p}, ZeroRhsDone<PcgInitEpi>{PcgInitEpi{tol}, B->d_ndone}, 1);
  } else {
    double *r = c->d_wk[1], *q = c->d_wk[2], *d = c->d_wk[3];
    bvec<2>(c, SqmrInit{b, m, x, r, q, d}, ZeroRhsDone<SqmrInitEpi>{SqmrInitEpi{tol}, B->d_ndone},
            1);
  }
  int launched = 0, ndone = 0;
  PM_TRY(bpoll(c, &ndone));
  const int chunk = 16;
  // graph of one chunk: kernels only.  The key holds every pointer and
  // launch parameter the captured kernels take, so a graph is never replayed
  // against a matrix, workspace or vector that has been reallocated.
  const PmDevMatrix &A = c->A;
  const uintptr_t key[PmBatch::kGraphKey] = {
      (uintptr_t)solver,    (uintptr_t)use32,     (uintptr_t)x,         (uintptr_t)m,
      (uintptr_t)c->d_wk[0], (uintptr_t)c->d_wk[1], (uintptr_t)c->d_wk[2], (uintptr_t)c->d_wk[3],
      (uintptr_t)A.val,     (uintptr_t)A.val32,   (uintptr_t)A.fr,      (uintptr_t)A.nr,
      (uintptr_t)A.colptr,  (uintptr_t)A.valptr,  (uintptr_t)A.ncol,    (uintptr_t)A.chptr,
      (uintptr_t)A.chunk,   (uintptr_t)A.neq,     (uintptr_t)B->G,
      (uintptr_t)c->spmv_cfg[use32 ? 1 : 0]};
  if (B->use_graph && (B->graph == nullptr || memcmp(key, B->graph_key, sizeof(key)) != 0)) {
    if (B->graph) (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
    hipGraph_t g;
    PM_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    bchunk(c, solver, use32, chunk, m, x);
    PM_CHECK_HIP(hipStreamEndCapture(c->stream, &g));
    PM_CHECK_HIP(hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    memcpy(B->graph_key, key, sizeof(key));
  }
  // PM_PROGRESS_SECONDS=s (opt-in): report the finished boxes on stderr every s seconds
  static const double progress_s = [] {
    const char *e = std::getenv("PM_PROGRESS_SECONDS");
    return e ? std::atof(e) : 0.0;
  }();
  auto last = std::chrono::steady_clock::now();
  int half_at = -1;
  // each time the open boxes drop to 3/4 of the last count, their
  // workgroups are handed out again among the open ones
  int open_last = B->nbox - ndone;
  bool reassigned = false;
  while (ndone < B->nbox && launched < maxit) {
    if (cap_stragglers && half_at >= 0 && launched >= 4 * std::max(half_at, 64)) break;
    if (B->use_graph) PM_CHECK_HIP(hipGraphLaunch(B->graph, c->stream));
    else bchunk(c, solver, use32, chunk, m, x);
    launched += chunk;
    (use32 ? info->spmv32 : info->spmv64) += chunk;
    PM_CHECK_HIP(hipGetLastError());
    PM_TRY(bpoll(c, &ndone));
    if (half_at < 0 && 2 * ndone >= B->nbox) half_at = launched;
    const int open_now = B->nbox - ndone;
    if (B->rebalance && open_now > 0 && 4 * open_now <= 3 * open_last) {
      PM_CHECK_HIP(hipMemcpy(B->h_S, B->d_S, sizeof(PmScalars) * B->nbox, hipMemcpyDeviceToHost));
      std::vector<int> open(B->nbox);
      for (int k = 0; k < B->nbox; k++) open[k] = B->h_S[k].done == 0;
      PM_TRY(bassign(c, &open));
      reassigned = true;
      open_last = open_now;
    }
    if (progress_s > 0.0) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last).count() >= progress_s) {
        std::fprintf(stderr, "[pm_batch] %s: %d of %d boxes done after %d iterations\n",
                     use32 ? "fp32" : "fp64", ndone, B->nbox, launched);
        std::fflush(stderr);
        last = now;
      }
    }
  }
  if (reassigned) PM_TRY(bassign(c, nullptr));  // the vector passes after a solve cover every box
  PM_CHECK_HIP(hipMemcpyAsync(B->h_S, B->d_S, sizeof(PmScalars) * B->nbox, hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  B->last_iters.assign(B->nbox, 0);
  int worst = 0, maxiter = 0;
  double maxres = 0.0;
  for (int k = 0; k < B->nbox; k++) {
    const PmScalars &s = B->h_S[k];
    if (s.done < 0) worst = s.done;
    else if (s.done == 0 && worst == 0) worst = PM_ERR_MAXIT;
    B->last_iters[k] = s.iter;
    maxiter = std::max(maxiter, s.iter);
    maxres = std::max(maxres, s.resid);
  }
  info->iters += maxiter;
  info->resid = maxres;
  info->used_solver = solver;
  return worst;
}

// per-box ||R_u||^2, ||R_p||^2 (red[0], red[1] of every box's scalars)
struct ResNormBody {
  const double *b;
  const unsigned char *isp;
  __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
    const double bi = b[i];
    if (isp[i]) acc[1] += bi * bi;
    else acc[0] += bi * bi;
  }
};
struct ZeroBoxBody {  // clears the right-hand side of converged boxes
  double *b;
  const int *active;
  const int *eq_box;
  __device__ void operator()(int64_t i, const PmScalars &, double *) const {
    if (!active[eq_box[i]]) b[i] = 0.0;
  }
};

int pm_batch_resnorms(pm_ctx *c, const double *b, std::vector<double> &ru2,
                      std::vector<double> &rp2) {
  PmBatch *B = c->batch;
  PM_TRY(breset(c));
  bvec<2>(c, ResNormBody{b, c->d_isp}, NoEpi{}, 1);
  PM_CHECK_HIP(hipMemcpyAsync(B->h_S, B->d_S, sizeof(PmScalars) * B->nbox, hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  ru2.resize(B->nbox);
  rp2.resize(B->nbox);
  for (int k = 0; k < B->nbox; k++) {
    ru2[k] = B->h_S[k].red[0];
    rp2[k] = B->h_S[k].red[1];
  }
  return PM_OK;
}

int pm_batch_unconverged(pm_ctx *c, std::vector<int> &open) {
  PmBatch *B = c->batch;
  open.assign(B->nbox, 0);
  for (int k = 0; k < B->nbox; k++) open[k] = B->h_S[k].done == 0;
  return PM_OK;
}

int pm_batch_mask_rhs(pm_ctx *c, double *b, const std::vector<int> &active) {
  PmBatch *B = c->batch;
  int *d_active = nullptr;
  PM_CHECK_HIP(hipMalloc(&d_active, sizeof(int) * B->nbox));
  PM_CHECK_HIP(hipMemcpyAsync(d_active, active.data(), sizeof(int) * B->nbox,
                              hipMemcpyHostToDevice, c->stream));
  bvec<0>(c, ZeroBoxBody{b, d_active, c->d_eq_box}, NoEpi{}, 1);
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(d_active);
  return PM_OK;
}

int pm_batch_define_dev(pm_ctx *c, int nbox, const int *elem_off) {
  if (!c->have_mesh) return PM_ERR_STATE;
  const PmHostMesh &m = c->mesh;
  if (nbox < 1 || elem_off[0] != 0 || elem_off[nbox] != m.numel) return PM_ERR_ARG;
  // equation range of every box from its elements' equations
  std::vector<int> lo(nbox, INT32_MAX), hi(nbox, -1);
  for (int b = 0; b < nbox; b++) {
    if (elem_off[b + 1] <= elem_off[b]) return PM_ERR_ARG;
    for (int e = elem_off[b]; e < elem_off[b + 1]; e++)
      for (int k = 0; k < m.nen; k++) {
        const int n = m.ix20[(size_t)e * 21 + k] - 1;
        const int nd = m.fnno[n] == 1 ? 4 : 3;
        for (int d = 0; d < nd; d++) {
          const int q = m.id[(size_t)n * 4 + d];
          if (q > 0) {
            lo[b] = std::min(lo[b], q - 1);
            hi[b] = std::max(hi[b], q - 1);
          }
        }
      }
  }
  PmBatch *B = new PmBatch();
  B->nbox = nbox;
  B->eq0.assign(nbox + 1, 0);
  for (int b = 0; b < nbox; b++) {
    if (hi[b] < 0 || lo[b] != B->eq0[b]) {  // boxes must hold consecutive equations
      delete B;
      return PM_ERR_ARG;
    }
    B->eq0[b + 1] = hi[b] + 1;
  }
  if (B->eq0[nbox] != m.neq) {
    delete B;
    return PM_ERR_ARG;
  }
  // node blocks of every box (blocks are in equation order)
  std::vector<int> eqblk(m.neq);
  PM_CHECK_HIP(hipMemcpy(eqblk.data(), c->A.eq_block, sizeof(int) * m.neq, hipMemcpyDeviceToHost));
  B->blk0.assign(nbox + 1, 0);
  for (int b = 0; b < nbox; b++) B->blk0[b] = eqblk[B->eq0[b]];
  B->blk0[nbox] = c->A.nrn;
  // workgroups in proportion to rows, at least one per box
  const int target = std::max(2048, nbox);
  B->wg0.assign(nbox + 1, 0);
  for (int b = 0; b < nbox; b++) {
    const double share = (double)(B->eq0[b + 1] - B->eq0[b]) / m.neq;
    B->wg0[b + 1] = B->wg0[b] + std::max(1, (int)(share * target));
  }
  B->G = B->wg0[nbox];
  std::vector<int> wg_box(B->G);
  for (int b = 0; b < nbox; b++)
    for (int g = B->wg0[b]; g < B->wg0[b + 1]; g++) wg_box[g] = b;
  std::vector<int> eq_box(m.neq);
  for (int b = 0; b < nbox; b++)
    for (int q = B->eq0[b]; q < B->eq0[b + 1]; q++) eq_box[q] = b;
  auto up = [&](int **d, const std::vector<int> &h) -> int {
    PM_CHECK_HIP(hipMalloc(d, sizeof(int) * h.size()));
    PM_CHECK_HIP(hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return PM_OK;
  };
  int rc = PM_OK;
  if ((rc = up(&B->d_wg_box, wg_box)) || (rc = up(&B->d_wg0, B->wg0)) ||
      (rc = up(&B->d_eq0, B->eq0)) || (rc = up(&B->d_blk0, B->blk0))) {
    delete B;
    return rc;
  }
  if (c->d_eq_box) (void)hipFree(c->d_eq_box);
  c->d_eq_box = nullptr;
  if ((rc = up(&c->d_eq_box, eq_box))) {
    delete B;
    return rc;
  }
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == PM_OK) rc = PM_ERR_HIP;
  };
  chk(hipMalloc(&B->d_counters, sizeof(unsigned int) * nbox));
  chk(hipMemset(B->d_counters, 0, sizeof(unsigned int) * nbox));
  chk(hipMalloc(&B->d_partials, sizeof(double) * 2 * B->G));
  chk(hipMalloc(&B->d_S, sizeof(PmScalars) * nbox));
  chk(hipHostMalloc((void **)&B->h_S, sizeof(PmScalars) * nbox, hipHostMallocDefault));
  chk(hipMalloc(&B->d_ndone, sizeof(int)));
  chk(hipHostMalloc((void **)&B->h_ndone, sizeof(int), hipHostMallocDefault));
  if (rc) {
    delete B;
    return rc;
  }
  if (c->batch) pm_batch_free(c->batch);
  c->batch = B;
  return PM_OK;
}

int pm_batch_nbox(pm_ctx *c) { return c->batch ? c->batch->nbox : 0; }

int pm_batch_last_iters(pm_ctx *c, int *iters) {
  if (!c->batch) return PM_ERR_STATE;
  for (int b = 0; b < c->batch->nbox; b++)
    iters[b] = b < (int)c->batch->last_iters.size() ? c->batch->last_iters[b] : 0;
  return PM_OK;
}

void pm_batch_set_graph(pm_ctx *c, int on) {
  if (c->batch) c->batch->use_graph = on;
}

// res = b - A x and per-box ||res||^2, ||b||^2
int pm_batch_true_res(pm_ctx *c, const double *b, const double *x, double *res,
                      std::vector<double> &r2, std::vector<double> &b2) {
  PmBatch *B = c->batch;
  PM_TRY(pm_spmv_dev(c, x, res, 0));
  PM_TRY(breset(c));
  struct TrueResB {
    const double *b;
    double *r;
    __device__ void operator()(int64_t i, const PmScalars &, double *acc) const {
      const double bi = b[i], ri = bi - r[i];
      r[i] = ri;
      acc[0] += ri * ri;
      acc[1] += bi * bi;
    }
  };
  bvec<2>(c, TrueResB{b, res}, NoEpi{}, 1);
  PM_CHECK_HIP(hipMemcpyAsync(B->h_S, B->d_S, sizeof(PmScalars) * B->nbox, hipMemcpyDeviceToHost,
                              c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  r2.resize(B->nbox);
  b2.resize(B->nbox);
  for (int k = 0; k < B->nbox; k++) {
    r2[k] = B->h_S[k].red[0];
    b2[k] = B->h_S[k].red[1];
  }
  return PM_OK;
}
