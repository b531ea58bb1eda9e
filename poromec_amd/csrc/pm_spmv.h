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
// over the LANES lanes finish them.
#pragma once
#include "pm_device.h"

// (LANES, UNR, PF) of each SpMV shape the tuner may pick; PF 1: the first
// pass's column indices are loaded one block ahead (pm_spmv_block64_pf);
// PF 2: one run of <= 4 consecutive columns per lane, UNR runs in flight
// (k_spmv_ch); PF 3: the same with the XCD's workgroups sweeping its blocks
// together; PF 4: PF 1 sweeping; PF 5: runs, one or two per lane by block
// width, sweeping (k_spmv_cha); PF 6: runs, LANES lanes per block (two
// blocks per wavefront), UNR runs per lane and pass, sweeping (k_spmv_chg)
#define PM_SPMV_NCFG 16
static constexpr int pm_spmv_lanes[PM_SPMV_NCFG] = {64, 64, 64, 64, 64, 32, 64, 64,
                                                    64, 64, 64, 64, 64, 64, 64, 32};
static constexpr int pm_spmv_unr[PM_SPMV_NCFG] = {1, 2, 3, 4, 5, 4, 3, 4, 5, 1, 2, 1, 2, 5, 1, 1};
static constexpr int pm_spmv_pf[PM_SPMV_NCFG] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 4, 5, 6};

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

// A whole wavefront works on one block when LANES == 64: its metadata is
// then wave-uniform, and moving it to scalar registers lets the loads use a
// scalar base with 32-bit lane offsets (fewer vector registers, more waves).
template <int LANES>
__device__ __forceinline__ PmBlockMeta pm_uniform(PmBlockMeta m) {
  if (LANES == 64) {
    m.cp = ((int64_t)__builtin_amdgcn_readfirstlane((int)(m.cp >> 32)) << 32) |
           (unsigned int)__builtin_amdgcn_readfirstlane((int)m.cp);
    m.vp = ((int64_t)__builtin_amdgcn_readfirstlane((int)(m.vp >> 32)) << 32) |
           (unsigned int)__builtin_amdgcn_readfirstlane((int)m.vp);
    m.c = __builtin_amdgcn_readfirstlane(m.c);
    m.n = __builtin_amdgcn_readfirstlane(m.n);
  }
  return m;
}

// one DPP move of a double (two 32-bit halves); lanes outside row_mask get 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double pm_dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over each group of LANES (16, 32 or 64) lanes with DPP moves instead
// of LDS permutes; the sum is valid in the group's last lane (15, 31 / 63)
template <int LANES>
__device__ __forceinline__ double pm_group_sum(double v) {
  v += pm_dpp<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += pm_dpp<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += pm_dpp<0x124, 0xf>(v);  // row_ror:4
  v += pm_dpp<0x128, 0xf>(v);  // row_ror:8  -> every lane holds its row of 16
  if (LANES >= 32) v += pm_dpp<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3 hold rows 0+1, 2+3
  if (LANES == 64) v += pm_dpp<0x143, 0xc>(v);  // row_bcast:31 -> row 3 holds all
  return v;
}

// The four row sums of a group of LANES (64 or 32) lanes finished together
// instead of one after another: two exchange steps inside each quad of lanes
// leave lane l with one row, row(l) = 2 (l & 1) + ((l >> 1) & 1), summed
// over its quad; two DPP rotations sum that row class over the lane's 16-lane
// DPP row, and gfx950's permlane16 swap (and for 64 lanes the permlane32
// swap) adds the group's DPP rows.  The group's lanes 0-3 end with rows 0, 2,
// 1, 3.  For 64 lanes that is about 33 VALU instructions against 72 for four
// pm_group_sum<64> (7 FP64 additions per lane against 24).
__device__ __forceinline__ int pm_row_of(int lane) { return 2 * (lane & 1) + ((lane >> 1) & 1); }

// v plus the value of the partner lane in the other 16-lane row of each pair
// of rows (SWAP32 false) or in the other half of the wavefront (true), for
// a 64-bit value as two 32-bit halves
template <bool SWAP32>
__device__ __forceinline__ double pm_swap_sum(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned int lo0 = (unsigned int)(b & 0xffffffffll), hi0 = (unsigned int)(b >> 32);
  const auto lo = SWAP32 ? __builtin_amdgcn_permlane32_swap(lo0, lo0, false, false)
                         : __builtin_amdgcn_permlane16_swap(lo0, lo0, false, false);
  const auto hi = SWAP32 ? __builtin_amdgcn_permlane32_swap(hi0, hi0, false, false)
                         : __builtin_amdgcn_permlane16_swap(hi0, hi0, false, false);
  const double a0 = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
  const double a1 = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
  return a0 + a1;
}

template <int LANES = 64>
__device__ __forceinline__ double pm_rows4_sum(int lane, const double a[4]) {
  const bool b0 = lane & 1, b1 = (lane >> 1) & 1;
  double k0 = b0 ? a[2] : a[0], k1 = b0 ? a[3] : a[1];
  const double g0 = b0 ? a[0] : a[2], g1 = b0 ? a[1] : a[3];
  k0 += pm_dpp<0xb1, 0xf>(g0);  // quad_perm [1,0,3,2]: partner lane l ^ 1
  k1 += pm_dpp<0xb1, 0xf>(g1);
  double s = b1 ? k1 : k0;
  const double g = b1 ? k0 : k1;
  s += pm_dpp<0x4e, 0xf>(g);   // quad_perm [2,3,0,1]: partner lane l ^ 2
  s += pm_dpp<0x124, 0xf>(s);  // row_ror:4
  s += pm_dpp<0x128, 0xf>(s);  // row_ror:8: the row class over the DPP row
  s = pm_swap_sum<false>(s);  // over 32 lanes
  return LANES == 64 ? pm_swap_sum<true>(s) : s;
}

// partial row sums of node block m, complete in the group's last lane.  Each
// pass covers UNR * LANES columns: every lane issues its UNR index loads and
// the value loads of the block's rows before the x gathers and the products,
// so a wave keeps a whole block's streams in flight instead of one 64-column
// slice at a time.  Past the end of the block a lane reloads the block's last
// column, past its row count row 0, and selects zero after the load:
// straight-line code, no branches between the loads.
// buffer resource over `bytes` bytes at p (raw, untyped); loads at offsets
// past `bytes` return zero
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pm_rsrc(const void *p, unsigned int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename VT, int AUX>
__device__ __forceinline__ VT pm_bload(__amdgpu_buffer_rsrc_t r, unsigned int off) {
  if constexpr (sizeof(VT) == 4)
    return __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
  else
    return __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
}

// The 64-lane form for FP32 values: the block's column list and values are
// addressed through wave-uniform buffer resources with 32-bit lane offsets
// (one register per load instead of a 64-bit address), and rows past the
// block's count read as zero from the hardware range check; x comes through
// a resource over the whole vector.  It keeps (64, 5) at 8 waves per SIMD
// (62 VGPRs against 92) and is ~3 % faster at 87^3; for FP64 values it was
// 7 % slower than plain global loads, which FP64 keeps.
template <typename VT, int UNR>
__device__ __forceinline__ void pm_spmv_block64(int lane, const PmBlockMeta &m,
                                                const int *__restrict__ ncol,
                                                const VT *__restrict__ val,
                                                __amdgpu_buffer_rsrc_t xr, double a[4]) {
  const int c = m.c, n = m.n;
  const __amdgpu_buffer_rsrc_t cr = pm_rsrc(ncol + m.cp, (unsigned int)c * 4u);
  const __amdgpu_buffer_rsrc_t vr =
      pm_rsrc(val + m.vp, (unsigned int)n * (unsigned int)c * (unsigned int)sizeof(VT));
  const unsigned int rowb = (unsigned int)c * (unsigned int)sizeof(VT);
  constexpr int NT = 2;  // non-temporal: the matrix streams through once
  for (int base = 0; base < c; base += UNR * 64) {
    unsigned int col[UNR];
    VT w0[UNR], w1[UNR], w2[UNR], w3[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const unsigned int j = (unsigned int)(base + lane + u * 64);
      col[u] = __builtin_amdgcn_raw_buffer_load_b32(cr, j * 4u, 0, NT);
      const unsigned int o = j * (unsigned int)sizeof(VT);
      w0[u] = pm_bload<VT, NT>(vr, o);
      w1[u] = pm_bload<VT, NT>(vr, o + rowb);
      w2[u] = pm_bload<VT, NT>(vr, o + 2u * rowb);
      w3[u] = pm_bload<VT, NT>(vr, o + 3u * rowb);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      // past the end of the block the column reads as 0 and the values as
      // the next row's: keep x out of those lanes
      const double xc = pm_bload<double, 0>(xr, col[u] * 8u);
      xv[u] = base + lane + u * 64 < c ? xc : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w0[u] * xv[u];
      a[1] += (double)w1[u] * xv[u];
      a[2] += (double)w2[u] * xv[u];
      a[3] += (double)w3[u] * xv[u];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<64>(a[q]);
}

// The 64-lane block with its first pass's column indices already in hand:
// on entry col[] holds block m's, on return block mn's (the next block of
// this wavefront).  Without them the x gathers wait one memory round trip
// for the index loads; with them the first pass issues its value loads and
// its gathers together, then the next block's index loads, and its sums wait
// only for the first two (loads complete in order, so the wait leaves the
// index loads in flight).  Buffer resources as in pm_spmv_block64: loads
// past a block's end, or of an empty next block, read zero.
template <typename VT, int UNR>
__device__ __forceinline__ void pm_spmv_block64_pf(int lane, const PmBlockMeta &m,
                                                   const PmBlockMeta &mn, unsigned int (&col)[UNR],
                                                   const int *__restrict__ ncol,
                                                   const VT *__restrict__ val,
                                                   __amdgpu_buffer_rsrc_t xr, double a[4]) {
  const int c = m.c, n = m.n;
  const __amdgpu_buffer_rsrc_t cr = pm_rsrc(ncol + m.cp, (unsigned int)c * 4u);
  const __amdgpu_buffer_rsrc_t vr =
      pm_rsrc(val + m.vp, (unsigned int)n * (unsigned int)c * (unsigned int)sizeof(VT));
  const __amdgpu_buffer_rsrc_t crn = pm_rsrc(ncol + mn.cp, (unsigned int)mn.c * 4u);
  const unsigned int rowb = (unsigned int)c * (unsigned int)sizeof(VT);
  constexpr int NT = 2;  // non-temporal: the matrix streams through once
  {
    VT w0[UNR], w1[UNR], w2[UNR], w3[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const unsigned int o = (unsigned int)(lane + u * 64) * (unsigned int)sizeof(VT);
      w0[u] = pm_bload<VT, NT>(vr, o);
      w1[u] = pm_bload<VT, NT>(vr, o + rowb);
      w2[u] = pm_bload<VT, NT>(vr, o + 2u * rowb);
      w3[u] = pm_bload<VT, NT>(vr, o + 3u * rowb);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const double xc = pm_bload<double, 0>(xr, col[u] * 8u);
      xv[u] = lane + u * 64 < c ? xc : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++)
      col[u] = __builtin_amdgcn_raw_buffer_load_b32(crn, (unsigned int)(lane + u * 64) * 4u, 0, NT);
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w0[u] * xv[u];
      a[1] += (double)w1[u] * xv[u];
      a[2] += (double)w2[u] * xv[u];
      a[3] += (double)w3[u] * xv[u];
    }
  }
  for (int base = UNR * 64; base < c; base += UNR * 64) {
    unsigned int cl[UNR];
    VT w0[UNR], w1[UNR], w2[UNR], w3[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const unsigned int j = (unsigned int)(base + lane + u * 64);
      cl[u] = __builtin_amdgcn_raw_buffer_load_b32(cr, j * 4u, 0, NT);
      const unsigned int o = j * (unsigned int)sizeof(VT);
      w0[u] = pm_bload<VT, NT>(vr, o);
      w1[u] = pm_bload<VT, NT>(vr, o + rowb);
      w2[u] = pm_bload<VT, NT>(vr, o + 2u * rowb);
      w3[u] = pm_bload<VT, NT>(vr, o + 3u * rowb);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const double xc = pm_bload<double, 0>(xr, cl[u] * 8u);
      xv[u] = base + lane + u * 64 < c ? xc : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w0[u] * xv[u];
      a[1] += (double)w1[u] * xv[u];
      a[2] += (double)w2[u] * xv[u];
      a[3] += (double)w3[u] * xv[u];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<64>(a[q]);
}

template <typename VT, int LANES, int UNR>
__device__ __forceinline__ void pm_spmv_block(int lane, const PmBlockMeta &m,
                                              const int *__restrict__ ncol,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ x, double a[4]) {
  const int *cols = ncol + m.cp;
  const VT *v0 = val + m.vp;
  const int c = m.c, n = m.n;
  // rows past the block's count reload row 0 and are zeroed after the load
  const int r1 = n > 1 ? c : 0, r2 = n > 2 ? 2 * c : 0, r3 = n > 3 ? 3 * c : 0;
  for (int base = 0; base < c; base += UNR * LANES) {  // same trip count in the group
    int col[UNR];
    bool ok[UNR];
    VT w0[UNR], w1[UNR], w2[UNR], w3[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int j = base + lane + u * LANES;
      ok[u] = j < c;
      const int jj = ok[u] ? j : c - 1;
      col[u] = __builtin_nontemporal_load(cols + jj);
      w0[u] = __builtin_nontemporal_load(v0 + jj);
      w1[u] = __builtin_nontemporal_load(v0 + r1 + jj);
      w2[u] = __builtin_nontemporal_load(v0 + r2 + jj);
      w3[u] = __builtin_nontemporal_load(v0 + r3 + jj);
    }
    double xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const double xc = x[col[u]];
      xv[u] = ok[u] ? xc : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[0] += (double)w0[u] * xv[u];
      a[1] += (double)(n > 1 ? w1[u] : VT(0)) * xv[u];
      a[2] += (double)(n > 2 ? w2[u] : VT(0)) * xv[u];
      a[3] += (double)(n > 3 ? w3[u] : VT(0)) * xv[u];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) a[q] = pm_group_sum<LANES>(a[q]);
}

// Runs of <= 4 consecutive columns, one per lane (k_spmv_ch, k_spmv_b_ch):
// a block's metadata, and one pass of UNR runs per lane.
struct PmRunMeta {
  int64_t vp, hp;
  int c, n, nh;
};

__device__ __forceinline__ PmRunMeta pm_run_meta(int64_t k, int64_t k1,
                                                 const int64_t *__restrict__ colptr,
                                                 const int64_t *__restrict__ valptr,
                                                 const int *__restrict__ nr,
                                                 const int64_t *__restrict__ chptr) {
  PmRunMeta m{0, 0, 0, 0, 0};
  if (k < k1) {
    m.vp = valptr[k];
    m.c = (int)(colptr[k + 1] - colptr[k]);
    m.n = nr[k];
    m.hp = chptr[k];
    m.nh = (int)(chptr[k + 1] - m.hp);
  }
  return m;
}

__device__ __forceinline__ unsigned int pm_lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned int)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
}

// one pass of UNR runs per lane over block m: d[] the descriptors, base the
// block columns before the pass; returns the columns the pass covered
// (NPF: the units of the next block's descriptors loaded ahead, UNR by default)
template <typename VT, int UNR, bool PREFETCH, int NPF = UNR>
__device__ __forceinline__ unsigned int pm_run_pass(int lane, int h0, const PmRunMeta &m,
                                                    unsigned int base, const unsigned int (&d)[UNR],
                                                    __amdgpu_buffer_rsrc_t vr, unsigned int rowb,
                                                    __amdgpu_buffer_rsrc_t xr,
                                                    __amdgpu_buffer_rsrc_t hrn,
                                                    unsigned int (&dn)[NPF], double a[4]) {
  // FP32 values stream with the non-temporal hint; FP64 values (two 16-byte
  // loads per row and run, into the same cache lines) without it: at 87^3 the
  // hint cost the FP64 product 13 % (3.71 against 3.23 ms) and left FP32 within 1 %
  constexpr int NT = sizeof(VT) == 4 ? 2 : 0;
  constexpr int VPR = sizeof(VT) == 4 ? 1 : 2;  // 16-byte loads per row and run
  unsigned int off[UNR];
  int len[UNR];
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const bool valid = h0 + lane + u * 64 < m.nh;
    const unsigned int l1 = d[u] & 3u;
    const unsigned long long bv = __ballot(valid);
    const unsigned long long b0 = __ballot(valid && (l1 & 1u));
    const unsigned long long b1 = __ballot(valid && (l1 & 2u));
    off[u] = base + pm_lanes_below(bv) + pm_lanes_below(b0) + 2u * pm_lanes_below(b1);
    len[u] = valid ? (int)l1 + 1 : 0;
    base += (unsigned int)(__popcll(bv) + __popcll(b0) + 2 * __popcll(b1));
  }
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  u4 w[UNR][4][VPR];
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const unsigned int o = off[u] * (unsigned int)sizeof(VT);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int h = 0; h < VPR; h++)
        w[u][q][h] = __builtin_bit_cast(
            u4, __builtin_amdgcn_raw_buffer_load_b128(vr, o + (unsigned int)q * rowb + 16u * h, 0, NT));
  }
  double xv[UNR][4];
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const unsigned int col = d[u] >> 2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double xc = pm_bload<double, 0>(xr, (col + (unsigned int)k) * 8u);
      xv[u][k] = k < len[u] ? xc : 0.0;
    }
  }
  if (PREFETCH) {
#pragma unroll
    for (int u = 0; u < NPF; u++)
      dn[u] = __builtin_amdgcn_raw_buffer_load_b32(hrn, (unsigned int)(lane + u * 64) * 4u, 0, NT);
  }
#pragma unroll
  for (int u = 0; u < UNR; u++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      VT v[4];
      __builtin_memcpy(v, &w[u][q][0], sizeof(v));
#pragma unroll
      for (int k = 0; k < 4; k++) a[q] += (double)v[k] * xv[u][k];
    }
  return base;
}
