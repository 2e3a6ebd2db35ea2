// Device helpers shared by the batched engines (batch.hip, lane.hip): H.v of
// one tile of one instance, both planes at once, and a fixed-order block sum.
#pragma once

#include <hip/hip_runtime.h>

#include "batch.hpp"

namespace mb {
namespace bdev {

constexpr int kWave = 64;

__device__ __forceinline__ int popc(uint32_t v) { return __builtin_popcount(v); }

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) lds[tid / kWave] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int w = 0; w < BS / kWave; ++w) r += lds[w];
  __syncthreads();
  return r;
}

// (H x)[e] for both planes of one instance, x planar with plane stride D.
__device__ __forceinline__ void hv_pair(const HvTables &t, const double *__restrict__ x, int64_t D,
                                        int64_t e, double &re, double &im) {
  const int64_t Ddn = t.d_dn;
  const int64_t u = e / Ddn, d = e - u * Ddn;
  double dg = t.diagU[popc(t.s_up[u] & t.s_dn[d])];
  if (t.has_onsite) dg = dg + t.up_on[u] + t.dn_on[d];
  const double *xr = x, *xi = x + D;
  double ur = 0.0, ui = 0.0;
  for (int32_t k = t.up_ptr[u]; k < t.up_ptr[u + 1]; ++k) {
    const int64_t f = (int64_t)t.up_tgt[k] * Ddn + d;
    const double a = t.up_amp[k];
    ur += a * xr[f];
    ui += a * xi[f];
  }
  double ar = dg * xr[e], ai = dg * xi[e];
  for (int32_t k = t.dn_ptr[d]; k < t.dn_ptr[d + 1]; ++k) {
    const int64_t f = u * Ddn + t.dn_tgt[k];
    const double a = t.dn_amp[k];
    ar += a * xr[f];
    ai += a * xi[f];
  }
  re = ur + ar;
  im = ui + ai;
}

// Row-tiled variant for instances with the compact down table: the tile is
// rows [u0, u0 + R) (R * D_dn elements); both planes of those rows are staged in
// LDS and down hops read LDS; up hops read whole neighbour rows (coalesced).
template <int BS>
__device__ __forceinline__ double hv_rows_tile(const BInst &I, const BTile &tl,
                                               const double *__restrict__ v,
                                               const double *__restrict__ vp, double *__restrict__ w,
                                               double bp, double *lds_row) {
  const HvTables &t = I.t;
  const int64_t D = I.D, Ddn = t.d_dn;
  const int64_t u0 = tl.e0 / Ddn;
  const int n = tl.n;
  double *lr = lds_row, *li = lds_row + n;
  for (int j = threadIdx.x; j < n; j += BS) {
    lr[j] = v[tl.e0 + j];
    li[j] = v[D + tl.e0 + j];
  }
  __syncthreads();
  const double h = t.mags[0];
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += BS) {
    const int64_t e = tl.e0 + j;
    const int64_t r = j / Ddn, d = j - r * Ddn, u = u0 + r;
    double dg = t.diagU[popc(t.s_up[u] & t.s_dn[d])];
    if (t.has_onsite) dg = dg + t.up_on[u] + t.dn_on[d];
    double ur = 0.0, ui = 0.0;
    for (int32_t k = t.up_ptr[u]; k < t.up_ptr[u + 1]; ++k) {
      const int64_t f = (int64_t)t.up_tgt[k] * Ddn + d;
      const double a = t.up_amp[k];
      ur += a * v[f];
      ui += a * v[D + f];
    }
    double ar = dg * lr[j], ai = dg * li[j];
    const int cnt = t.dn_cnt[d];
    const uint16_t *ct = t.dn_comp + d * 32;
    const uint32_t sm = t.dn_csign[d];
    const int64_t rb = r * Ddn;
    for (int k = 0; k < cnt; ++k) {
      const double a = ((sm >> k) & 1u) ? -h : h;
      const int64_t q = rb + ct[k];
      ar += a * lr[q];
      ai += a * li[q];
    }
    double re = ur + ar, im = ui + ai;
    if (bp != 0.0) {
      re -= bp * vp[e];
      im -= bp * vp[D + e];
    }
    w[e] = re;
    w[D + e] = im;
    s += lr[j] * re + li[j] * im;
  }
  return s;
}

}  // namespace bdev
}  // namespace mb
