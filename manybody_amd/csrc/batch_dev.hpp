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
// rows [u0, u0 + R) (R * D_dn elements).  LDS holds both planes of those rows
// (row stride D_dn + 1: a zero slot ends each row, the target of the unused
// slots of the compact table) and the up-hop lists of the R rows; down hops
// read LDS in groups of 8 slots up to the wave's largest hop count (no per-lane
// predicate, signs by flipping the sign bit), up hops gather whole neighbour
// rows of x from global memory (coalesced along d).  The input is x with
// v = sigma x (the normalisation of the previous Lanczos step applied on the
// fly); v of the tile is written to vout when it is given.  Returns the tile's
// partial v . w for w = H v - bp vp.  LDS layout: rows_tile_lds (batch.hpp).
__device__ __forceinline__ double flip_sign(double v, uint32_t sbit) {
  return __hiloint2double(__double2hiint(v) ^ (int)sbit, __double2loint(v));
}

template <int BS>
__device__ __forceinline__ double hv_rows_tile(const BInst &I, const BTile &tl,
                                               const double *__restrict__ x, double sigma,
                                               double *__restrict__ vout,
                                               const double *__restrict__ vp, double *__restrict__ w,
                                               double bp, double *lds_row) {
  constexpr int K = 32, G = 8;
  const HvTables &t = I.t;
  const int64_t D = I.D;
  const int Ddn = (int)t.d_dn, S = Ddn + 1, MU = t.max_up_row;
  const int64_t u0 = tl.e0 / Ddn;
  const int n = tl.n, R = n / Ddn;
  double *lr = lds_row, *li = lds_row + R * S, *ua = li + R * S;
  int *ut = reinterpret_cast<int *>(ua + R * MU), *un = ut + R * MU;
  for (int j = threadIdx.x; j < n; j += BS) {
    const int r = j / Ddn, d = j - r * Ddn;
    const double a = sigma * x[tl.e0 + j], b = sigma * x[D + tl.e0 + j];
    lr[r * S + d] = a;
    li[r * S + d] = b;
    if (vout) {
      vout[tl.e0 + j] = a;
      vout[D + tl.e0 + j] = b;
    }
  }
  for (int r = threadIdx.x; r < R; r += BS) {
    lr[r * S + Ddn] = 0.0;
    li[r * S + Ddn] = 0.0;
    const int64_t u = u0 + r;
    un[r] = t.up_ptr[u + 1] - t.up_ptr[u];
  }
  for (int q = threadIdx.x; q < R * MU; q += BS) {
    const int r = q / MU, k = q - r * MU;
    const int64_t u = u0 + r;
    const int32_t b = t.up_ptr[u];
    if (k < t.up_ptr[u + 1] - b) {
      ut[q] = t.up_tgt[b + k];
      ua[q] = t.up_amp[b + k];
    }
  }
  __syncthreads();
  const double h = t.mags[0];
  const int lane = threadIdx.x & (kWave - 1);
  double s = 0.0;
  for (int j0 = (int)threadIdx.x - lane; j0 < n; j0 += BS) {
    const int j = j0 + lane;
    const bool ok = j < n;
    const int jj = ok ? j : j0;
    const int r = jj / Ddn, d = jj - r * Ddn;
    const int64_t u = u0 + r, e = tl.e0 + jj;
    const uint32_t sm = t.dn_csign[d];
    int wm = t.dn_cnt[d];
    for (int off = kWave / 2; off > 0; off >>= 1) wm = max(wm, __shfl_xor(wm, off, kWave));
    wm = __builtin_amdgcn_readfirstlane(wm);
    uint4 q[K / 8];
#pragma unroll
    for (int c = 0; c < K / 8; ++c) q[c] = c * 8 < wm ? comp_chunk(t.dn_comp, d, c) : make_uint4(0, 0, 0, 0);
    double dg = t.diagU[popc(t.s_up[u] & t.s_dn[d])];
    if (t.has_onsite) dg = dg + t.up_on[u] + t.dn_on[d];
    // up hops: rows of x from global memory
    double ur = 0.0, ui = 0.0;
    const int nu = un[r];
    for (int k = 0; k < nu; ++k) {
      const int64_t f = (int64_t)ut[r * MU + k] * Ddn + d;
      const double a = ua[r * MU + k];
      ur += a * x[f];
      ui += a * x[D + f];
    }
    // down hops: LDS
    const double *rr = lr + r * S, *ri = li + r * S;
    double hr0 = 0.0, hr1 = 0.0, hi0 = 0.0, hi1 = 0.0;
#pragma unroll
    for (int g = 0; g < K / G; ++g) {
      if (g * G < wm) {
        double vr[G], vi[G];
#pragma unroll
        for (int c = 0; c < G; ++c) {
          const int k = g * G + c;
          const uint32_t wd = reinterpret_cast<const uint32_t *>(&q[k / 8])[(k % 8) / 2];
          const uint32_t idx = (k & 1) ? (wd >> 16) : (wd & 0xFFFFu);
          vr[c] = rr[idx];
          vi[c] = ri[idx];
        }
#pragma unroll
        for (int c = 0; c < G; ++c) {
          const int k = g * G + c;
          const uint32_t sb = (sm << (31 - k)) & 0x80000000u;
          if (c & 1) {
            hr1 += flip_sign(vr[c], sb);
            hi1 += flip_sign(vi[c], sb);
          } else {
            hr0 += flip_sign(vr[c], sb);
            hi0 += flip_sign(vi[c], sb);
          }
        }
      }
    }
    const double xr = rr[d], xi = ri[d];
    double re = sigma * ur + (dg * xr + h * (hr0 + hr1));
    double im = sigma * ui + (dg * xi + h * (hi0 + hi1));
    if (ok) {
      if (bp != 0.0) {
        re -= bp * vp[e];
        im -= bp * vp[D + e];
      }
      w[e] = re;
      w[D + e] = im;
      s += xr * re + xi * im;
    }
  }
  return s;
}

}  // namespace bdev
}  // namespace mb
