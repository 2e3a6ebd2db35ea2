// Sparse export of a compiled expression (expression.cpp): every element emits
// its raw contributions (output row i, source column j, amplitude) keyed
// col_ref * D + row_ref in the reference's basis order; a radix sort by key
// and a merge of equal keys give the nonzeros in arma::SpMat iteration order
// (column-major, rows ascending), and entries whose merged value is below
// 1000 eps are dropped, as ExpressionBase::add_to erases them
// (expression_base.h:85-97).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "expression.hpp"

namespace mb {

namespace {

constexpr int kBlk = 256;
constexpr double kTol = 1000.0 * 2.220446049250313e-16;

struct Binom33 {
  uint64_t c[kMaxSites + 1][kMaxSites + 1];
};

__device__ __forceinline__ int64_t colex(const Binom33 &B, uint32_t bits) {
  int64_t r = 0;
  int k = 0;
  while (bits) {
    const int p = __builtin_ctz(bits);
    r += (int64_t)B.c[p][k + 1];
    ++k;
    bits &= bits - 1;
  }
  return r;
}

// Visits the contributions of element (w, tid): f(source element, re, im).
template <typename F>
__device__ __forceinline__ void for_each_contribution(const ExprTables &t, const Binom33 &B, int g,
                                                      int c, F &&f) {
  const int h = t.row_dg0[g] + c;
  const uint32_t d = t.dn_bits[h];
  const int64_t base = t.row_base[g], i = base + c;
  f(i, t.diag_up[2 * g] + t.diag_dn[2 * h], t.diag_up[2 * g + 1] + t.diag_dn[2 * h + 1]);
  for (int64_t e = t.uo_ptr[g]; e < t.uo_ptr[g + 1]; ++e)
    f(t.uo_base[e] + c, t.uo_amp[2 * e], t.uo_amp[2 * e + 1]);
  for (int64_t e = t.do_ptr[h]; e < t.do_ptr[h + 1]; ++e)
    f(base + t.do_col[e], t.do_amp[2 * e], t.do_amp[2 * e + 1]);
  for (int64_t e = t.mx_ptr[g]; e < t.mx_ptr[g + 1]; ++e) {
    const int k = t.mx_grp[e];
    const int grp = t.mx_dgrp[e];
    const double su = t.mx_sign[e];
    for (int l = t.dp_ptr[k]; l < t.dp_ptr[k + 1]; ++l) {
      if ((d & t.dp_touched[l]) != t.dp_out[l]) continue;
      const uint32_t d2 = d ^ t.dp_flip[l];
      int64_t col;
      if (t.momentum) {
        if (t.dn_mom[d2] != grp) continue;
        col = t.dn_loc[d2];
      } else {
        col = colex(B, d2);
      }
      const double sg = (__builtin_popcount(d & t.dp_parity[l]) & 1) ? -su : su;
      f(t.mx_base[e] + col, sg * t.dp_c[2 * l], sg * t.dp_c[2 * l + 1]);
    }
  }
}

__device__ void load_b(Binom33 &B, const ExprTables &t) {
  if (t.momentum) return;
  for (int i = threadIdx.x; i < (kMaxSites + 1) * (kMaxSites + 1); i += blockDim.x)
    (&B.c[0][0])[i] = t.binom[i];
  __syncthreads();
}

__global__ __launch_bounds__(kBlk) void export_count_kernel(ExprTables t, int64_t *__restrict__ cnt) {
  __shared__ Binom33 B;
  load_b(B, t);
  const int w = blockIdx.x, g = t.wg_row[w], c = t.wg_c0[w] + threadIdx.x;
  if (c >= t.row_ncols[g]) return;
  int64_t n = 0;
  for_each_contribution(t, B, g, c, [&](int64_t, double, double) { ++n; });
  cnt[t.row_base[g] + c] = n;
}

__global__ __launch_bounds__(kBlk) void export_emit_kernel(ExprTables t, const int64_t *__restrict__ off,
                                                           const int64_t *__restrict__ ref,
                                                           uint64_t *__restrict__ keys,
                                                           int64_t *__restrict__ idx,
                                                           double *__restrict__ vals) {
  __shared__ Binom33 B;
  load_b(B, t);
  const int w = blockIdx.x, g = t.wg_row[w], c = t.wg_c0[w] + threadIdx.x;
  if (c >= t.row_ncols[g]) return;
  const int64_t i = t.row_base[g] + c;
  const uint64_t row_ref = (uint64_t)ref[i], D = (uint64_t)t.D;
  int64_t o = off[i];
  for_each_contribution(t, B, g, c, [&](int64_t j, double re, double im) {
    keys[o] = (uint64_t)ref[j] * D + row_ref;
    idx[o] = o;
    vals[2 * o] = re;
    vals[2 * o + 1] = im;
    ++o;
  });
}

// run heads of the sorted keys: sum the run, keep it if not below 1000 eps
__global__ __launch_bounds__(kBlk) void export_merge_kernel(int64_t n, const uint64_t *__restrict__ keys,
                                                            const int64_t *__restrict__ order,
                                                            const double *__restrict__ vals,
                                                            int64_t *__restrict__ keep,
                                                            double *__restrict__ sums) {
  for (int64_t p = (int64_t)blockIdx.x * kBlk + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlk) {
    int64_t k = 0;
    if (p == 0 || keys[p] != keys[p - 1]) {
      double re = 0.0, im = 0.0;
      for (int64_t q = p; q < n && keys[q] == keys[p]; ++q) {
        re += vals[2 * order[q]];
        im += vals[2 * order[q] + 1];
      }
      sums[2 * p] = re;
      sums[2 * p + 1] = im;
      k = (re * re + im * im >= kTol * kTol) ? 1 : 0;
    }
    keep[p] = k;
  }
}

__global__ __launch_bounds__(kBlk) void export_compact_kernel(int64_t n, const uint64_t *__restrict__ keys,
                                                              const int64_t *__restrict__ keep,
                                                              const int64_t *__restrict__ pos,
                                                              const double *__restrict__ sums,
                                                              uint64_t *__restrict__ out_keys,
                                                              double *__restrict__ out_vals) {
  for (int64_t p = (int64_t)blockIdx.x * kBlk + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlk) {
    if (!keep[p]) continue;
    const int64_t q = pos[p];
    out_keys[q] = keys[p];
    out_vals[2 * q] = sums[2 * p];
    out_vals[2 * q + 1] = sums[2 * p + 1];
  }
}

// Rectangular export: the reference index of output word s2 in the row basis,
// or -1 when it is not a state of that basis.  Words have 2L bits (monomial
// orbitals are below L), so both halves index the 2^L lookups.
__device__ __forceinline__ int64_t rect_target(const RectSide &r, uint64_t s2) {
  const uint32_t u2 = (uint32_t)(s2 & ((uint64_t{1} << r.L) - 1)), d2 = (uint32_t)(s2 >> r.L);
  const int32_t g = r.up_row[u2];
  if (g < 0 || r.dn_grp[d2] != r.row_cgrp[g]) return -1;
  return r.ref[r.row_base[g] + r.dn_loc[d2]];
}

// One thread per column-basis element (row blockIdx.x, column blockIdx.y * kBlk
// + tid): counts (EMIT false) or emits its contributions, one per monomial that
// applies and lands in the row basis.
template <bool EMIT>
__global__ __launch_bounds__(kBlk) void rect_kernel(RectSide row, RectSide col,
                                                    const RectMono *__restrict__ mono, int n_mono,
                                                    const int64_t *__restrict__ off,
                                                    int64_t *__restrict__ cnt,
                                                    uint64_t *__restrict__ keys,
                                                    int64_t *__restrict__ idx,
                                                    double *__restrict__ vals) {
  const int64_t g = blockIdx.x;
  const int c = (int)blockIdx.y * kBlk + (int)threadIdx.x;
  if (c >= col.row_nc[g]) return;
  const int64_t pos = col.row_base[g] + c;
  const uint64_t s =
      (uint64_t)col.row_up[g] | ((uint64_t)col.dn_bits[col.row_dg0[g] + c] << col.L);
  int64_t o = EMIT ? off[pos] : 0;
  const uint64_t cref = EMIT ? (uint64_t)col.ref[pos] : 0;
  for (int m = 0; m < n_mono; ++m) {
    const RectMono a = mono[m];
    if ((s & a.touched) != a.in) continue;
    const int64_t rr = rect_target(row, s ^ (a.in ^ a.out));
    if (rr < 0) continue;
    if constexpr (EMIT) {
      const double sg = (__builtin_popcountll(s & a.parity) & 1) ? -1.0 : 1.0;
      keys[o] = cref * (uint64_t)row.D + (uint64_t)rr;
      idx[o] = o;
      vals[2 * o] = sg * a.re;
      vals[2 * o + 1] = sg * a.im;
    }
    ++o;
  }
  if constexpr (!EMIT) cnt[pos] = o;
}

unsigned grid_of(int64_t n) {
  const int64_t g = (n + kBlk - 1) / kBlk;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 16));
}

// exclusive prefix sum of n int64 values; returns the total
int64_t exclusive_sum(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
  if (n == 0) return 0;
  size_t tmp = 0;
  MB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
  DevBuf t(std::max<size_t>(tmp, 1));
  MB_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, (int)n, s));
  int64_t last_off = 0, last_in = 0;
  MB_HIP(hipMemcpyAsync(&last_off, out + n - 1, 8, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  return last_off + last_in;
}

}  // namespace

int64_t sort_merge(int64_t n, uint64_t key_max, DevBuf &k0, DevBuf &k1, DevBuf &i0, DevBuf &i1,
                   DevBuf &v0, DevBuf &keys, DevBuf &vals, hipStream_t s);

int64_t expr_export(const ExprTables &t, const int64_t *ref_of_pos, DevBuf &keys, DevBuf &vals,
                    hipStream_t s) {
  const int64_t D = t.D;
  if (D == 0 || t.nwg == 0) return 0;
  if (D >= (int64_t{1} << 31)) throw std::invalid_argument("sector too large to export");
  DevBuf cnt((size_t)D * 8), off((size_t)D * 8);
  hipLaunchKernelGGL(export_count_kernel, dim3((unsigned)t.nwg), dim3(kBlk), 0, s, t,
                     cnt.as<int64_t>());
  MB_HIP(hipGetLastError());
  const int64_t n = exclusive_sum(cnt.as<int64_t>(), off.as<int64_t>(), D, s);
  if (n >= (int64_t{1} << 31)) throw std::invalid_argument("too many matrix entries to export");
  DevBuf k0((size_t)n * 8), k1((size_t)n * 8), i0((size_t)n * 8), i1((size_t)n * 8),
      v0((size_t)n * 16);
  hipLaunchKernelGGL(export_emit_kernel, dim3((unsigned)t.nwg), dim3(kBlk), 0, s, t,
                     off.as<int64_t>(), ref_of_pos, k0.as<uint64_t>(), i0.as<int64_t>(),
                     v0.as<double>());
  MB_HIP(hipGetLastError());
  return sort_merge(n, (uint64_t)D * (uint64_t)D - 1, k0, k1, i0, i1, v0, keys, vals, s);
}

// Sorts n raw (key, value) contributions by key (stable: equal keys keep their
// emission order), sums each run and keeps the sums not below 1000 eps: keys /
// vals get the kept entries in key order; returns their number.
int64_t sort_merge(int64_t n, uint64_t key_max, DevBuf &k0, DevBuf &k1, DevBuf &i0, DevBuf &i1,
                   DevBuf &v0, DevBuf &keys, DevBuf &vals, hipStream_t s) {
  int bits = 1;
  while (bits < 64 && (key_max >> bits)) ++bits;
  size_t tmp = 0;
  MB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                            i0.as<int64_t>(), i1.as<int64_t>(), (int)n, 0, bits,
                                            s));
  {
    DevBuf tb(std::max<size_t>(tmp, 1));
    MB_HIP(hipcub::DeviceRadixSort::SortPairs(tb.p, tmp, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                              i0.as<int64_t>(), i1.as<int64_t>(), (int)n, 0,
                                              bits, s));
  }
  // merge runs: k0 / i0 reused as keep flags and positions, v sums in a new buffer
  DevBuf sums((size_t)n * 16), keep((size_t)n * 8), pos((size_t)n * 8);
  hipLaunchKernelGGL(export_merge_kernel, dim3(grid_of(n)), dim3(kBlk), 0, s, n,
                     k1.as<uint64_t>(), i1.as<int64_t>(), v0.as<double>(), keep.as<int64_t>(),
                     sums.as<double>());
  MB_HIP(hipGetLastError());
  const int64_t m = exclusive_sum(keep.as<int64_t>(), pos.as<int64_t>(), n, s);
  keys.alloc((size_t)std::max<int64_t>(m, 1) * 8);
  vals.alloc((size_t)std::max<int64_t>(m, 1) * 16);
  hipLaunchKernelGGL(export_compact_kernel, dim3(grid_of(n)), dim3(kBlk), 0, s, n,
                     k1.as<uint64_t>(), keep.as<int64_t>(), pos.as<int64_t>(), sums.as<double>(),
                     keys.as<uint64_t>(), vals.as<double>());
  MB_HIP(hipGetLastError());
  MB_HIP(hipStreamSynchronize(s));
  return m;
}

int64_t rect_export(const RectSide &row, const RectSide &col, const RectMono *monos, int n_mono,
                    DevBuf &keys, DevBuf &vals, hipStream_t s) {
  const int64_t Dc = col.D;
  if (Dc == 0 || row.D == 0 || n_mono == 0 || col.nrows == 0 || col.max_nc == 0) return 0;
  if (Dc >= (int64_t{1} << 31) || col.nrows >= (int64_t{1} << 31))
    throw std::invalid_argument("column basis too large to export");
  DevBuf cnt((size_t)Dc * 8), off((size_t)Dc * 8);
  MB_HIP(hipMemsetAsync(cnt.p, 0, (size_t)Dc * 8, s));
  const dim3 grid((unsigned)col.nrows, (unsigned)((col.max_nc + kBlk - 1) / kBlk));
  hipLaunchKernelGGL(rect_kernel<false>, grid, dim3(kBlk), 0, s, row, col, monos, n_mono,
                     nullptr, cnt.as<int64_t>(), nullptr, nullptr, nullptr);
  MB_HIP(hipGetLastError());
  const int64_t n = exclusive_sum(cnt.as<int64_t>(), off.as<int64_t>(), Dc, s);
  if (n == 0) return 0;
  if (n >= (int64_t{1} << 31)) throw std::invalid_argument("too many matrix entries to export");
  DevBuf k0((size_t)n * 8), k1((size_t)n * 8), i0((size_t)n * 8), i1((size_t)n * 8),
      v0((size_t)n * 16);
  hipLaunchKernelGGL(rect_kernel<true>, grid, dim3(kBlk), 0, s, row, col, monos, n_mono,
                     off.as<int64_t>(), nullptr, k0.as<uint64_t>(), i0.as<int64_t>(),
                     v0.as<double>());
  MB_HIP(hipGetLastError());
  return sort_merge(n, (uint64_t)Dc * (uint64_t)row.D - 1, k0, k1, i0, i1, v0, keys, vals, s);
}

}  // namespace mb
