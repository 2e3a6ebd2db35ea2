// Observables of a state on the device (the step after a ground-state solve):
//
//  * opposite-spin correlation G_updn(r) = (1/N) sum_i n_{i,up} n_{i+r,dn}
//    (HubbardModel*::opposite_spin_correlation, hubbard_model.h:40-53,
//    105-120, 184-200) for every lattice displacement r at once.  The operator
//    is diagonal in the occupation basis: <x|G(r)|x> = sum_(u,d) |x|^2 popc(s_u &
//    P_r(s_d)) / N with P_r the down string moved back by r (bit i of P_r(s) is
//    bit i + r of s), so one pass over |x|^2 with a table of P_r(s_d) for all r
//    and d gives every r; sums are fixed-order two-level reductions.
//  * Rayleigh quotient E = <x|H|x> / <x|x> and residual |Hx - E x| / |x|, the
//    latter summed directly from (Hx - E x) so that a small residual is not lost
//    to cancellation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "krylov.hpp"
#include "mb_internal.hpp"
#include "observables.hpp"

namespace mb {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kObsBlocks = 512;
constexpr int kMaxR = kMaxSites;  // displacements (one per site)

__device__ double wave_sum(double v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// partial[r * gridDim.x + blk] for r < N: sum |x|^2 popc(s_u & P_r(s_d));
// partial[kMaxR * gridDim.x + blk]: sum |x|^2.
__global__ __launch_bounds__(kBlock) void corr_partial_kernel(const uint32_t *__restrict__ s_up,
                                                              const uint32_t *__restrict__ sh,
                                                              const double *__restrict__ x,
                                                              int64_t Dup, int64_t Ddn, int N,
                                                              double *partial) {
  __shared__ double lds[kMaxR + 1][kBlock / kWave];
  double acc[kMaxR + 1];
#pragma unroll
  for (int r = 0; r <= kMaxR; ++r) acc[r] = 0.0;
  const int64_t D = Dup * Ddn;
  // rows u strided over blocks, columns d over threads: no index division
  for (int64_t u = blockIdx.x; u < Dup; u += gridDim.x) {
    const uint32_t su = s_up[u];
    for (int64_t d = threadIdx.x; d < Ddn; d += kBlock) {
      const int64_t e = u * Ddn + d;
      const double p = x[e] * x[e] + x[D + e] * x[D + e];
      acc[kMaxR] += p;
#pragma unroll
      for (int r = 0; r < kMaxR; ++r)
        if (r < N) acc[r] += p * (double)__builtin_popcount(su & sh[(int64_t)r * Ddn + d]);
    }
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int r = 0; r <= kMaxR; ++r) {
    const double v = wave_sum(acc[r]);
    if (lane == 0) lds[r][w] = v;
  }
  __syncthreads();
  if (threadIdx.x <= kMaxR) {
    double s = 0.0;
    for (int i = 0; i < kBlock / kWave; ++i) s += lds[threadIdx.x][i];
    partial[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// out[j] = sum of the nb partials of row j (one block per row, fixed order)
__global__ __launch_bounds__(kBlock) void rows_final_kernel(const double *partial, int nb,
                                                            double *out) {
  __shared__ double lds[kBlock / kWave];
  const double *p = partial + (int64_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) s += p[i];
  s = wave_sum(s);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock / kWave; ++i) t += lds[i];
    out[blockIdx.x] = t;
  }
}

// partial rows: 0 = Re <x|y>, 1 = |x|^2 ; with y = H x
__global__ __launch_bounds__(kBlock) void rayleigh_partial_kernel(const double *__restrict__ x,
                                                                  const double *__restrict__ y,
                                                                  int64_t n, double *partial) {
  __shared__ double lds[2][kBlock / kWave];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    a += x[i] * y[i];
    b += x[i] * x[i];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) {
    lds[0][w] = a;
    lds[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int i = 0; i < kBlock / kWave; ++i) s += lds[threadIdx.x][i];
    partial[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// partial[blk] = sum |y - E x|^2 (E read from the device)
__global__ __launch_bounds__(kBlock) void residual_partial_kernel(const double *__restrict__ x,
                                                                  const double *__restrict__ y,
                                                                  int64_t n, const double *sums,
                                                                  double *partial) {
  __shared__ double lds[kBlock / kWave];
  const double E = sums[0] / sums[1];
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double q = y[i] - E * x[i];
    a += q * q;
  }
  a = wave_sum(a);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kBlock / kWave; ++i) s += lds[i];
    partial[blockIdx.x] = s;
  }
}

}  // namespace

std::vector<double> opposite_spin_correlation(const Hubbard &h, const void *x, hipStream_t s) {
  if (h.ndim < 1) throw std::invalid_argument("the model has no lattice (made from a bond list)");
  const Sector &sec = *h.sec;
  const int N = sec.L;
  if (N > kMaxR) throw std::invalid_argument("at most 32 sites");
  // P_r(s) for every displacement r and down string s: bit i = bit (i + r) of s
  const int Lx = h.lsize[0], Ly = h.lsize[1], Lz = h.lsize[2];
  std::vector<uint32_t> sh((size_t)N * sec.d_dn);
  for (int r = 0; r < N; ++r) {
    const int rx = r % Lx, ry = (r / Lx) % Ly, rz = r / (Lx * Ly);
    std::vector<int> m(N);
    for (int z = 0; z < Lz; ++z)
      for (int y = 0; y < Ly; ++y)
        for (int xx = 0; xx < Lx; ++xx)
          m[xx + Lx * y + Lx * Ly * z] =
              ((xx + rx) % Lx) + Lx * ((y + ry) % Ly) + Lx * Ly * ((z + rz) % Lz);
    for (int64_t d = 0; d < sec.d_dn; ++d) {
      const uint32_t sd = sec.s_dn[d];
      uint32_t p = 0;
      for (int i = 0; i < N; ++i) p |= ((sd >> m[i]) & 1u) << i;
      sh[(size_t)r * sec.d_dn + d] = p;
    }
  }
  DevBuf dsh = upload(sh);
  DevBuf part((size_t)(kMaxR + 1) * kObsBlocks * sizeof(double));
  DevBuf sums((size_t)(kMaxR + 1) * sizeof(double));
  hipLaunchKernelGGL(corr_partial_kernel, dim3(kObsBlocks), dim3(kBlock), 0, s,
                     sec.d_s_up.as<uint32_t>(), dsh.as<uint32_t>(), (const double *)x, sec.d_up,
                     sec.d_dn, N, part.as<double>());
  hipLaunchKernelGGL(rows_final_kernel, dim3(kMaxR + 1), dim3(kBlock), 0, s,
                     (const double *)part.as<double>(), kObsBlocks, sums.as<double>());
  MB_HIP(hipGetLastError());
  std::vector<double> hs(kMaxR + 1);
  MB_HIP(hipMemcpyAsync(hs.data(), sums.p, sizeof(double) * (kMaxR + 1), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (!(hs[kMaxR] > 0.0)) throw std::runtime_error("Input vector has zero norm");
  std::vector<double> g(N);
  for (int r = 0; r < N; ++r) g[r] = hs[r] / ((double)N * hs[kMaxR]);
  return g;
}

void energy_residual(const HvOp &op, const void *x, double *energy, double *residual,
                     hipStream_t s) {
  const int64_t D = op.dim();
  DevBuf y((size_t)D * 16);
  DevBuf part((size_t)2 * kObsBlocks * sizeof(double));
  DevBuf sums(4 * sizeof(double));
  EpiArgs a;
  op.hv((const double *)x, y.as<double>(), Epi::Plain, a, s);
  hipLaunchKernelGGL(rayleigh_partial_kernel, dim3(kObsBlocks), dim3(kBlock), 0, s, (const double *)x,
                     (const double *)y.as<double>(), 2 * D, part.as<double>());
  hipLaunchKernelGGL(rows_final_kernel, dim3(2), dim3(kBlock), 0, s, (const double *)part.as<double>(),
                     kObsBlocks, sums.as<double>());
  hipLaunchKernelGGL(residual_partial_kernel, dim3(kObsBlocks), dim3(kBlock), 0, s,
                     (const double *)x, (const double *)y.as<double>(), 2 * D,
                     (const double *)sums.as<double>(), part.as<double>());
  hipLaunchKernelGGL(rows_final_kernel, dim3(1), dim3(kBlock), 0, s, (const double *)part.as<double>(),
                     kObsBlocks, sums.as<double>() + 2);
  MB_HIP(hipGetLastError());
  double hs[3];
  MB_HIP(hipMemcpyAsync(hs, sums.p, sizeof(hs), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (!(hs[1] > 0.0)) throw std::runtime_error("Input vector has zero norm");
  *energy = hs[0] / hs[1];
  *residual = std::sqrt(hs[2] / hs[1]);
}

}  // namespace mb
