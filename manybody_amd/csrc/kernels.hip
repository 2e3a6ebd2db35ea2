// HIP kernels for gfx950 (CDNA4, wave64): matrix-free Hubbard H.v, the
// reference-order <-> product-order permutation, fused Lanczos vector steps,
// Chebyshev (KPM) steps and seeded vector fills.
//
// Device layout: a state vector is the D_up x D_dn matrix x[u * D_dn + d]
// (u, d = combinatorial ranks of the up / down bitstrings).  Up hops move whole
// rows with one sign per row pair; down hops gather inside a row.  Reductions
// are two-level and atomic-free (fixed grid, fixed tree), so results do not
// depend on scheduling.
#include <hip/hip_runtime.h>

#include "krylov.hpp"
#include "mb_internal.hpp"

namespace mb {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

__device__ __forceinline__ int popc(uint32_t v) { return __builtin_popcount(v); }

// ------------------------------------------------------------ reductions

// Fixed-order block reduction of one double (kBlock threads).
__device__ double block_sum(double v, double *lds) {
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) lds[tid / kWave] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int w = 0; w < kBlock / kWave; ++w) r += lds[w];
  __syncthreads();
  return r;  // valid in thread 0
}

// ------------------------------------------------------------ H.v core

template <typename R>
struct Tabs;
template <>
struct Tabs<double> {
  __device__ static const double *up_amp(const HvTables &t) { return t.up_amp; }
  __device__ static const double *dn_amp(const HvTables &t) { return t.dn_amp; }
  __device__ static const double *up_on(const HvTables &t) { return t.up_on; }
  __device__ static const double *dn_on(const HvTables &t) { return t.dn_on; }
  __device__ static const double *diag(const HvTables &t) { return t.diagU; }
};
template <>
struct Tabs<float> {
  __device__ static const float *up_amp(const HvTables &t) { return t.up_ampf; }
  __device__ static const float *dn_amp(const HvTables &t) { return t.dn_ampf; }
  __device__ static const float *up_on(const HvTables &t) { return t.up_onf; }
  __device__ static const float *dn_on(const HvTables &t) { return t.dn_onf; }
  __device__ static const float *diag(const HvTables &t) { return t.diagUf; }
};

template <typename R>
struct C2 {
  R re, im;
};

// (H x)[e] for element e = u * D_dn + d of complex vector r (element-major
// multi-vector layout x[e * NV + r]).  Summation order: diagonal, up hops in
// table order, down hops in table order.
template <typename R, int NV>
__device__ __forceinline__ C2<R> hv_elem(const HvTables &t, const C2<R> *__restrict__ x,
                                         int64_t u, int64_t d, int r) {
  const int64_t Ddn = t.d_dn;
  const int64_t e = u * Ddn + d;
  R dg = Tabs<R>::diag(t)[popc(t.s_up[u] & t.s_dn[d])];
  if (t.has_onsite) dg = dg + Tabs<R>::up_on(t)[u] + Tabs<R>::dn_on(t)[d];
  const C2<R> xe = x[e * NV + r];
  R ar = dg * xe.re, ai = dg * xe.im;
  const int32_t ub = t.up_ptr[u], ue = t.up_ptr[u + 1];
  const R *uamp = Tabs<R>::up_amp(t);
  for (int32_t p = ub; p < ue; ++p) {
    const C2<R> v = x[((int64_t)t.up_tgt[p] * Ddn + d) * NV + r];
    const R a = uamp[p];
    ar += a * v.re;
    ai += a * v.im;
  }
  const int32_t db = t.dn_ptr[d], de = t.dn_ptr[d + 1];
  const R *damp = Tabs<R>::dn_amp(t);
  for (int32_t p = db; p < de; ++p) {
    const C2<R> v = x[(u * Ddn + t.dn_tgt[p]) * NV + r];
    const R a = damp[p];
    ar += a * v.re;
    ai += a * v.im;
  }
  return {ar, ai};
}

template <typename R, int NV>
__global__ __launch_bounds__(kBlock) void hv_kernel(HvTables t, const C2<R> *__restrict__ x,
                                                    C2<R> *__restrict__ y) {
  const int64_t total = t.d_up * t.d_dn * NV;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += stride) {
    const int64_t e = g / NV;
    const int r = (int)(g - e * NV);
    const int64_t u = e / t.d_dn;
    const int64_t d = e - u * t.d_dn;
    y[g] = hv_elem<R, NV>(t, x, u, d, r);
  }
}

int grid_for(int64_t work, int cap) {
  int64_t g = (work + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

constexpr int kGridCap = 1 << 20;

// ------------------------------------------------------------ permutation

struct BinomLds {
  uint64_t c[kMaxSites + 1][kMaxSites + 1];
};

__device__ void load_binom(BinomLds &b, const uint64_t *g) {
  for (int i = threadIdx.x; i < (kMaxSites + 1) * (kMaxSites + 1); i += blockDim.x)
    (&b.c[0][0])[i] = g[i];
  __syncthreads();
}

__device__ __forceinline__ uint64_t cb(const BinomLds &b, int n, int k) {
  return (n < 0 || k < 0 || k > n) ? 0ull : b.c[n][k];
}

// Reference basis index (DFS position, basis.cpp:267-302) of (up, dn); the
// same counting as Sector::ref_index on the host.
__device__ int64_t ref_rank(const BinomLds &B, int L, int n_up, int n_dn, uint32_t up,
                            uint32_t dn) {
  int64_t rank = 0;
  int U = n_up, D = n_dn;
  int prev_o = -1, prev_s = -1;
  for (int o = 0; o < L; ++o) {
    const uint32_t bit = 1u << o;
    for (int s = 0; s < 2; ++s) {
      if (!((s == 0 ? up : dn) & bit)) continue;
      if (U > 0) {
        const int hi = (s == 0) ? o - 1 : L - 1;
        for (int i = prev_o + 1; i <= hi; ++i)
          rank += (int64_t)(cb(B, L - 1 - i, U - 1) * cb(B, L - i, D));
      }
      if (s == 1 && D > 0) {
        int lo = (prev_s == 0) ? prev_o : prev_o + 1;
        if (lo < 0) lo = 0;
        for (int i = lo; i < o; ++i) rank += (int64_t)(cb(B, L - 1 - i, U) * cb(B, L - 1 - i, D - 1));
      }
      if (s == 0) --U; else --D;
      prev_o = o;
      prev_s = s;
    }
  }
  return rank;
}

template <bool kGather>
__global__ __launch_bounds__(kBlock) void perm_kernel(PermTables p, const double2 *__restrict__ src,
                                                      double2 *__restrict__ dst) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t total = p.d_up * p.d_dn;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += stride) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const int64_t r = ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    if (r < 0 || r >= total) continue;  // unreachable for a valid sector; never index out of range
    if (kGather)
      dst[e] = src[r];
    else
      dst[r] = src[e];
  }
}

// SplitMix64 finalizer over a Weyl sequence: the value for counter c of seed s.
__device__ __forceinline__ uint64_t mix64(uint64_t s, uint64_t c) {
  uint64_t z = s + (c + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit_open(uint64_t z) {  // (0, 1]
  return ((double)(z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double unit_half(uint64_t z) {  // [0, 1)
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kBlock) void fill_gaussian_kernel(PermTables p, uint64_t seed,
                                                               double2 *__restrict__ x) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t total = p.d_up * p.d_dn;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += stride) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const uint64_t i = (uint64_t)ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    const double u1 = unit_open(mix64(seed, 2 * i));
    const double u2 = unit_half(mix64(seed, 2 * i + 1));
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    x[e] = make_double2(rad * cos(th), rad * sin(th));
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void fill_phase_kernel(PermTables p, uint64_t seed, int nvec,
                                                            C2<R> *__restrict__ x) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t total = p.d_up * p.d_dn;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += stride) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const uint64_t i = (uint64_t)ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    for (int r = 0; r < nvec; ++r) {
      const double ph = 6.283185307179586 * unit_half(mix64(seed, i * (uint64_t)nvec + r));
      x[e * nvec + r] = {(R)cos(ph), (R)sin(ph)};
    }
  }
}

// ------------------------------------------------------------ Lanczos

constexpr int kLanczosGrid = 4096;

__global__ __launch_bounds__(kBlock) void norm_partial_kernel(const double *__restrict__ x,
                                                              int64_t n, double *partial) {
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += x[i] * x[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__device__ double final_sum(const double *partial, int n, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += partial[i];
  return block_sum(s, lds);
}

__global__ __launch_bounds__(kBlock) void sqrt_final_kernel(const double *partial, int n,
                                                            double *out) {
  __shared__ double lds[kBlock / kWave];
  const double s = final_sum(partial, n, lds);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}

__global__ __launch_bounds__(kBlock) void scale_div_kernel(const double *__restrict__ x,
                                                           const double *den,
                                                           double *__restrict__ y, int64_t n) {
  const double dv = den[0];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = x[i] / dv;
}

// w = H v - beta_{step-1} v_prev ; partial[b] = sum Re(conj(v) w)
__global__ __launch_bounds__(kBlock) void lanczos_hv_kernel(HvTables t,
                                                            const C2<double> *__restrict__ v,
                                                            const C2<double> *__restrict__ vp,
                                                            C2<double> *__restrict__ w,
                                                            LanczosState st, int step) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double bp = step > 0 ? st.beta[step - 1] : 0.0;
  const int64_t total = t.d_up * t.d_dn;
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / t.d_dn, d = e - u * t.d_dn;
    C2<double> h = hv_elem<double, 1>(t, v, u, d, 0);
    if (bp != 0.0) {
      const C2<double> q = vp[e];
      h.re -= bp * q.re;
      h.im -= bp * q.im;
    }
    w[e] = h;
    const C2<double> ve = v[e];
    s += ve.re * h.re + ve.im * h.im;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) st.partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void lanczos_finish_alpha_kernel(LanczosState st, int step,
                                                                      int nblocks) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double s = final_sum(st.partial, nblocks, lds);
  if (threadIdx.x == 0) {
    st.alpha[step] = s;
    st.flags[1] = step + 1;
  }
}

// w -= alpha v ; partial = |w|^2
__global__ __launch_bounds__(kBlock) void lanczos_axpy_norm_kernel(const double *__restrict__ v,
                                                                   double *__restrict__ w, int64_t n,
                                                                   LanczosState st, int step) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double a = st.alpha[step];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double q = w[i] - a * v[i];
    w[i] = q;
    s += q * q;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) st.partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void lanczos_finish_beta_kernel(LanczosState st, int step,
                                                                     int nblocks) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double s = final_sum(st.partial, nblocks, lds);
  if (threadIdx.x == 0) {
    const double b = sqrt(s);
    st.beta[step] = b;
    if (b <= kTol) st.flags[0] = 1;  // breakdown: later kernels of this run are no-ops
  }
}

__global__ __launch_bounds__(kBlock) void lanczos_normalize_kernel(double *__restrict__ w, int64_t n,
                                                                   LanczosState st, int step) {
  if (st.flags[0]) return;
  const double b = st.beta[step];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    w[i] = w[i] / b;
}

// pass two: w = H v - beta_prev v_prev - alpha v
__global__ __launch_bounds__(kBlock) void recon_hv_kernel(HvTables t, const C2<double> *__restrict__ v,
                                                          const C2<double> *__restrict__ vp,
                                                          C2<double> *__restrict__ w, double alpha,
                                                          double bp) {
  const int64_t total = t.d_up * t.d_dn;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / t.d_dn, d = e - u * t.d_dn;
    C2<double> h = hv_elem<double, 1>(t, v, u, d, 0);
    if (bp != 0.0) {
      const C2<double> q = vp[e];
      h.re -= bp * q.re;
      h.im -= bp * q.im;
    }
    const C2<double> ve = v[e];
    h.re -= alpha * ve.re;
    h.im -= alpha * ve.im;
    w[e] = h;
  }
}

// w /= beta ; x += y * w   (complex y)
__global__ __launch_bounds__(kBlock) void recon_update_kernel(C2<double> *__restrict__ w,
                                                              C2<double> *__restrict__ x, double beta,
                                                              double yr, double yi, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    C2<double> q = w[e];
    q.re = q.re / beta;
    q.im = q.im / beta;
    w[e] = q;
    C2<double> o = x[e];
    o.re += q.re * yr - q.im * yi;
    o.im += q.re * yi + q.im * yr;
    x[e] = o;
  }
}

__global__ __launch_bounds__(kBlock) void scale_complex_kernel(const C2<double> *__restrict__ x,
                                                               double yr, double yi,
                                                               C2<double> *__restrict__ o, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const C2<double> q = x[e];
    o[e] = {q.re * yr - q.im * yi, q.re * yi + q.im * yr};
  }
}

// ------------------------------------------------------------ Chebyshev

constexpr int kChebGrid = 2048;

// tnext = scale2 * (H tn - shift tn) - tprev  (first: tnext = scale2 * (H tn - shift tn))
// partial[b][r] = <tnext|tn>, partial[b][NV + r] = <tnext|tnext>   (FP64)
template <typename R, int NV>
__global__ __launch_bounds__(kBlock) void cheb_kernel(HvTables t, const C2<R> *__restrict__ tn,
                                                      const C2<R> *__restrict__ tprev,
                                                      C2<R> *__restrict__ tnext, R scale2, R shift,
                                                      int first, double *partial) {
  __shared__ double lds[2 * NV][kBlock / kWave];
  const int64_t total = t.d_up * t.d_dn * NV;
  double s_cross = 0.0, s_self = 0.0;
  // kBlock is a multiple of NV and so is the grid stride: each thread keeps one r
  const int r = threadIdx.x % NV;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += (int64_t)gridDim.x * kBlock) {
    const int64_t e = g / NV;
    const int64_t u = e / t.d_dn, d = e - u * t.d_dn;
    C2<R> h = hv_elem<R, NV>(t, tn, u, d, r);
    const C2<R> c = tn[g];
    C2<R> o;
    o.re = scale2 * (h.re - shift * c.re);
    o.im = scale2 * (h.im - shift * c.im);
    if (!first) {
      const C2<R> q = tprev[g];
      o.re -= q.re;
      o.im -= q.im;
    }
    tnext[g] = o;
    s_cross += (double)o.re * (double)c.re + (double)o.im * (double)c.im;
    s_self += (double)o.re * (double)o.re + (double)o.im * (double)o.im;
  }
  // reduce per r: lanes with equal r are NV apart
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off >= NV; off >>= 1) {
    s_cross += __shfl_down(s_cross, off, kWave);
    s_self += __shfl_down(s_self, off, kWave);
  }
  if ((tid & (kWave - 1)) < NV) {
    lds[r][tid / kWave] = s_cross;
    lds[NV + r][tid / kWave] = s_self;
  }
  __syncthreads();
  if (tid < 2 * NV) {
    double a = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) a += lds[tid][w];
    partial[(int64_t)blockIdx.x * 2 * NV + tid] = a;
  }
}

// out[j] = sum_b partial[b * width + j]
__global__ __launch_bounds__(kBlock) void reduce_partials_kernel(const double *partial, int nblocks,
                                                                 int width, double *out) {
  __shared__ double lds[kBlock / kWave];
  for (int j = 0; j < width; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += kBlock) s += partial[(int64_t)i * width + j];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[j] = s;
  }
}

// partial[b][NV + r] = sum |x_r|^2 over the block's elements, partial[b][r] = 0
template <typename R, int NV>
__global__ __launch_bounds__(kBlock) void sumsq_multi_kernel(const C2<R> *__restrict__ x, int64_t D,
                                                             double *partial) {
  __shared__ double lds[NV][kBlock / kWave];
  const int64_t total = D * NV;
  const int r = threadIdx.x % NV;
  double acc = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total; g += (int64_t)gridDim.x * kBlock) {
    const C2<R> c = x[g];
    acc += (double)c.re * (double)c.re + (double)c.im * (double)c.im;
  }
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off >= NV; off >>= 1) acc += __shfl_down(acc, off, kWave);
  if ((tid & (kWave - 1)) < NV) lds[r][tid / kWave] = acc;
  __syncthreads();
  if (tid < NV) {
    double a = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) a += lds[tid][w];
    partial[(int64_t)blockIdx.x * 2 * NV + NV + tid] = a;
    partial[(int64_t)blockIdx.x * 2 * NV + tid] = 0.0;
  }
}

}  // namespace

// ------------------------------------------------------------ launchers

void launch_hv_c128(const HvTables &t, const void *x, void *y, hipStream_t s) {
  const int64_t D = t.d_up * t.d_dn;
  hipLaunchKernelGGL((hv_kernel<double, 1>), dim3(grid_for(D, kGridCap)), dim3(kBlock), 0, s, t,
                     (const C2<double> *)x, (C2<double> *)y);
  MB_HIP(hipGetLastError());
}

void launch_hv_c64(const HvTables &t, const void *x, void *y, hipStream_t s) {
  const int64_t D = t.d_up * t.d_dn;
  hipLaunchKernelGGL((hv_kernel<float, 1>), dim3(grid_for(D, kGridCap)), dim3(kBlock), 0, s, t,
                     (const C2<float> *)x, (C2<float> *)y);
  MB_HIP(hipGetLastError());
}

void launch_hv_c64_multi(const HvTables &t, int nvec, const void *x, void *y, hipStream_t s) {
  const int64_t W = t.d_up * t.d_dn * nvec;
  const dim3 g(grid_for(W, kGridCap)), b(kBlock);
  const auto *xi = (const C2<float> *)x;
  auto *yo = (C2<float> *)y;
  switch (nvec) {
    case 1: hipLaunchKernelGGL((hv_kernel<float, 1>), g, b, 0, s, t, xi, yo); break;
    case 2: hipLaunchKernelGGL((hv_kernel<float, 2>), g, b, 0, s, t, xi, yo); break;
    case 4: hipLaunchKernelGGL((hv_kernel<float, 4>), g, b, 0, s, t, xi, yo); break;
    case 8: hipLaunchKernelGGL((hv_kernel<float, 8>), g, b, 0, s, t, xi, yo); break;
    case 16: hipLaunchKernelGGL((hv_kernel<float, 16>), g, b, 0, s, t, xi, yo); break;
    default: throw std::invalid_argument("nvec must be 1, 2, 4, 8 or 16");
  }
  MB_HIP(hipGetLastError());
}

void launch_gather_c128(const PermTables &p, const void *x_ref, void *x_dev, hipStream_t s) {
  hipLaunchKernelGGL((perm_kernel<true>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock), 0,
                     s, p, (const double2 *)x_ref, (double2 *)x_dev);
  MB_HIP(hipGetLastError());
}

void launch_scatter_c128(const PermTables &p, const void *x_dev, void *x_ref, hipStream_t s) {
  hipLaunchKernelGGL((perm_kernel<false>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock),
                     0, s, p, (const double2 *)x_dev, (double2 *)x_ref);
  MB_HIP(hipGetLastError());
}

void launch_fill_gaussian_c128(const PermTables &p, uint64_t seed, void *x, hipStream_t s) {
  hipLaunchKernelGGL(fill_gaussian_kernel, dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock),
                     0, s, p, seed, (double2 *)x);
  MB_HIP(hipGetLastError());
}

void launch_fill_phase_c64_multi(const PermTables &p, uint64_t seed, int nvec, void *x,
                                 hipStream_t s) {
  hipLaunchKernelGGL((fill_phase_kernel<float>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)),
                     dim3(kBlock), 0, s, p, seed, nvec, (C2<float> *)x);
  MB_HIP(hipGetLastError());
}

void launch_fill_phase_c128_multi(const PermTables &p, uint64_t seed, int nvec, void *x,
                                  hipStream_t s) {
  hipLaunchKernelGGL((fill_phase_kernel<double>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)),
                     dim3(kBlock), 0, s, p, seed, nvec, (C2<double> *)x);
  MB_HIP(hipGetLastError());
}

void launch_norm_c128(const double *x, int64_t n, double *partial, double *out, hipStream_t s) {
  hipLaunchKernelGGL(norm_partial_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, x, n, partial);
  hipLaunchKernelGGL(sqrt_final_kernel, dim3(1), dim3(kBlock), 0, s, (const double *)partial,
                     kRedBlocks, out);
  MB_HIP(hipGetLastError());
}

void launch_scale_div_c128(const double *x, const double *den, double *y, int64_t n,
                           hipStream_t s) {
  hipLaunchKernelGGL(scale_div_kernel, dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s, x, den, y, n);
  MB_HIP(hipGetLastError());
}

static int lanczos_grid(int64_t D) { return grid_for(D, kLanczosGrid); }

void launch_lanczos_hv(const HvTables &t, const double *v, const double *vprev, double *w,
                       const LanczosState &st, int step, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_hv_kernel, dim3(lanczos_grid(t.d_up * t.d_dn)), dim3(kBlock), 0, s, t,
                     (const C2<double> *)v, (const C2<double> *)vprev, (C2<double> *)w, st, step);
  MB_HIP(hipGetLastError());
}

int lanczos_partial_blocks(int64_t D) { return lanczos_grid(D); }

void launch_lanczos_finish_alpha(const LanczosState &st, int step, int nblocks, hipStream_t s) {
  // nblocks = grid of the preceding lanczos_hv launch (one partial per block)
  hipLaunchKernelGGL(lanczos_finish_alpha_kernel, dim3(1), dim3(kBlock), 0, s, st, step, nblocks);
  MB_HIP(hipGetLastError());
}

void launch_lanczos_axpy_norm(const double *v, double *w, int64_t n, const LanczosState &st,
                              int step, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_axpy_norm_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, v, w, n, st,
                     step);
  MB_HIP(hipGetLastError());
}

void launch_lanczos_finish_beta(const LanczosState &st, int step, int k, hipStream_t s) {
  (void)k;
  hipLaunchKernelGGL(lanczos_finish_beta_kernel, dim3(1), dim3(kBlock), 0, s, st, step, kRedBlocks);
  MB_HIP(hipGetLastError());
}

void launch_lanczos_normalize(double *w, int64_t n, const LanczosState &st, int step,
                              hipStream_t s) {
  hipLaunchKernelGGL(lanczos_normalize_kernel, dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s, w, n,
                     st, step);
  MB_HIP(hipGetLastError());
}

void launch_recon_hv(const HvTables &t, const double *v, const double *vprev, double *w,
                     double alpha, double beta_prev, hipStream_t s) {
  hipLaunchKernelGGL(recon_hv_kernel, dim3(grid_for(t.d_up * t.d_dn, kGridCap)), dim3(kBlock), 0, s,
                     t, (const C2<double> *)v, (const C2<double> *)vprev, (C2<double> *)w, alpha,
                     beta_prev);
  MB_HIP(hipGetLastError());
}

void launch_recon_update(double *w, double *x, double beta, double yr, double yi, int64_t n,
                         hipStream_t s) {
  hipLaunchKernelGGL(recon_update_kernel, dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s,
                     (C2<double> *)w, (C2<double> *)x, beta, yr, yi, n);
  MB_HIP(hipGetLastError());
}

void launch_scale_complex_c128(const double *x, double yr, double yi, double *out, int64_t n,
                               hipStream_t s) {
  hipLaunchKernelGGL(scale_complex_kernel, dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s,
                     (const C2<double> *)x, yr, yi, (C2<double> *)out, n);
  MB_HIP(hipGetLastError());
}

int cheb_partial_blocks() { return kChebGrid; }

template <typename R>
static void launch_cheb(const HvTables &t, int nvec, const R *tn, const R *tprev, R *tnext, R scale2,
                        R shift, int first, double *partial, hipStream_t s) {
  const dim3 g(kChebGrid), b(kBlock);
  const auto *a = (const C2<R> *)tn;
  const auto *p = (const C2<R> *)tprev;
  auto *o = (C2<R> *)tnext;
  switch (nvec) {
    case 1: hipLaunchKernelGGL((cheb_kernel<R, 1>), g, b, 0, s, t, a, p, o, scale2, shift, first, partial); break;
    case 2: hipLaunchKernelGGL((cheb_kernel<R, 2>), g, b, 0, s, t, a, p, o, scale2, shift, first, partial); break;
    case 4: hipLaunchKernelGGL((cheb_kernel<R, 4>), g, b, 0, s, t, a, p, o, scale2, shift, first, partial); break;
    case 8: hipLaunchKernelGGL((cheb_kernel<R, 8>), g, b, 0, s, t, a, p, o, scale2, shift, first, partial); break;
    case 16: hipLaunchKernelGGL((cheb_kernel<R, 16>), g, b, 0, s, t, a, p, o, scale2, shift, first, partial); break;
    default: throw std::invalid_argument("nvec must be 1, 2, 4, 8 or 16");
  }
  MB_HIP(hipGetLastError());
}

void launch_cheb_c64(const HvTables &t, int nvec, const float *tn, const float *tprev, float *tnext,
                     float scale2, float shift, int first, double *partial, hipStream_t s) {
  launch_cheb<float>(t, nvec, tn, tprev, tnext, scale2, shift, first, partial, s);
}

void launch_cheb_c128(const HvTables &t, int nvec, const double *tn, const double *tprev,
                      double *tnext, double scale2, double shift, int first, double *partial,
                      hipStream_t s) {
  launch_cheb<double>(t, nvec, tn, tprev, tnext, scale2, shift, first, partial, s);
}

void launch_reduce_partials(const double *partial, int nblocks, int width, double *out,
                            hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, s, partial, nblocks, width,
                     out);
  MB_HIP(hipGetLastError());
}

}  // namespace mb

namespace mb {

template <typename R>
static void sumsq_launch(int nvec, const void *x, int64_t D, double *partial, hipStream_t s) {
  const dim3 g(kChebGrid), b(kBlock);
  const auto *a = (const C2<R> *)x;
  switch (nvec) {
    case 1: hipLaunchKernelGGL((sumsq_multi_kernel<R, 1>), g, b, 0, s, a, D, partial); break;
    case 2: hipLaunchKernelGGL((sumsq_multi_kernel<R, 2>), g, b, 0, s, a, D, partial); break;
    case 4: hipLaunchKernelGGL((sumsq_multi_kernel<R, 4>), g, b, 0, s, a, D, partial); break;
    case 8: hipLaunchKernelGGL((sumsq_multi_kernel<R, 8>), g, b, 0, s, a, D, partial); break;
    case 16: hipLaunchKernelGGL((sumsq_multi_kernel<R, 16>), g, b, 0, s, a, D, partial); break;
    default: throw std::invalid_argument("nvec must be 1, 2, 4, 8 or 16");
  }
  MB_HIP(hipGetLastError());
}

void launch_sumsq_multi(bool fp32, int nvec, const void *x, int64_t D, double *partial,
                        hipStream_t s) {
  if (fp32)
    sumsq_launch<float>(nvec, x, D, partial, s);
  else
    sumsq_launch<double>(nvec, x, D, partial, s);
}

}  // namespace mb
