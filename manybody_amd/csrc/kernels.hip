// HIP kernels for gfx950 (CDNA4, wave64): matrix-free Hubbard H.v, the
// reference-order <-> device-order permutation, fused Lanczos and Chebyshev
// epilogues and seeded vector fills.
//
// H = H_up (x) 1 + 1 (x) H_dn + diagonal acts on the D_up x D_dn matrix of each
// real plane.  Up hops combine whole rows with one signed amplitude per row
// pair (row u of H x gathers rows u' of x); down hops gather inside a row.  The
// two need different locality, so H.v runs as two kernels:
//
//  1. up_strip_kernel: the plane is cut into column strips of 256 bytes per row
//     (32 doubles / 64 floats).  One strip of all D_up rows (2.9 MB for the
//     4 x 4, 7 + 7 sector) stays resident in the L2 of one XCD while all rows
//     of that strip are produced, so every x element is fetched from HBM once
//     and its ~17 reuses hit L2.  Workgroup -> (strip, row chunk) is XCD-aware:
//     workgroups are dispatched round-robin over the 8 XCDs, so the chunks of a
//     strip are given ids congruent mod 8.  The hop lists of a workgroup's 32
//     rows are staged in LDS as 32-bit byte offsets of the target rows (a
//     plane is < 4 GiB) and amplitudes, so a gather is one add on a uniform
//     base.  Writes y = (H_up x).
//  2. row kernel: a plane row x[u, :] is staged in LDS (91.5 KB for D_dn =
//     11,440 doubles), the diagonal and the down hops are applied from LDS, the
//     strip result is added and the epilogue (plain, Lanczos, reconstruction,
//     Chebyshev) applied, with one partial sum per row for the fused inner
//     products.  row_kernel_sorted (persistent, one workgroup per CU, strings in
//     hop-count order so that a wave's lanes share trip count and sign
//     pattern) is used where its tables exist; row_kernel_compact (one
//     workgroup per row, padded trip counts, per-lane sign masks) otherwise.
//
// Reductions are two-level and atomic-free (fixed partial slots, fixed tree),
// so results do not depend on scheduling.  A one-kernel generic path (thread
// per element, both tables from global memory) covers sectors whose rows do not
// fit in LDS or whose tables cannot be packed.
#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdlib>
#include <cstring>

#include "krylov.hpp"
#include "mb_internal.hpp"
#include "momentum.hpp"
#include "conductivity.hpp"
#include "expression.hpp"

namespace mb {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kMaxRowLds = 160 * 1024 - 1024;  // bytes of LDS a staged row may use
constexpr int kGenericGrid = 2048;              // blocks per plane of the generic path
constexpr int kCompactK = 32;                   // entries per string of the compact down table
#ifndef MB_STRIP_LANES
#define MB_STRIP_LANES 16
#endif
constexpr int kStripLanes = MB_STRIP_LANES;     // lanes per row of the strip kernel (16 B each)

__device__ __forceinline__ int popc(uint32_t v) { return __builtin_popcount(v); }

// Fixed-order sum over the workgroup (blockDim multiple of 64); valid in thread 0.
__device__ double wg_sum(double v, double *lds) {
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) lds[tid / kWave] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r += lds[w];
  __syncthreads();
  return r;
}

template <typename T>
struct Tabs;
template <>
struct Tabs<double> {
  __device__ static const double *up_amp(const HvTables &t) { return t.up_amp; }
  __device__ static const double *dn_amp(const HvTables &t) { return t.dn_amp; }
  __device__ static const double *up_on(const HvTables &t) { return t.up_on; }
  __device__ static const double *dn_on(const HvTables &t) { return t.dn_on; }
  __device__ static const double *diag(const HvTables &t) { return t.diagU; }
  __device__ static const double *mags(const HvTables &t) { return t.mags; }
};
template <>
struct Tabs<float> {
  __device__ static const float *up_amp(const HvTables &t) { return t.up_ampf; }
  __device__ static const float *dn_amp(const HvTables &t) { return t.dn_ampf; }
  __device__ static const float *up_on(const HvTables &t) { return t.up_onf; }
  __device__ static const float *dn_on(const HvTables &t) { return t.dn_onf; }
  __device__ static const float *diag(const HvTables &t) { return t.diagUf; }
  __device__ static const float *mags(const HvTables &t) { return t.magsf; }
};

// ------------------------------------------------------------ epilogue

struct Partial2 {
  double s0 = 0.0, s1 = 0.0;
};

// Per-launch constants of the epilogue: bp multiplies prev; sx scales H x and
// x when the Lanczos vectors are kept unnormalised (EpiArgs::lazy).
template <typename T>
struct EpiK {
  T bp = T(0), sx = T(1);
};

template <typename T, Epi E>
__device__ __forceinline__ EpiK<T> epi_consts(const EpiArgs &a) {
  EpiK<T> k;
  if constexpr (E == Epi::Lanczos) {
    if (a.step > 0) {
      const double b1 = a.beta_dev[a.step - 1];
      if (a.lazy) {  // x = b1 v, prev = b2 v_prev (b2 = beta[step - 2], 1 at step 1)
        const double b2 = a.step > 1 ? a.beta_dev[a.step - 2] : 1.0;
        k.sx = (T)(1.0 / b1);
        k.bp = (T)(b1 / b2);
      } else {
        k.bp = (T)b1;
      }
    }
  } else if constexpr (E == Epi::Recon) {
    k.bp = (T)a.beta;
  }
  return k;
}

// Applies the epilogue to r = (H x)[i] at plane offset i; xv = x[i].
template <typename T, Epi E>
__device__ __forceinline__ void epilogue(T r, T xv, int64_t i, T *__restrict__ y,
                                         const EpiArgs &a, const EpiK<T> &bp, Partial2 &acc) {
  if constexpr (E == Epi::Plain) {
    y[i] = r;
  } else if constexpr (E == Epi::Lanczos) {
    if (a.lazy) {
      r *= bp.sx;
      xv *= bp.sx;
    }
    if (bp.bp != T(0)) r -= bp.bp * static_cast<const T *>(a.prev)[i];
    y[i] = r;
    acc.s0 += (double)xv * (double)r;
  } else if constexpr (E == Epi::Recon) {
    if (bp.bp != T(0)) r -= bp.bp * static_cast<const T *>(a.prev)[i];
    r -= (T)a.alpha * xv;
    y[i] = r;
  } else {  // Cheb
    T o = (T)a.scale2 * (r - (T)a.shift * xv);
    if (!a.first) o -= static_cast<const T *>(a.prev)[i];
    y[i] = o;
    acc.s0 += (double)o * (double)xv;
    acc.s1 += (double)o * (double)o;
  }
}


// ------------------------------------------------------------ kernel 1: up hops

// VEC consecutive values of one row, loaded/stored as one 16-byte access when VEC > 1.
template <typename T, int VEC>
struct Vec {
  T v[VEC];
  __device__ __forceinline__ void load(const T *p) {
    if constexpr (VEC * sizeof(T) == 16) {
      const uint4 q = *reinterpret_cast<const uint4 *>(p);
      __builtin_memcpy(v, &q, 16);
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) v[j] = p[j];
    }
  }
};

// 16 lanes per row, 4 rows per wave, 32 rows per workgroup.  The workgroup's hop
// lists are staged in LDS; y is written with non-temporal stores so that the
// output stream does not evict the strip of x from L2.
template <typename T, int VEC, int MAXH>
__global__ __launch_bounds__(kBlock) void up_strip_kernel(HvTables t, int P, const T *__restrict__ x,
                                                          T *__restrict__ y, int nJ, int nC,
                                                          const int *flags) {
  constexpr int LPR = kStripLanes, RPW = kWave / LPR, RC = 32, W = LPR * VEC;
  __shared__ uint32_t s_off[RC][MAXH];  // byte offset of the target row in the plane
  __shared__ T s_amp[RC][MAXH];
  __shared__ int s_n[RC];
  if (flags && flags[0]) return;
  const int S = P * nJ;
  const int i = blockIdx.x;
  const int xcd = i & 7, q = i >> 3;
  const int s = xcd + 8 * (q / nC);
  const int c = q % nC;
  if (s >= S) return;
  const int p = s / nJ, J = s - p * nJ;
  const int u_begin = c * RC;
  const int u_end = (int)min((int64_t)u_begin + RC, t.d_up);
  // hop lists of the workgroup's rows; from the 16-bit table where there is
  // one magnitude (every strip re-reads the lists: 2 bytes a hop instead of 12
  // keeps them from displacing the strip in L2)
  const T *__restrict__ amp = Tabs<T>::up_amp(t);
  const T mag = (T)t.up_mag;
  for (int e = threadIdx.x; e < RC * MAXH; e += kBlock) {
    const int r = e / MAXH, k = e % MAXH;
    const int u = u_begin + r;
    if (u < u_end) {
      const int b = t.up_ptr[u], n = t.up_ptr[u + 1] - b;
      if (k == 0) s_n[r] = n;
      if (k < n) {
        uint32_t tg;
        T a;
        if (t.up_c16_ok) {
          const uint32_t c = t.up_c16[b + k];
          tg = c & 0x7FFFu;
          a = (c & 0x8000u) ? -mag : mag;
        } else {
          tg = (uint32_t)t.up_tgt[b + k];
          a = amp[b + k];
        }
        s_off[r][k] = (uint32_t)((uint64_t)tg * (uint64_t)t.d_dn * sizeof(T));
        s_amp[r][k] = a;
      }
    } else if (k == 0) {
      s_n[r] = 0;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int64_t Ddn = t.d_dn;
  const int64_t col = (int64_t)J * W + VEC * (lane % LPR);
  const bool col_ok = col < Ddn;  // D_dn is a multiple of VEC
  const int64_t plane = (int64_t)p * t.d_up * Ddn;
  // gathers: uniform plane base + 32-bit byte offset (row offset from LDS + column)
  const char *__restrict__ xb = reinterpret_cast<const char *>(x + plane);
  const uint32_t colb = (uint32_t)(col * (int64_t)sizeof(T));
  T *__restrict__ yp = y + plane;
  const int sub = lane / LPR;
  for (int rb = wv * RPW; rb < RC; rb += (kBlock / kWave) * RPW) {
    const int r = rb + sub;
    const int u = u_begin + r;
    const bool ok = col_ok && u < u_end;
    const int n = ok ? s_n[r] : 0;
    T acc[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) acc[j] = T(0);
#pragma unroll 4
    for (int k = 0; k < n; ++k) {
      const T a = s_amp[r][k];
      Vec<T, VEC> v;
      v.load(reinterpret_cast<const T *>(xb + (s_off[r][k] + colb)));
#pragma unroll
      for (int j = 0; j < VEC; ++j) acc[j] += a * v.v[j];
    }
    if (ok) {
      T *dst = yp + (int64_t)u * Ddn + col;
#pragma unroll
      for (int j = 0; j < VEC; ++j) __builtin_nontemporal_store(acc[j], dst + j);
    }
  }
}

// ------------------------------------------------------------ kernel 2: row

// Shared prologue of the row kernels: stage x[plane row] in LDS.
template <typename T>
__device__ __forceinline__ void stage_row(T *row, const T *__restrict__ x, int64_t base, int64_t Ddn) {
  for (int64_t d = threadIdx.x; d < Ddn; d += blockDim.x) row[d] = x[base + d];
  __syncthreads();
}

template <typename T, Epi E>
__device__ __forceinline__ void row_reduce(const Partial2 &acc2, const EpiArgs &a, int64_t wg,
                                           double (*red)[1024 / kWave]) {
  if constexpr (E == Epi::Lanczos || E == Epi::Cheb) {
    const double s0 = wg_sum(acc2.s0, red[0]);
    double s1 = 0.0;
    if constexpr (E == Epi::Cheb) s1 = wg_sum(acc2.s1, red[1]);
    if (threadIdx.x == 0) {
      if constexpr (E == Epi::Lanczos) {
        a.partial[wg] = s0;
      } else {
        a.partial[2 * wg] = s0;
        a.partial[2 * wg + 1] = s1;
      }
    }
  }
}

// Flips the sign of v when sbit (0 or the sign-bit mask) is set.
__device__ __forceinline__ double flip_sign(double v, uint32_t sbit) {
  return __hiloint2double(__double2hiint(v) ^ (int)sbit, __double2loint(v));
}
__device__ __forceinline__ float flip_sign(float v, uint32_t sbit) {
  return __int_as_float(__float_as_int(v) ^ (int)sbit);
}

// Compact down table, one amplitude magnitude h: 32 uint16 target indices per
// string (fetched as up to four 16-byte loads, as many as the group needs), unused slots pointing at a zero slot
// row[D_dn]; bit k of a per-string mask is the sign of slot k; per group of 64
// strings (one wave) the largest hop count, so every lane runs the same number
// of slots (rounded up to 8) with no per-lane predicate.  The 8 LDS reads of a
// group are issued together; signs are applied by flipping the sign bit.
template <typename T, Epi E>
__global__ __launch_bounds__(1024) void row_kernel_compact(HvTables t, const T *__restrict__ x,
                                                           T *__restrict__ y, EpiArgs a) {
  constexpr int K = kCompactK, G = 8;
  extern __shared__ double smem[];
  __shared__ double red[2][1024 / kWave];
  T *row = reinterpret_cast<T *>(smem);
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int64_t wg = blockIdx.x;  // = p * D_up + u
  const int64_t Ddn = t.d_dn, Dup = t.d_up;
  const int64_t u = wg % Dup;
  const int64_t base = wg * Ddn;
  if (threadIdx.x == 0) row[Ddn] = T(0);
  stage_row(row, x, base, Ddn);
  const uint32_t su = t.s_up[u];
  const T *__restrict__ diag = Tabs<T>::diag(t);
  const T h = Tabs<T>::mags(t)[0];
  const EpiK<T> bp = epi_consts<T, E>(a);
  const T on_u = t.has_onsite ? Tabs<T>::up_on(t)[u] : T(0);
  const int lane = threadIdx.x & (kWave - 1);
  Partial2 acc2;
  for (int64_t d0 = (int64_t)threadIdx.x - lane; d0 < Ddn; d0 += blockDim.x) {
    const int64_t d = d0 + lane;  // tables are padded to a multiple of 64 strings
    const bool ok = d < Ddn;
    const int wm = __builtin_amdgcn_readfirstlane((int)t.dn_cgmax[d0 / kWave]);
    uint4 q[K / 8];
#pragma unroll
    for (int j = 0; j < K / 8; ++j) q[j] = comp_chunk(t.dn_comp, d, j);  // all four: see comp_chunk
    const uint32_t sm = t.dn_csign[d];
    const T yin = ok ? y[base + d] : T(0);
    const T xv = row[ok ? d : Ddn];
    T dg = diag[popc(su & t.s_dn[ok ? d : 0])];
    if (t.has_onsite) dg = dg + on_u + Tabs<T>::dn_on(t)[ok ? d : 0];
    T h0 = T(0), h1 = T(0);
#pragma unroll
    for (int g = 0; g < K / G; ++g) {
      if (g * G < wm) {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int k = g * G + j;
          const uint32_t w = reinterpret_cast<const uint32_t *>(&q[k / 8])[(k % 8) / 2];
          v[j] = row[(k & 1) ? (w >> 16) : (w & 0xFFFFu)];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int k = g * G + j;
          const T fv = flip_sign(v[j], (sm << (31 - k)) & 0x80000000u);
          if (j & 1) h1 += fv; else h0 += fv;
        }
      }
    }
    if (ok) {
      // (H_up x) from kernel 1, then the diagonal and the down hops
      const T r = yin + (dg * xv + h * (h0 + h1));
      epilogue<T, E>(r, xv, base + d, y, a, bp, acc2);
    }
  }
  row_reduce<T, E>(acc2, a, wg, red);
}

// NP consecutive values of T at a 16-byte (or, for 8 bytes, 8-byte) aligned address.
template <typename T, int NP>
__device__ __forceinline__ void load_pack(const T *p, T (&v)[NP]) {
  constexpr int B = NP * (int)sizeof(T);
  if constexpr (B % 16 == 0) {
#pragma unroll
    for (int j = 0; j < B / 16; ++j) {
      const uint4 q = reinterpret_cast<const uint4 *>(p)[j];
      __builtin_memcpy(reinterpret_cast<char *>(v) + 16 * j, &q, 16);
    }
  } else {
    static_assert(B == 8, "pack of 8, 16 or 32 bytes");
    const uint2 q = *reinterpret_cast<const uint2 *>(p);
    __builtin_memcpy(v, &q, 8);
  }
}
template <typename T, int NP>
__device__ __forceinline__ void store_pack(T *p, const T (&v)[NP]) {
  constexpr int B = NP * (int)sizeof(T);
  if constexpr (B % 16 == 0) {
#pragma unroll
    for (int j = 0; j < B / 16; ++j) {
      uint4 q;
      __builtin_memcpy(&q, reinterpret_cast<const char *>(v) + 16 * j, 16);
      reinterpret_cast<uint4 *>(p)[j] = q;
    }
  } else {
    uint2 q;
    __builtin_memcpy(&q, v, 8);
    *reinterpret_cast<uint2 *>(p) = q;
  }
}

// row_kernel_compact over NP planes of the same row at once: the rows of NP
// consecutive planes (for one complex vector its re and im planes, for several
// vectors several of them) are staged in LDS interleaved, NP values per string,
// so each compact-table entry is read from L2 once for NP planes and each hop
// is one LDS read of NP values.  Same sums in the same order as
// row_kernel_compact, so the results are bit-identical; one partial-sum slot
// per plane row as there.  Workgroup = plane group * D_up + row.
template <typename T, Epi E, int NP>
__global__ __launch_bounds__(1024) void row_kernel_mp(HvTables t, const T *__restrict__ x,
                                                      T *__restrict__ y, EpiArgs a) {
  constexpr int K = kCompactK, G = 8, NS = (E == Epi::Cheb) ? 2 : 1;
  extern __shared__ uint4 smem16[];  // 16-byte aligned for the pack reads
  __shared__ double red[2 * NP][1024 / kWave];
  T *row = reinterpret_cast<T *>(smem16);  // (D_dn + 1) * NP, string-major; string D_dn is zero
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int64_t Ddn = t.d_dn, Dup = t.d_up, D = Dup * Ddn;
  const int64_t wg = blockIdx.x;
  const int64_t u = wg % Dup, pg = wg / Dup;
  const int64_t base0 = pg * NP * D + u * Ddn;  // plane pg * NP, row u
  for (int64_t d = threadIdx.x; d < Ddn; d += blockDim.x) {
    T v[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) v[p] = x[base0 + p * D + d];
    store_pack<T, NP>(row + d * NP, v);
  }
  if (threadIdx.x < NP) row[Ddn * NP + threadIdx.x] = T(0);
  __syncthreads();
  const uint32_t su = t.s_up[u];
  const T *__restrict__ diag = Tabs<T>::diag(t);
  const T h = Tabs<T>::mags(t)[0];
  const EpiK<T> bp = epi_consts<T, E>(a);
  const T on_u = t.has_onsite ? Tabs<T>::up_on(t)[u] : T(0);
  const int lane = threadIdx.x & (kWave - 1);
  Partial2 acc2[NP];
  for (int64_t d0 = (int64_t)threadIdx.x - lane; d0 < Ddn; d0 += blockDim.x) {
    const int64_t d = d0 + lane;  // tables are padded to a multiple of 64 strings
    const bool ok = d < Ddn;
    const int wm = __builtin_amdgcn_readfirstlane((int)t.dn_cgmax[d0 / kWave]);
    uint4 q[K / 8];
#pragma unroll
    for (int j = 0; j < K / 8; ++j) q[j] = comp_chunk(t.dn_comp, d, j);  // all four: see comp_chunk
    const uint32_t sm = t.dn_csign[d];
    T dg = diag[popc(su & t.s_dn[ok ? d : 0])];
    if (t.has_onsite) dg = dg + on_u + Tabs<T>::dn_on(t)[ok ? d : 0];
    T h0[NP], h1[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) h0[p] = h1[p] = T(0);
#pragma unroll
    for (int g = 0; g < K / G; ++g) {
      if (g * G < wm) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int k = g * G + j;
          const uint32_t w = reinterpret_cast<const uint32_t *>(&q[k / 8])[(k % 8) / 2];
          const uint32_t tg = (k & 1) ? (w >> 16) : (w & 0xFFFFu);
          T v[NP];
          load_pack<T, NP>(row + (int64_t)tg * NP, v);
          const uint32_t sb = (sm << (31 - k)) & 0x80000000u;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const T fv = flip_sign(v[p], sb);
            if (j & 1) h1[p] += fv; else h0[p] += fv;
          }
        }
      }
    }
    if (ok) {
      T xv[NP];
      load_pack<T, NP>(row + d * NP, xv);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int64_t i = base0 + p * D + d;
        const T r = y[i] + (dg * xv[p] + h * (h0[p] + h1[p]));
        epilogue<T, E>(r, xv[p], i, y, a, bp, acc2[p]);
      }
    }
  }
  if constexpr (E == Epi::Lanczos || E == Epi::Cheb) {
    // per plane: wave sums, then the waves in order (fixed tree, as wg_sum)
    const int wv = threadIdx.x / kWave;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int c = 0; c < NS; ++c) {
        double v = c == 0 ? acc2[p].s0 : acc2[p].s1;
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
        if (lane == 0) red[NS * p + c][wv] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x < NS * NP) {
      double r = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r += red[threadIdx.x][w];
      const int p = threadIdx.x / NS, c = threadIdx.x % NS;
      const int64_t prow = (pg * NP + p) * Dup + u;  // the plane row's slot
      if constexpr (E == Epi::Lanczos)
        a.partial[prow] = r;
      else
        a.partial[2 * prow + c] = r;
    }
  }
}

// H.v in one pass for sectors whose plane group (NP planes of D_up x D_dn) is
// small enough to stay in the Infinity Cache: row_kernel_mp with the up hops
// formed in the same kernel instead of by a separate strip pass.  The row's up
// hop list is staged in LDS; each column then sums its up-source values
// x[u', d] (read from L2 or the Infinity Cache, coalesced over the columns) in
// hop order, exactly as up_strip_kernel does, and adds the down hops and the
// diagonal as row_kernel_mp adds them to the strip result, so the two paths
// give identical results.  HBM traffic per H.v: x once, y once (and prev for
// the Lanczos, reconstruction and Chebyshev epilogues), against x twice and the
// strip result written and read back for the two passes.
constexpr int kFusedMaxHops = 64;  // up hops per row the fused kernel stages

template <typename T, Epi E, int NP>
__global__ __launch_bounds__(1024) void row_kernel_fused(HvTables t, const T *__restrict__ x,
                                                         T *__restrict__ y, EpiArgs a) {
  constexpr int K = kCompactK, G = 8, NS = (E == Epi::Cheb) ? 2 : 1;
  extern __shared__ uint4 smem16[];  // 16-byte aligned for the pack reads
  __shared__ double red[2 * NP][1024 / kWave];
  __shared__ uint32_t s_off[kFusedMaxHops];  // byte offset of the source row in the plane
  __shared__ T s_amp[kFusedMaxHops];
  T *row = reinterpret_cast<T *>(smem16);  // (D_dn + 1) * NP, string-major; string D_dn is zero
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int64_t Ddn = t.d_dn, Dup = t.d_up, D = Dup * Ddn;
  const int64_t wg = blockIdx.x;
  const int64_t u = wg % Dup, pg = wg / Dup;
  const int64_t base0 = pg * NP * D + u * Ddn;  // plane pg * NP, row u
  const int hb = t.up_ptr[u], nh = t.up_ptr[u + 1] - hb;
  const T *__restrict__ amp = Tabs<T>::up_amp(t);
  const T mag = (T)t.up_mag;
  for (int k = threadIdx.x; k < nh; k += blockDim.x) {
    uint32_t tg;
    T am;
    if (t.up_c16_ok) {
      const uint32_t c = t.up_c16[hb + k];
      tg = c & 0x7FFFu;
      am = (c & 0x8000u) ? -mag : mag;
    } else {
      tg = (uint32_t)t.up_tgt[hb + k];
      am = amp[hb + k];
    }
    s_off[k] = (uint32_t)((uint64_t)tg * (uint64_t)Ddn * sizeof(T));
    s_amp[k] = am;
  }
  for (int64_t d = threadIdx.x; d < Ddn; d += blockDim.x) {
    T v[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) v[p] = x[base0 + p * D + d];
    store_pack<T, NP>(row + d * NP, v);
  }
  if (threadIdx.x < NP) row[Ddn * NP + threadIdx.x] = T(0);
  // plane group base as bytes: uniform 64-bit base + 32-bit row offset + column
  const char *__restrict__ xg = reinterpret_cast<const char *>(x + pg * NP * D);
  __syncthreads();
  const uint32_t su = t.s_up[u];
  const T *__restrict__ diag = Tabs<T>::diag(t);
  const T h = Tabs<T>::mags(t)[0];
  const EpiK<T> bp = epi_consts<T, E>(a);
  const T on_u = t.has_onsite ? Tabs<T>::up_on(t)[u] : T(0);
  const int lane = threadIdx.x & (kWave - 1);
  Partial2 acc2[NP];
  for (int64_t d0 = (int64_t)threadIdx.x - lane; d0 < Ddn; d0 += blockDim.x) {
    const int64_t d = d0 + lane;  // tables are padded to a multiple of 64 strings
    const bool ok = d < Ddn;
    const int64_t dc = ok ? d : Ddn - 1;  // in bounds for the up-source loads
    // up hops, in hop order as the strip kernel sums them
    T up[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) up[p] = T(0);
#pragma unroll 4
    for (int k = 0; k < nh; ++k) {
      const T am = s_amp[k];
      const uint32_t off = s_off[k] + (uint32_t)(dc * (int64_t)sizeof(T));
#pragma unroll
      for (int p = 0; p < NP; ++p)
        up[p] += am * *reinterpret_cast<const T *>(xg + (int64_t)p * D * (int64_t)sizeof(T) + off);
    }
    const int wm = __builtin_amdgcn_readfirstlane((int)t.dn_cgmax[d0 / kWave]);
    uint4 q[K / 8];
#pragma unroll
    for (int j = 0; j < K / 8; ++j) q[j] = comp_chunk(t.dn_comp, d, j);  // all four: see comp_chunk
    const uint32_t sm = t.dn_csign[d];
    T dg = diag[popc(su & t.s_dn[ok ? d : 0])];
    if (t.has_onsite) dg = dg + on_u + Tabs<T>::dn_on(t)[ok ? d : 0];
    T h0[NP], h1[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) h0[p] = h1[p] = T(0);
#pragma unroll
    for (int g = 0; g < K / G; ++g) {
      if (g * G < wm) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int k = g * G + j;
          const uint32_t w = reinterpret_cast<const uint32_t *>(&q[k / 8])[(k % 8) / 2];
          const uint32_t tg = (k & 1) ? (w >> 16) : (w & 0xFFFFu);
          T v[NP];
          load_pack<T, NP>(row + (int64_t)tg * NP, v);
          const uint32_t sb = (sm << (31 - k)) & 0x80000000u;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const T fv = flip_sign(v[p], sb);
            if (j & 1) h1[p] += fv; else h0[p] += fv;
          }
        }
      }
    }
    if (ok) {
      T xv[NP];
      load_pack<T, NP>(row + d * NP, xv);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int64_t i = base0 + p * D + d;
        const T r = up[p] + (dg * xv[p] + h * (h0[p] + h1[p]));
        epilogue<T, E>(r, xv[p], i, y, a, bp, acc2[p]);
      }
    }
  }
  if constexpr (E == Epi::Lanczos || E == Epi::Cheb) {
    const int wv = threadIdx.x / kWave;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int c = 0; c < NS; ++c) {
        double v = c == 0 ? acc2[p].s0 : acc2[p].s1;
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
        if (lane == 0) red[NS * p + c][wv] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x < NS * NP) {
      double r = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r += red[threadIdx.x][w];
      const int p = threadIdx.x / NS, c = threadIdx.x % NS;
      const int64_t prow = (pg * NP + p) * Dup + u;  // the plane row's slot
      if constexpr (E == Epi::Lanczos)
        a.partial[prow] = r;
      else
        a.partial[2 * prow + c] = r;
    }
  }
}

// Epilogue with the values it reads from memory loaded beforehand: pv = prev[i]
// (ignored where the epilogue does not read it).
template <typename T, Epi E>
__device__ __forceinline__ void epilogue_pv(T r, T xv, T pv, int64_t i, T *__restrict__ y,
                                            const EpiArgs &a, const EpiK<T> &bp, Partial2 &acc) {
  if constexpr (E == Epi::Plain) {
    y[i] = r;
  } else if constexpr (E == Epi::Lanczos) {
    if (a.lazy) {
      r *= bp.sx;
      xv *= bp.sx;
    }
    if (bp.bp != T(0)) r -= bp.bp * pv;
    y[i] = r;
    acc.s0 += (double)xv * (double)r;
  } else if constexpr (E == Epi::Recon) {
    if (bp.bp != T(0)) r -= bp.bp * pv;
    r -= (T)a.alpha * xv;
    y[i] = r;
  } else {  // Cheb
    T o = (T)a.scale2 * (r - (T)a.shift * xv);
    if (!a.first) o -= pv;
    y[i] = o;
    acc.s0 += (double)o * (double)xv;
    acc.s1 += (double)o * (double)o;
  }
}

// Coefficient of the previous Lanczos step's axpy (x = w - c v), as
// lanczos_axpy_norm_kernel forms it for unnormalised vectors.
template <typename T>
__device__ __forceinline__ T fold_coef(const EpiArgs &a) {
  const int st = a.fold_step;
  return st > 0 ? (T)(a.alpha_dev[st] / a.beta_dev[st - 1]) : (T)a.alpha_dev[st];
}

// x[i] = w[i] - c v[i] (the axpy's arithmetic), both read once: non-temporal
// loads (2.62 against 2.68 ms for the 4 x 4, 7 + 7 row pass).
template <typename T>
__device__ __forceinline__ T fold_val(const T *w, const T *v, T c, int64_t i) {
  return __builtin_nontemporal_load(w + i) - c * __builtin_nontemporal_load(v + i);
}

template <typename T, Epi E>
__device__ __forceinline__ bool reads_prev(const EpiArgs &a, const EpiK<T> &bp) {
  if constexpr (E == Epi::Lanczos || E == Epi::Recon) return bp.bp != T(0);
  if constexpr (E == Epi::Cheb) return !a.first;
  return false;
}

// Block-staged up hops.  Up strings that agree from bit k upward are a
// contiguous run of rows (a block; the strings are in colex order), and a hop
// between two sites below k stays inside its block.  One workgroup takes one
// block and one column strip (LPR lanes x 16 bytes per row): it stages the
// strip of the block's rows and the block's hop entries in LDS, then each row
// sums the hops that leave the block from L2, where the strip of every row
// stays as for up_strip_kernel (the blocks of a strip share an XCD), and the
// in-block hops from LDS, in the order of the row's entries (out-of-block ones
// first).  Small blocks and many workgroups per CU measured best: on the 4 x 4,
// 7 + 7 sector k = 8 (blocks of <= 70 rows, 12 of 32 bonds inside, ~21 KB of
// LDS, 7 workgroups per CU) gives 1.65 ms against 2.17 ms for up_strip_kernel;
// larger blocks keep more hops inside but leave fewer workgroups per CU
// (k = 9: 2.0 ms, k = 10: 2.1 ms), and 128-byte strips or 512 threads were
// slower at every k.
constexpr int kUbStage = 8;     // staging loads per thread in flight
constexpr int kUbLanes = 16;    // lanes per row, 16 bytes each: 256-byte strips
constexpr int kUbThreads = 256;
static_assert(kUbLanes * 16 == kUpBlockRowBytes, "the host sizes blocks for this strip width");

// LZ: the up pass of a row-first Lanczos step: y = Lanczos epilogue of
// (up hops) x + z, z the row pass's (down hops + diagonal) x, with one partial
// sum per workgroup in a.partial (the string's own value comes from the
// staged rows, since a block's output rows are its staged rows).
template <typename T, int LPR = kUbLanes, int NT = kUbThreads, bool LZ = false>
__global__ __launch_bounds__(NT) void up_block_kernel(HvTables t, int P, const T *__restrict__ x,
                                                      T *__restrict__ y, int nJ, const int *flags,
                                                      const T *__restrict__ z, EpiArgs a) {
  constexpr int VEC = 16 / sizeof(T), W = LPR * VEC;
  extern __shared__ uint4 ub_smem[];
  __shared__ double red[LZ ? NT / kWave : 1];
  if (flags && flags[0]) return;
  const int nb = t.ub_nb;
  const int i = blockIdx.x;
  const int xcd = i & 7, q = i >> 3;
  const int s = xcd + 8 * (q / nb);
  const int b = q % nb;
  if (s >= P * nJ) {
    if constexpr (LZ) {
      if (threadIdx.x == 0) a.partial[i] = 0.0;  // a pad workgroup of the XCD map
    }
    return;
  }
  const int p = s / nJ, J = s - p * nJ;
  const int r0 = t.ub_blk[2 * b], nr = t.ub_blk[2 * b + 1];
  uint4 *sx = ub_smem;  // row r of the block at sx[r * LPR + lane]
  uint32_t *smeta = reinterpret_cast<uint32_t *>(ub_smem + t.ub_rows * LPR);
  uint16_t *sent = reinterpret_cast<uint16_t *>(smeta + t.ub_rows);
  const int e0 = t.up_ptr[r0], ne = t.up_ptr[r0 + nr] - e0;
  const int64_t Ddn = t.d_dn;
  const uint32_t rowb = (uint32_t)(Ddn * (int64_t)sizeof(T));  // a plane is < 4 GiB
  const int ln = threadIdx.x & (LPR - 1);
  const int64_t col = (int64_t)J * W + VEC * ln;
  const bool col_ok = col < Ddn;  // D_dn is a multiple of VEC
  const int64_t plane = (int64_t)p * t.d_up * Ddn;
  const char *__restrict__ xb = reinterpret_cast<const char *>(x + plane);
  const uint32_t colb = (uint32_t)(col * (int64_t)sizeof(T));
  // staging loads issued kUbStage at a time before any is stored
  for (int eb = threadIdx.x; eb < nr * LPR; eb += kUbStage * NT) {
    uint4 v[kUbStage];
#pragma unroll
    for (int j = 0; j < kUbStage; ++j) {
      const int e = eb + j * NT;
      const uint32_t r = (uint32_t)(e / LPR);  // e % LPR == ln
      v[j] = (col_ok && e < nr * LPR)
                 ? *reinterpret_cast<const uint4 *>(xb + ((uint32_t)r0 + r) * rowb + colb)
                 : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < kUbStage; ++j)
      if (eb + j * NT < nr * LPR) sx[eb + j * NT] = v[j];
  }
  for (int r = threadIdx.x; r < nr; r += NT) {
    const int u = r0 + r;
    smeta[r] = (uint32_t)(t.up_ptr[u] - e0) | (uint32_t)t.ub_nin[u] << 16 |
               (uint32_t)(t.up_ptr[u + 1] - t.up_ptr[u]) << 24;
  }
  for (int e = threadIdx.x; e < ne; e += NT) sent[e] = t.up_c16[e0 + e];
  __syncthreads();
  const T mag = (T)t.up_mag;
  T *__restrict__ yp = y + plane;
  EpiK<T> bk;
  bool rp = false;
  if constexpr (LZ) {
    bk = epi_consts<T, Epi::Lanczos>(a);
    rp = reads_prev<T, Epi::Lanczos>(a, bk);
  }
  Partial2 part;
  for (int r = threadIdx.x / LPR; r < nr; r += NT / LPR) {
    const uint32_t m = smeta[r];
    const int beg = (int)(m & 0xFFFFu), nin = (int)((m >> 16) & 0xFFu), ntot = (int)(m >> 24);
    const int nout = ntot - nin;
    T acc[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) acc[j] = T(0);
#pragma unroll 4
    for (int k = 0; k < nout; ++k) {  // out-of-block hops from L2
      const uint32_t c = sent[beg + k];
      const T a = (c & 0x8000u) ? -mag : mag;
      Vec<T, VEC> v;
      v.load(reinterpret_cast<const T *>(xb + ((c & 0x7FFFu) * rowb + colb)));
#pragma unroll
      for (int j = 0; j < VEC; ++j) acc[j] += a * v.v[j];
    }
#pragma unroll 4
    for (int k = nout; k < ntot; ++k) {  // in-block hops from the staged rows
      const uint32_t c = sent[beg + k];
      const T a = (c & 0x8000u) ? -mag : mag;
      const uint4 w = sx[((int)(c & 0x7FFFu) - r0) * LPR + ln];
      T v[VEC];
      __builtin_memcpy(v, &w, 16);
#pragma unroll
      for (int j = 0; j < VEC; ++j) acc[j] += a * v[j];
    }
    if constexpr (LZ) {
      if (col_ok) {
        // w is written once: a non-temporal store, so the stream does not
        // evict the strip of x from L2 (non-temporal loads of z and prev were
        // slower: 2.70 against 2.57 ms)
        const int64_t e = plane + (int64_t)(r0 + r) * Ddn + col;
        T zv[VEC], pv[VEC];
        {
          Vec<T, VEC> zq, pq;
          zq.load(z + e);
          if (rp) pq.load(static_cast<const T *>(a.prev) + e);
#pragma unroll
          for (int j = 0; j < VEC; ++j) {
            zv[j] = zq.v[j];
            pv[j] = rp ? pq.v[j] : T(0);
          }
        }
        const uint4 xw = sx[r * LPR + ln];
        T xv[VEC];
        __builtin_memcpy(xv, &xw, 16);
#pragma unroll
        for (int j = 0; j < VEC; ++j) {  // the Lanczos epilogue (epilogue_pv) with a streaming store
          T rr = acc[j] + zv[j], xx = xv[j];
          if (a.lazy) {
            rr *= bk.sx;
            xx *= bk.sx;
          }
          if (bk.bp != T(0)) rr -= bk.bp * pv[j];
          __builtin_nontemporal_store(rr, y + e + j);
          part.s0 += (double)xx * (double)rr;
        }
      }
    } else if (col_ok) {
      T *dst = yp + (int64_t)(r0 + r) * Ddn + col;
#pragma unroll
      for (int j = 0; j < VEC; ++j) __builtin_nontemporal_store(acc[j], dst + j);
    }
  }
  if constexpr (LZ) {
    const double tot = wg_sum(part.s0, red);
    if (threadIdx.x == 0) a.partial[i] = tot;
  }
}

// Persistent row kernel over the sorted compact table.  One workgroup of 1024
// threads per CU walks plane rows w, w + gridDim.x, ...  Per row:
//  A. every thread takes one slot per round (kSortedRound slots): the string of
//     the slot and the wave's sign pattern (both loaded once per kernel: slots
//     do not depend on the row), its hop targets (requested a round ahead, in
//     an order that spreads a half-wave's reads over the LDS bank pairs); the
//     trip count is the wave's shared hop count, so lanes do no padded work
//     beyond rounding to a multiple of four; the diagonal comes from a copy of
//     diagU in LDS.  Results stay in registers.  Then the next row is
//     requested into registers; it lands during B/C.
//  B/C. in two halves of kSortedHalf strings (the LDS left over holds half a
//     row): results are written to an LDS buffer at their string, then one
//     coalesced sweep adds the strip result and applies the epilogue (y, and
//     prev for the fused epilogues, loaded before the half's barrier).
//  D. the prefetched next row replaces the staged one.
constexpr int kSortedHalf = kSortedRound * (kSortedRounds / 2);

template <typename T, Epi E>
__global__ __launch_bounds__(1024) void row_kernel_sorted(HvTables t, const T *__restrict__ x,
                                                          T *__restrict__ y, EpiArgs a,
                                                          int64_t nrows) {
  constexpr int K = kCompactK, G = 4, NT = kSortedRound, R = kSortedRounds, H = R / 2;
  extern __shared__ double smem[];
  __shared__ double red[2][NT / kWave];
  __shared__ T s_diag[kMaxSites + 1];
  const int Ddn = (int)t.d_dn;
  const int64_t Dup = t.d_up;
  T *row = reinterpret_cast<T *>(smem);  // Ddn + 1 entries, zero slot at Ddn
  T *scr = row + Ddn + 1;                // kSortedHalf entries
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  if constexpr (E == Epi::Down) {
    if (a.flags && a.flags[0]) return;
  }
  int64_t w = blockIdx.x;
  if (w >= nrows) return;
  const int tid = threadIdx.x;
  const int nr = (int)(t.srt_np / NT);  // slot rounds (<= R)
  if (tid < t.n_diag) s_diag[tid] = Tabs<T>::diag(t)[tid];
  // Down with a fold: rows are staged as x = fold_w - c fold_v (the previous
  // Lanczos step's axpy, same arithmetic) and written out; ss = |x row|^2
  const bool fold = E == Epi::Down && a.fold;
  const T fc = fold ? fold_coef<T>(a) : T(0);
  const T *__restrict__ fw = static_cast<const T *>(a.fold_w);
  const T *__restrict__ fv = static_cast<const T *>(a.fold_v);
  T *__restrict__ fx = static_cast<T *>(a.fold_x);
  double ss = 0.0;
  {
    const T *__restrict__ xr = x + w * Ddn;  // uniform row base; lanes index it with tid
    T v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int64_t i = w * Ddn + j * NT + tid;
      v[j] = j * NT + tid < Ddn ? (fold ? fold_val(fw, fv, fc, i) : xr[j * NT + tid]) : T(0);
    }
    if (fold) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j * NT + tid < Ddn) {
          __builtin_nontemporal_store(v[j], fx + w * Ddn + j * NT + tid);
          ss += (double)v[j] * (double)v[j];
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j * NT + tid < Ddn) row[j * NT + tid] = v[j];
    if (tid == 0) row[Ddn] = T(0);
  }
  __syncthreads();
  const T h = Tabs<T>::mags(t)[0];
  const EpiK<T> bp = epi_consts<T, E>(a);
  const bool rp = reads_prev<T, E>(a, bp);
  // The slots do not depend on the row: each thread's slot strings and the
  // waves' sign patterns per round (both packed two per word) are loaded once,
  // so a round waits on one L2 round trip (its table chunks and the strings'
  // bits); loaded per round, the sign pattern had to arrive before the chunks
  // were requested, and the string before its bits.
  uint32_t dd[R / 2], pns[R / 2];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (j % 2 == 0) {
      dd[j / 2] = (uint32_t)Ddn | ((uint32_t)Ddn << 16);
      pns[j / 2] = 0;
    }
    if (j < nr) {
      const uint32_t d = t.srt_perm[j * NT + tid];
      const uint32_t pn = (uint32_t)__builtin_amdgcn_readfirstlane((int)t.srt_pn[(j * NT + tid) >> 6]);
      dd[j / 2] = (j & 1) ? ((dd[j / 2] & 0xFFFFu) | (d << 16)) : ((dd[j / 2] & 0xFFFF0000u) | d);
      pns[j / 2] |= pn << (16 * (j & 1));
    }
  }
  for (; w < nrows; w += gridDim.x) {
    const int64_t nx = w + gridDim.x;
    const int64_t u = w % Dup;
    const uint32_t su = t.s_up[u];
    const T on_u = t.has_onsite ? Tabs<T>::up_on(t)[u] : T(0);
    // An opaque zero added per row to the table base and the slot data:
    // otherwise the per-round addresses, sign selects and branch conditions
    // derived from them are hoisted out of the row loop and spilled.
    int zero = 0;
    asm volatile("" : "+s"(zero));
    const uint16_t *comp_b = t.srt_comp + zero;
    // A: hop sums of every slot of this thread.  A round's table chunks are
    // requested during the round before (3.52 -> 3.48 ms for the 4 x 4, 7 + 7
    // H.v), the first round's here.
    T res[R];
    uint4 qn[K / 8];
#pragma unroll
    for (int k = 0; k < K / 8; ++k) qn[k] = comp_chunk(comp_b, tid, k);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      res[j] = T(0);
      if (j < nr) {
        // uniform base of this round's table; lanes index it with tid
        const uint16_t *__restrict__ comp = comp_b + (int64_t)j * NT * K;
        const int d = (int)(((dd[j / 2] + (uint32_t)zero) >> (16 * (j & 1))) & 0xFFFFu);
        // wave-uniform sign pattern: slots [0, np_) positive, [np_, nt) negative
        const int pn = (int)(((pns[j / 2] + (uint32_t)zero) >> (16 * (j & 1))) & 0xFFFFu);
        const int np_ = pn & 0xFF, nt = np_ + (pn >> 8);
        uint4 q[K / 8];
#pragma unroll
        for (int k = 0; k < K / 8; ++k) q[k] = qn[k];
        const int dc = d < Ddn ? d : 0;
        const uint32_t sd = t.s_dn[dc];
        if (j + 1 < nr) {  // the next round's chunks, as many as its group needs
          const int pn1 = (int)(((pns[(j + 1) / 2] + (uint32_t)zero) >> (16 * ((j + 1) & 1))) & 0xFFFFu);
          const int nt1 = (pn1 & 0xFF) + (pn1 >> 8);
#pragma unroll
          for (int k = 0; k < K / 8; ++k)
            qn[k] = k * 8 < nt1 ? comp_chunk(comp + NT * K, tid, k) : make_uint4(0, 0, 0, 0);
        }
        T h0 = T(0), h1 = T(0);
#pragma unroll
        for (int g = 0; g < K / G; ++g) {
          if (g * G < nt) {
            T v[G];
#pragma unroll
            for (int i = 0; i < G; ++i) {
              const int k = g * G + i;
              const uint32_t e = reinterpret_cast<const uint32_t *>(&q[k / 8])[(k % 8) / 2];
              v[i] = row[(k & 1) ? (e >> 16) : (e & 0xFFFFu)];
            }
#pragma unroll
            for (int i = 0; i < G; ++i) {
              const int k = g * G + i;
              const T sg = k < np_ ? T(1) : T(-1);  // uniform: a scalar operand of the fma
              if (i & 1) h1 = fma(v[i], sg, h1); else h0 = fma(v[i], sg, h0);
            }
          }
        }
        T dg = s_diag[popc(su & sd)];
        if (t.has_onsite) dg = dg + on_u + Tabs<T>::dn_on(t)[dc];
        res[j] = dg * row[d] + h * (h0 + h1);
      }
      // keep the rounds apart: hoisting loads further only adds live registers
      __builtin_amdgcn_sched_barrier(0);
    }
    // next row into registers (consumed in D).  Issued after A: vmcnt retires in
    // order, so loads in flight during A would hold up its table loads.
    T pre[R];
    {
      const int64_t rn = nx < nrows ? nx : w;
      const T *__restrict__ xn = x + rn * Ddn;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int64_t i = rn * Ddn + j * NT + tid;
        pre[j] = j * NT + tid < Ddn ? (fold ? fold_val(fw, fv, fc, i) : xn[j * NT + tid]) : T(0);
      }
    }
    T *__restrict__ yr = y + w * Ddn;
    const T *__restrict__ pr = rp ? static_cast<const T *>(a.prev) + w * Ddn : nullptr;
    Partial2 acc2;
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      const int lo = hf * kSortedHalf;
      T yin[H], pv[H];
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const int d = (hf * H + j) * NT + tid;
        yin[j] = (E != Epi::Down && d < Ddn) ? yr[d] : T(0);
        pv[j] = (rp && d < Ddn) ? pr[d] : T(0);
      }
      // B: results of strings [lo, lo + kSortedHalf) into the LDS buffer
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int d = (int)((dd[j / 2] >> (16 * (j & 1))) & 0xFFFFu);
        if (d < Ddn && d >= lo && d < lo + kSortedHalf) scr[d - lo] = res[j];
      }
      __syncthreads();
      // C: coalesced sweep over the half
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const int d = (hf * H + j) * NT + tid;
        if (d < Ddn) {
          if constexpr (E == Epi::Down)
            __builtin_nontemporal_store(scr[d - lo], yr + d);
          else
            epilogue_pv<T, E>(yin[j] + scr[d - lo], row[d], pv[j], w * Ddn + d, y, a, bp, acc2);
        }
      }
      __syncthreads();
    }
    row_reduce<T, E>(acc2, a, w, red);
    if (fold) {  // |x row|^2 of this row; the next row's staged values are in pre
      const double r = wg_sum(ss, red[0]);
      if (tid == 0) a.fold_partial[w] = r;
      ss = 0.0;
      if (nx < nrows) {
#pragma unroll
        for (int j = 0; j < R; ++j)
          if (j * NT + tid < Ddn) {
            __builtin_nontemporal_store(pre[j], fx + nx * Ddn + j * NT + tid);
            ss += (double)pre[j] * (double)pre[j];
          }
      }
    }
    // D: next row
    if (nx < nrows) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j * NT + tid < Ddn) row[j * NT + tid] = pre[j];
    }
    __syncthreads();
  }
}

// General down table (any number of amplitude magnitudes): column-major uint32
// entries target << 8 | sign << 7 | magnitude index.
template <typename T, Epi E>
__global__ __launch_bounds__(1024) void row_kernel_ell(HvTables t, const T *__restrict__ x,
                                                       T *__restrict__ y, EpiArgs a) {
  extern __shared__ double smem[];
  __shared__ double red[2][1024 / kWave];
  T *row = reinterpret_cast<T *>(smem);
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int64_t wg = blockIdx.x;
  const int64_t Ddn = t.d_dn, Dup = t.d_up;
  const int64_t u = wg % Dup;
  const int64_t base = wg * Ddn;
  stage_row(row, x, base, Ddn);
  const uint32_t su = t.s_up[u];
  const T *__restrict__ diag = Tabs<T>::diag(t);
  const T *__restrict__ mags = Tabs<T>::mags(t);
  const EpiK<T> bp = epi_consts<T, E>(a);
  const T on_u = t.has_onsite ? Tabs<T>::up_on(t)[u] : T(0);
  Partial2 acc2;
  for (int64_t d = threadIdx.x; d < Ddn; d += blockDim.x) {
    const T xv = row[d];
    T dg = diag[popc(su & t.s_dn[d])];
    if (t.has_onsite) dg = dg + on_u + Tabs<T>::dn_on(t)[d];
    T acc = dg * xv;
    const int cnt = t.dn_cnt[d];
    for (int k = 0; k < cnt; ++k) {
      const uint32_t e = t.dn_ell[(int64_t)k * Ddn + d];
      const T m = mags[e & 127u];
      acc += ((e & 128u) ? -m : m) * row[e >> 8];
    }
    const T r = y[base + d] + acc;
    epilogue<T, E>(r, xv, base + d, y, a, bp, acc2);
  }
  row_reduce<T, E>(acc2, a, wg, red);
}

// ------------------------------------------------------------ generic path

// Thread per element of plane blockIdx.y; both tables read from global memory.
template <typename T, Epi E>
__global__ __launch_bounds__(kBlock) void generic_kernel(HvTables t, const T *__restrict__ x,
                                                         T *__restrict__ y, EpiArgs a) {
  __shared__ double red[2][kBlock / kWave];
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int64_t Ddn = t.d_dn, D = t.d_up * Ddn;
  const int64_t p = blockIdx.y;
  const T *__restrict__ xp = x + p * D;
  const EpiK<T> bp = epi_consts<T, E>(a);
  Partial2 acc2;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / Ddn, d = e - u * Ddn;
    T dg = Tabs<T>::diag(t)[popc(t.s_up[u] & t.s_dn[d])];
    if (t.has_onsite) dg = dg + Tabs<T>::up_on(t)[u] + Tabs<T>::dn_on(t)[d];
    const T xv = xp[e];
    T up = T(0);
    for (int32_t k = t.up_ptr[u]; k < t.up_ptr[u + 1]; ++k)
      up += Tabs<T>::up_amp(t)[k] * xp[(int64_t)t.up_tgt[k] * Ddn + d];
    T acc = dg * xv;
    for (int32_t k = t.dn_ptr[d]; k < t.dn_ptr[d + 1]; ++k)
      acc += Tabs<T>::dn_amp(t)[k] * xp[u * Ddn + t.dn_tgt[k]];
    epilogue<T, E>(up + acc, xv, p * D + e, y, a, bp, acc2);
  }
  if constexpr (E == Epi::Lanczos || E == Epi::Cheb) {
    const double s0 = wg_sum(acc2.s0, red[0]);
    double s1 = 0.0;
    if constexpr (E == Epi::Cheb) s1 = wg_sum(acc2.s1, red[1]);
    if (threadIdx.x == 0) {
      const int64_t slot = p * gridDim.x + blockIdx.x;
      if constexpr (E == Epi::Lanczos) {
        a.partial[slot] = s0;
      } else {
        a.partial[2 * slot] = s0;
        a.partial[2 * slot + 1] = s1;
      }
    }
  }
}

int grid_for(int64_t work, int cap) {
  int64_t g = (work + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

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

// Reference order is interleaved complex128; device order is planar (re, im).
template <bool kGather>
__global__ __launch_bounds__(kBlock) void perm_kernel(PermTables p, const double *__restrict__ src,
                                                      double *__restrict__ dst) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t D = p.d_up * p.d_dn;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const int64_t r = ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    if (r < 0 || r >= D) continue;  // unreachable for a valid sector; never index out of range
    if (kGather) {
      dst[e] = src[2 * r];
      dst[D + e] = src[2 * r + 1];
    } else {
      dst[2 * r] = src[e];
      dst[2 * r + 1] = src[D + e];
    }
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
                                                               double *__restrict__ x) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t D = p.d_up * p.d_dn;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const uint64_t i = (uint64_t)ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    const double u1 = unit_open(mix64(seed, 2 * i));
    const double u2 = unit_half(mix64(seed, 2 * i + 1));
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    x[e] = rad * cos(th);
    x[D + e] = rad * sin(th);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fill_phase_kernel(PermTables p, uint64_t seed, int nvec,
                                                            T *__restrict__ x) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t D = p.d_up * p.d_dn;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    const uint64_t i = (uint64_t)ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
    for (int r = 0; r < nvec; ++r) {
      const double ph = 6.283185307179586 * unit_half(mix64(seed, i * (uint64_t)nvec + r));
      x[(int64_t)(2 * r) * D + e] = (T)cos(ph);
      x[(int64_t)(2 * r + 1) * D + e] = (T)sin(ph);
    }
  }
}

// The same phases keyed by a device position -> reference index map (any
// basis layout): vector r only, planes (re, im) of D values.
template <typename T>
__global__ __launch_bounds__(kBlock) void fill_phase_ref_kernel(const int64_t *__restrict__ ref,
                                                                int64_t D, uint64_t seed,
                                                                int nvec, int r,
                                                                T *__restrict__ x) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const uint64_t i = (uint64_t)ref[e];
    const double ph = 6.283185307179586 * unit_half(mix64(seed, i * (uint64_t)nvec + r));
    x[e] = (T)cos(ph);
    x[D + e] = (T)sin(ph);
  }
}

// out0 = sum of partial[2 w], out1 = sum of partial[2 w + 1], w < n (fixed order)
// out[b] = the sum of in[b c, (b + 1) c), c = ceil(n / gridDim.x), in a fixed order.
__global__ __launch_bounds__(kBlock) void chunk_sum_kernel(const double *__restrict__ in, int64_t n,
                                                           double *__restrict__ out) {
  __shared__ double lds[kBlock / kWave];
  const int64_t c = (n + gridDim.x - 1) / gridDim.x, lo = (int64_t)blockIdx.x * c;
  const int64_t hi = lo + c < n ? lo + c : n;
  double s = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) s += in[i];
  s = wg_sum(s, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void reduce_pairs_kernel(const double *__restrict__ partial,
                                                              int64_t n, double *out0,
                                                              double *out1) {
  __shared__ double lds[kBlock / kWave];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t w = threadIdx.x; w < n; w += kBlock) {
    s0 += partial[2 * w];
    s1 += partial[2 * w + 1];
  }
  s0 = wg_sum(s0, lds);
  s1 = wg_sum(s1, lds);
  if (threadIdx.x == 0) {
    *out0 = s0;
    *out1 = s1;
  }
}

// ------------------------------------------------------------ vector steps

template <typename T>
__global__ __launch_bounds__(kBlock) void sumsq_partial_kernel(const T *__restrict__ x, int64_t n,
                                                               double *partial) {
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += (double)x[i] * (double)x[i];
  s = wg_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Planar complex dot <a|b> = sum conj(a) b: partial[blk] = Re, partial[nb + blk] = Im.
__global__ __launch_bounds__(kBlock) void cdot_partial_kernel(const double *__restrict__ a,
                                                              const double *__restrict__ b,
                                                              int64_t D, double *partial) {
  __shared__ double lds[kBlock / kWave];
  double re = 0.0, im = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < D; i += (int64_t)gridDim.x * kBlock) {
    const double ar = a[i], ai = a[D + i], br = b[i], bi = b[D + i];
    re += ar * br + ai * bi;
    im += ar * bi - ai * br;
  }
  re = wg_sum(re, lds);
  im = wg_sum(im, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = re;
    partial[gridDim.x + blockIdx.x] = im;
  }
}

__device__ double final_sum(const double *partial, int64_t n, double *lds) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBlock) s += partial[i];
  return wg_sum(s, lds);
}

__global__ __launch_bounds__(kBlock) void final_kernel(const double *partial, int64_t n, double *out,
                                                       int take_sqrt) {
  __shared__ double lds[kBlock / kWave];
  const double s = final_sum(partial, n, lds);
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(s) : s;
}

// y = x / den[0], x in FP64, y in T (FP32 Krylov vectors are rounded here).
template <typename T>
__global__ __launch_bounds__(kBlock) void scale_div_kernel(const double *__restrict__ x,
                                                           const double *den,
                                                           T *__restrict__ y, int64_t n) {
  const double dv = den[0];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = (T)(x[i] / dv);
}

__global__ __launch_bounds__(kBlock) void lanczos_finish_alpha_kernel(LanczosState st, int step,
                                                                      int64_t nparts) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double s = final_sum(st.partial, nparts, lds);
  if (threadIdx.x == 0) {
    st.alpha[step] = s;
    st.flags[1] = step + 1;
  }
}

// w -= alpha v ; partial = |w|^2.  lazy: v is stored unnormalised as
// beta[step - 1] v (step > 0), so alpha is divided by that norm.
template <typename T>
__global__ __launch_bounds__(kBlock) void lanczos_axpy_norm_kernel(const T *__restrict__ v,
                                                                   T *__restrict__ w, int64_t n,
                                                                   LanczosState st, int step,
                                                                   int lazy) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const T a = (lazy && step > 0) ? (T)(st.alpha[step] / st.beta[step - 1]) : (T)st.alpha[step];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T q = w[i] - a * v[i];
    w[i] = q;
    s += (double)q * (double)q;
  }
  s = wg_sum(s, lds);
  if (threadIdx.x == 0) st.partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void lanczos_finish_beta_kernel(LanczosState st, int step,
                                                                     int64_t nparts) {
  __shared__ double lds[kBlock / kWave];
  if (st.flags[0]) return;
  const double s = final_sum(st.partial, nparts, lds);
  if (threadIdx.x == 0) {
    const double b = sqrt(s);
    st.beta[step] = b;
    if (b <= kTol) st.flags[0] = 1;  // breakdown: later kernels of this run are no-ops
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void lanczos_normalize_kernel(T *__restrict__ w, int64_t n,
                                                                   LanczosState st, int step) {
  if (st.flags[0]) return;
  const T b = (T)st.beta[step];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    w[i] = w[i] / b;
}

// Reorthogonalisation against m stored Krylov vectors V_j (planar, stride
// 2 D): partial[(2 j + c) * gridDim.x + blk] = Re / Im of the block's share of
// <V_j|w>.  Fixed blocks and a fixed tree, so the coefficients are reproducible.
__global__ __launch_bounds__(kBlock) void proj_partial_kernel(const double *__restrict__ V,
                                                              const double *__restrict__ w,
                                                              int64_t D, double *partial,
                                                              const int *flags) {
  __shared__ double lds[kBlock / kWave];
  if (flags[0]) return;
  const int j = blockIdx.y;
  const double *vj = V + (int64_t)j * 2 * D;
  double re = 0.0, im = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const double vr = vj[e], vi = vj[D + e], wr = w[e], wi = w[D + e];
    re += vr * wr + vi * wi;
    im += vr * wi - vi * wr;
  }
  re = wg_sum(re, lds);
  im = wg_sum(im, lds);
  if (threadIdx.x == 0) {
    partial[(2 * (int64_t)j) * gridDim.x + blockIdx.x] = re;
    partial[(2 * (int64_t)j + 1) * gridDim.x + blockIdx.x] = im;
  }
}

// c[2 j + c] = sum of the nb partials of coefficient (j, c); one block per j.
__global__ __launch_bounds__(kBlock) void proj_final_kernel(const double *partial, int nb,
                                                            double *c, const int *flags) {
  __shared__ double lds[kBlock / kWave];
  if (flags[0]) return;
  const int j = blockIdx.x;
  const double re = final_sum(partial + (2 * (int64_t)j) * nb, nb, lds);
  const double im = final_sum(partial + (2 * (int64_t)j + 1) * nb, nb, lds);
  if (threadIdx.x == 0) {
    c[2 * j] = re;
    c[2 * j + 1] = im;
  }
}

// w -= sum_j c_j V_j (complex c_j, j in order).
__global__ __launch_bounds__(kBlock) void proj_sub_kernel(const double *__restrict__ V, int m,
                                                          const double *__restrict__ c,
                                                          double *__restrict__ w, int64_t D,
                                                          const int *flags) {
  if (flags[0]) return;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < m; ++j) {
      const double *vj = V + (int64_t)j * 2 * D;
      const double cr = c[2 * j], ci = c[2 * j + 1], vr = vj[e], vi = vj[D + e];
      sr += cr * vr - ci * vi;
      si += cr * vi + ci * vr;
    }
    w[e] -= sr;
    w[D + e] -= si;
  }
}

// w /= beta ; x += y * w   (complex y, planar vectors; w in T, x accumulated in FP64)
template <typename T>
__global__ __launch_bounds__(kBlock) void recon_update_kernel(T *__restrict__ w,
                                                              double *__restrict__ x, double beta,
                                                              double yr, double yi, int64_t D) {
  const T bt = (T)beta;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const T wr = w[e] / bt, wi = w[D + e] / bt;
    w[e] = wr;
    w[D + e] = wi;
    x[e] += (double)wr * yr - (double)wi * yi;
    x[D + e] += (double)wr * yi + (double)wi * yr;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scale_complex_kernel(const T *__restrict__ x,
                                                               double yr, double yi,
                                                               double *__restrict__ o, int64_t D) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const double xr = x[e], xi = x[D + e];
    o[e] = xr * yr - xi * yi;
    o[D + e] = xr * yi + xi * yr;
  }
}

// One power-method step on the planar vectors v, w (w = H v on entry):
// w <- sign w + shift v in place, with per-block partial sums of the three
// quantities power_method reads (linear_operator_utils.h:18-50):
// Re sum_i v_i w_i and Re sum_i v_i v_i (arma::dot, no conjugation) and
// sum_i |w_i|^2.
__global__ __launch_bounds__(kBlock) void power_step_kernel(const double *__restrict__ v,
                                                            double *__restrict__ w, int64_t D,
                                                            double sign, double shift,
                                                            double *__restrict__ partial) {
  __shared__ double red[kBlock / kWave];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const double vr = v[e], vi = v[D + e];
    const double wr = sign * w[e] + shift * vr, wi = sign * w[D + e] + shift * vi;
    w[e] = wr;
    w[D + e] = wi;
    a0 += vr * wr - vi * wi;
    a1 += vr * vr - vi * vi;
    a2 += wr * wr + wi * wi;
  }
  a0 = wg_sum(a0, red);
  a1 = wg_sum(a1, red);
  a2 = wg_sum(a2, red);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = a0;
    partial[3 * blockIdx.x + 1] = a1;
    partial[3 * blockIdx.x + 2] = a2;
  }
}

// One Chebyshev step of the KPM projector over both planes (n = 2 D doubles):
// T_next = s (y - b T) (first step) or s (y - b T) - T_prev, with y = H T,
// written over T_prev; result += c T_next.
__global__ __launch_bounds__(kBlock) void cheb_accum_kernel(const double *__restrict__ y,
                                                            const double *__restrict__ t,
                                                            double *__restrict__ tp,
                                                            double *__restrict__ res, int64_t n,
                                                            double sc, double b, int first,
                                                            double c) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    double tn = sc * (y[e] - b * t[e]);
    if (!first) tn -= tp[e];
    tp[e] = tn;
    res[e] += c * tn;
  }
}

// x = sum_{j < m} y_j V_j for Krylov vectors V_j kept in HBM (planar, stride
// 2 D, in T) and complex coefficients y (re, im pairs on the device), summed
// in FP64 in the order j = 0, 1, ...: one pass over the m vectors in place of
// regenerating them with m - 1 more H.v.
template <typename T>
__global__ __launch_bounds__(kBlock) void basis_combine_kernel(const T *__restrict__ V, int m,
                                                               int64_t D,
                                                               const double *__restrict__ y,
                                                               double *__restrict__ x) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    double xr = 0.0, xi = 0.0;
    for (int j = 0; j < m; ++j) {
      const double vr = V[(int64_t)j * 2 * D + e], vi = V[(int64_t)j * 2 * D + D + e];
      const double yr = y[2 * j], yi = y[2 * j + 1];
      xr += vr * yr - vi * yi;
      xi += vr * yi + vi * yr;
    }
    x[e] = xr;
    x[D + e] = xi;
  }
}

// out[r] = sum over the slots of planes 2r, 2r+1 of partial[2 slot]; out[nvec + r] the odd entries
__global__ __launch_bounds__(kBlock) void reduce_cheb_kernel(const double *partial, int64_t rows,
                                                             int nvec, double *out) {
  __shared__ double lds[kBlock / kWave];
  const int r = blockIdx.x;
  for (int j = 0; j < 2; ++j) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < 2 * rows; i += kBlock)
      s += partial[2 * ((int64_t)2 * r * rows + i) + j];
    s = wg_sum(s, lds);
    if (threadIdx.x == 0) out[j * nvec + r] = s;
  }
}

// ------------------------------------------------------------ momentum sectors

// H.v of the factorized momentum-space operator (momentum.cpp).  One workgroup
// per output up string g (block p = its momentum, row r inside the block); its
// columns are the down strings of momentum K - p.  For each q != 0 the source
// rows of g under rho_{q,up} (block p - q) are first combined in LDS with their
// signs, comb[d] = sum_u s_u x[u, d] (coalesced loads, both planes), then each
// column gathers its rho_{-q,dn} sources from comb: sum_u sum_d s_u s_d x[u, d]
// = sum_d s_d comb[d].  The sum is scaled by U/N and the diagonal t (eps_up +
// eps_dn) + (U/N) N_up N_dn added, then the epilogue runs with one partial sum
// per plane.
template <Epi E, int CPT>
__global__ __launch_bounds__(kBlock) void mom_hv_kernel(MomTables t, const double *__restrict__ x,
                                                        double *__restrict__ y, EpiArgs a) {
  extern __shared__ double smem[];  // 2 * max_cols: re then im of the combined source row
  __shared__ double red[2][kBlock / kWave];
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  // each XCD takes a contiguous range of output rows (workgroups are dispatched
  // round-robin over the 8 XCDs), so rows of one block share its L2
  const int per = (int)((t.nrows + 7) / 8);
  const int g = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (g >= t.nrows) return;  // uniform over the workgroup
  const int N = t.N, tid = threadIdx.x;
  const int p = t.up_grp[g];
  const int nc = t.ncols[p];
  const int64_t D = t.D;
  const int64_t base = t.blk_off[p] + (int64_t)t.up_loc[g] * nc;
  const int dg0 = t.dn_off[t.msub[t.K * N + p]];  // first down string of the columns
  double ar[CPT], ai[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) ar[j] = ai[j] = 0.0;
  for (int q = 1; q < N; ++q) {
    const int ps = t.msub[p * N + q];
    const int ncs = t.ncols[ps];
    const int64_t gq = (int64_t)g * N + q;
    const int cu = t.up_cnt[gq];
    if (cu == 0) continue;  // uniform over the workgroup
    const uint16_t *__restrict__ us = t.up_src + gq * t.mu;
    const double *__restrict__ xb = x + t.blk_off[ps];
    __syncthreads();  // the previous comb is no longer read
    for (int c = tid; c < ncs; c += kBlock) {
      double cr = 0.0, ci = 0.0;
      for (int i = 0; i < cu; ++i) {
        const uint32_t e = us[i];
        const int64_t o = (int64_t)(e & 0x7FFFu) * ncs + c;
        const double vr = xb[o], vi = xb[D + o];
        if (e & 0x8000u) {
          cr -= vr;
          ci -= vi;
        } else {
          cr += vr;
          ci += vi;
        }
      }
      smem[c] = cr;
      smem[ncs + c] = ci;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = tid + j * kBlock;
      if (c < nc) {
        const int64_t hq = (int64_t)(dg0 + c) * N + q;
        const int cd = t.dn_cnt[hq];
        const uint16_t *__restrict__ ds = t.dn_src + hq * t.md;
        for (int l = 0; l < cd; ++l) {
          const uint32_t e2 = ds[l];
          const int lc = (int)(e2 & 0x7FFFu);
          if (e2 & 0x8000u) {
            ar[j] -= smem[lc];
            ai[j] -= smem[ncs + lc];
          } else {
            ar[j] += smem[lc];
            ai[j] += smem[ncs + lc];
          }
        }
      }
    }
  }
  const double uo = t.U / N, eu = t.eps_up[g];
  const EpiK<double> bp = epi_consts<double, E>(a);
  Partial2 accr, acci;
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = tid + j * kBlock;
    if (c < nc) {
      const int64_t i = base + c;
      const double xr = x[i], xi = x[D + i];
      const double dg = t.t * (eu + t.eps_dn[dg0 + c]) + t.q0;
      epilogue<double, E>(dg * xr + uo * ar[j], xr, i, y, a, bp, accr);
      epilogue<double, E>(dg * xi + uo * ai[j], xi, D + i, y, a, bp, acci);
    }
  }
  if constexpr (E == Epi::Lanczos) {
    const double sr = wg_sum(accr.s0, red[0]);
    const double si = wg_sum(acci.s0, red[1]);
    if (tid == 0) {
      a.partial[g] = sr;
      a.partial[t.nrows + g] = si;
    }
  }
}

// The same operator as mom_hv_kernel, one plane at a time (H is real, so the
// planes do not mix), in a persistent launch of one workgroup per CU.  Work
// items are (block p, plane); XCD x takes items x, x + 8, ... (workgroups go
// round-robin over the 8 XCDs), and the W workgroups of an XCD split the rows of
// the block, rows w, w + W, ... (at most kLockRows each).  All workgroups of an
// XCD run through q = 1 .. N-1 at about the same pace, so at any time the XCD
// reads about one source block plane (4 MB at 4x4 7+7), each of whose rows is
// combined into about four output rows.
//
// The workgroup is two groups of 6 waves; group h owns rows kLockHalf h ..
// kLockHalf h + kLockHalf - 1 and, per thread, kLockCols columns of them, whose
// sums stay in registers.  For each q a group first stages its rows: each row's
// signed up-source rows combined into an LDS row, comb[j][d] = sum_u s_u x[u, d]
// (one wave per row, loads through a buffer resource, a batch of 64-column
// chunks in flight); then it gathers: each owned column adds its down sources,
// sum_d s_d comb[j][d], for all of the group's rows at once.  The groups run one
// step apart, so one group's loads overlap the other's LDS gathers, with one
// barrier per step.  The rows' up-source lists for every q are staged in LDS
// when the item starts.
constexpr int kLockBlock = 768;  // threads per workgroup
constexpr int kLockGroup = 384;  // threads per group (6 waves)
constexpr int kLockCols = 2;     // columns per thread: at most 768 columns
constexpr int kLockRows = 24;    // rows per workgroup and item
constexpr int kLockHalf = 12;    // rows per group
constexpr int kLockS = 768;      // LDS doubles per comb row
constexpr int kLockSrc = 8;      // source list entries per (string, q)
constexpr int kLockQ = 16;       // momenta per item whose up-source lists are staged

constexpr size_t lock_lds_bytes() {  // the groups' comb rows, then the staged lists and counts
  return (size_t)kLockRows * kLockS * sizeof(double) +
         (size_t)kLockRows * kLockQ * (kLockSrc * sizeof(uint16_t) + sizeof(uint32_t));
}

// Word 3 of a raw buffer resource (32-bit data format, no swizzle): loads
// through it take a per-lane byte offset and a uniform one, no 64-bit address.
constexpr int kBufWord3 = 0x00020000;

// Buffer resource over `bytes` bytes from p, built from wave-uniform words so
// that it lives in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lock_rsrc(const double *p, int bytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  void *base = reinterpret_cast<void *>(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, __builtin_amdgcn_readfirstlane(bytes),
                                           kBufWord3);
}

// Loads of chunks c0 .. c0 + CB - 1 (64 columns each) of the NS source rows at
// byte offsets so[i]; chunks past the row and lanes past nc load the last column
// (in bounds, unused), so a batch is always CB * NS loads.
template <int NS, int CB>
__device__ __forceinline__ void lock_load(__amdgpu_buffer_rsrc_t rs, const int (&so)[kLockSrc], int nc,
                                          int c0, int lane, double (&v)[CB][NS]) {
#pragma unroll
  for (int b = 0; b < CB; ++b) {
    const int off = min((c0 + b) * kWave + lane, nc - 1) * 8;
#pragma unroll
    for (int i = 0; i < NS; ++i)
      v[b][i] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, so[i], 0));
  }
}

// Signed sums of loaded chunks in list order (exact +-1 products) into crow.
template <int NS, int CB>
__device__ __forceinline__ void lock_sum(const double (&sg)[kLockSrc], int nc, int c0, int lane,
                                         const double (&v)[CB][NS], double *__restrict__ crow) {
#pragma unroll
  for (int b = 0; b < CB; ++b) {
    const int d = (c0 + b) * kWave + lane;
    if (d < nc) {
      double sum = sg[0] * v[b][0];
#pragma unroll
      for (int i = 1; i < NS; ++i) sum = fma(sg[i], v[b][i], sum);
      crow[d] = sum;
    }
  }
}

// Chunks per batch for NS sources: at most 16 loads, and 12 chunks (a 715 to
// 768 column row) in an even number of batches where possible.
template <int NS>
constexpr int lock_cb() {
  return NS <= 2 ? 6 : (NS <= 5 ? 3 : 2);
}

// One row of comb: the NS signed source rows summed over the nc columns, in
// batches of CB 64-column chunks, the next batch's loads issued before the
// current batch is summed (two batches in flight).
template <int NS>
__device__ __forceinline__ void lock_stage_row(__amdgpu_buffer_rsrc_t rs, const int (&so)[kLockSrc],
                                               const double (&sg)[kLockSrc], int nc, int lane,
                                               double *__restrict__ crow) {
  constexpr int CB = lock_cb<NS>();
  const int ncol = (nc + kWave - 1) / kWave * kWave;  // columns of the whole chunks
  double va[CB][NS], vb[CB][NS];
  lock_load<NS, CB>(rs, so, nc, 0, lane, va);
  int c0 = 0;
  for (; (c0 + CB) * kWave < ncol; c0 += 2 * CB) {
    lock_load<NS, CB>(rs, so, nc, c0 + CB, lane, vb);
    lock_sum<NS, CB>(sg, nc, c0, lane, va, crow);
    lock_load<NS, CB>(rs, so, nc, c0 + 2 * CB, lane, va);
    lock_sum<NS, CB>(sg, nc, c0 + CB, lane, vb, crow);
  }
  if (c0 * kWave < ncol) lock_sum<NS, CB>(sg, nc, c0, lane, va, crow);
}

// Staging of one group: rows k = k0 + j (j < kLockHalf, k < nr) of the
// workgroup, momentum q, into comb rows j; gw is the wave within the group.
__device__ __forceinline__ void lock_stage_half(const double *__restrict__ xs, int ncs, int nrs, int q,
                                                int k0, int nr, int gw, int lane, const uint4 *ent,
                                                const uint32_t *ecnt, double *__restrict__ comb) {
  if (ncs <= 0) return;
  const __amdgpu_buffer_rsrc_t rs = lock_rsrc(xs, (int)((int64_t)nrs * ncs * 8));
  const int k1 = min(k0 + kLockHalf, nr);
  for (int k = k0 + gw; k < k1; k += kLockGroup / kWave) {
    const uint4 ev = ent[k * kLockQ + q];
    const int cu = __builtin_amdgcn_readfirstlane(ecnt[k * kLockQ + q]);
    const uint32_t ew[4] = {(uint32_t)__builtin_amdgcn_readfirstlane(ev.x),
                            (uint32_t)__builtin_amdgcn_readfirstlane(ev.y),
                            (uint32_t)__builtin_amdgcn_readfirstlane(ev.z),
                            (uint32_t)__builtin_amdgcn_readfirstlane(ev.w)};
    int so[kLockSrc];
    double sg[kLockSrc];
#pragma unroll
    for (int i = 0; i < kLockSrc; ++i) {
      const uint32_t e = (ew[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
      so[i] = (int)(e & 0x7FFFu) * ncs * 8;
      sg[i] = (e & 0x8000u) ? -1.0 : 1.0;
    }
    double *__restrict__ crow = comb + (k - k0) * kLockS;
    switch (cu) {
      case 1: lock_stage_row<1>(rs, so, sg, ncs, lane, crow); break;
      case 2: lock_stage_row<2>(rs, so, sg, ncs, lane, crow); break;
      case 3: lock_stage_row<3>(rs, so, sg, ncs, lane, crow); break;
      case 4: lock_stage_row<4>(rs, so, sg, ncs, lane, crow); break;
      case 5: lock_stage_row<5>(rs, so, sg, ncs, lane, crow); break;
      case 6: lock_stage_row<6>(rs, so, sg, ncs, lane, crow); break;
      case 7: lock_stage_row<7>(rs, so, sg, ncs, lane, crow); break;
      case 8: lock_stage_row<8>(rs, so, sg, ncs, lane, crow); break;
      default:  // no sources under this q
        for (int d = lane; d < ncs; d += kWave) crow[d] = 0.0;
    }
  }
}

// Gathering of one group: for each owned column c < nc, its down sources under
// q (8 packed entries, cd of them valid) read from the group's comb rows.
__device__ __forceinline__ void lock_gather_half(const MomTables &t, int q, int dg0, int nc, int gt,
                                                 const double *__restrict__ comb,
                                                 double (&acc)[kLockCols][kLockHalf]) {
  const int N = t.N;
  uint32_t de[kLockCols][4];
  int cd[kLockCols];
#pragma unroll
  for (int j = 0; j < kLockCols; ++j) {
    const int c = gt + j * kLockGroup;
    de[j][0] = de[j][1] = de[j][2] = de[j][3] = 0u;
    cd[j] = 0;
    if (c < nc) {
      const int64_t hq = (int64_t)(dg0 + c) * N + q;
      cd[j] = t.dn_cnt[hq];
      const uint4 v = *reinterpret_cast<const uint4 *>(t.dn_src + hq * kLockSrc);
      de[j][0] = v.x;
      de[j][1] = v.y;
      de[j][2] = v.z;
      de[j][3] = v.w;
    }
  }
#pragma unroll
  for (int j = 0; j < kLockCols; ++j) {
#pragma unroll
    for (int l = 0; l < kLockSrc; ++l) {
      const uint32_t e = (de[j][l >> 1] >> ((l & 1) * 16)) & 0xFFFFu;
      if (l < cd[j]) {
        const double sg = (e & 0x8000u) ? -1.0 : 1.0;
        const double *__restrict__ cb = comb + (e & 0x7FFFu);
#pragma unroll
        for (int r = 0; r < kLockHalf; ++r) acc[j][r] = fma(sg, cb[r * kLockS], acc[j][r]);
      }
    }
  }
}

template <Epi E>
__global__ __launch_bounds__(kLockBlock) void mom_hv_lock_kernel(MomTables t, const double *__restrict__ x,
                                                                 double *__restrict__ y, EpiArgs a) {
  constexpr int kWaves = kLockBlock / kWave;
  extern __shared__ double lds[];  // comb: [kLockRows][kLockS], group h's rows from h * kLockHalf
  uint4 *ent = reinterpret_cast<uint4 *>(lds + kLockRows * kLockS);        // [row][q]
  uint32_t *ecnt = reinterpret_cast<uint32_t *>(ent + kLockRows * kLockQ);  // [row][q]
  __shared__ double red[kWaves];
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int N = t.N, tid = threadIdx.x, lane = tid & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int grp = wave / (kLockGroup / kWave);        // 0 or 1, wave-uniform
  const int gw = wave - grp * (kLockGroup / kWave);  // wave within the group
  const int gt = tid - grp * kLockGroup;             // thread within the group
  const int k0 = grp * kLockHalf;                    // the group's first row
  double *__restrict__ comb = lds + k0 * kLockS;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3, W = gridDim.x >> 3;
  const int64_t D = t.D;
  const double uo = t.U / N;
  const EpiK<double> bp = epi_consts<double, E>(a);
  for (int item = xcd; item < 2 * N; item += 8) {
    const int p = item >> 1, pl = item & 1;
    const int r0 = t.up_off[p];
    const int nblk = t.up_off[p + 1] - r0;
    const int nr = nblk > w ? (nblk - w + W - 1) / W : 0;  // rows w, w + W, ... of block p
    const int nc = t.ncols[p];
    const int dg0 = t.dn_off[t.msub[t.K * N + p]];  // first down string of the columns
    const double *__restrict__ xp = x + (int64_t)pl * D;
    __syncthreads();  // the previous item no longer reads the staged lists
    for (int i = tid; i < nr * N; i += kLockBlock) {
      const int k = i / N, q = i - k * N;
      const int64_t gq = (int64_t)(r0 + w + k * W) * N + q;
      ent[k * kLockQ + q] = *reinterpret_cast<const uint4 *>(t.up_src + gq * kLockSrc);
      ecnt[k * kLockQ + q] = t.up_cnt[gq];
    }
    __syncthreads();
    double acc[kLockCols][kLockHalf];
#pragma unroll
    for (int j = 0; j < kLockCols; ++j)
#pragma unroll
      for (int k = 0; k < kLockHalf; ++k) acc[j][k] = 0.0;
    // Each group alternates staging and gathering, q = 1 .. N-1: its own step
    // u = 2 (q - 1) stages q, u + 1 gathers it.  Group 1 runs one step behind
    // group 0, so in every workgroup step one group stages while the other gathers.
    if (nr > 0) {  // uniform
      for (int s = 0; s <= 2 * N - 2; ++s) {
        const int u = s - grp;
        if (u >= 0 && k0 < nr) {  // uniform over the group
          const int q = u / 2 + 1;
          if ((u & 1) == 0) {
            if (q < N) {
              const int ps = t.msub[p * N + q];
              lock_stage_half(xp + t.blk_off[ps], t.ncols[ps], t.up_off[ps + 1] - t.up_off[ps], q, k0,
                              nr, gw, lane, ent, ecnt, comb);
            }
          } else {
            lock_gather_half(t, q, dg0, nc, gt, comb, acc);
          }
        }
        __syncthreads();
      }
    }
    Partial2 ap;
    const int64_t base = (int64_t)pl * D + t.blk_off[p];
#pragma unroll
    for (int j = 0; j < kLockCols; ++j) {
      const int c = gt + j * kLockGroup;
      if (c < nc) {
        const double ed = t.eps_dn[dg0 + c];
#pragma unroll
        for (int r = 0; r < kLockHalf; ++r) {
          const int k = k0 + r;
          if (k < nr) {
            const int row = w + k * W;
            const int64_t i = base + (int64_t)row * nc + c;
            const double xv = x[i];
            const double dg = t.t * (t.eps_up[r0 + row] + ed) + t.q0;
            epilogue<double, E>(dg * xv + uo * acc[j][r], xv, i, y, a, bp, ap);
          }
        }
      }
    }
    if constexpr (E == Epi::Lanczos) {
      const double sum = wg_sum(ap.s0, red);
      if (tid == 0) a.partial[item * W + w] = sum;
    }
  }
}

// One-body momentum transfer sum_s w(s) c+_{s+Q} c_s (both spins), sector S ->
// T = S + Q (conductivity.cpp).  One workgroup per up string g (momentum p in
// both sectors): target row g of block p.  Up moves read rows of the source
// block p - Q at the same column (its column group is the target's); down moves
// read the source row g of block p, staged in LDS, at the gathered columns.
__global__ __launch_bounds__(kBlock) void transfer_kernel(TransferTables t, const double *__restrict__ x,
                                                          double *__restrict__ y) {
  extern __shared__ double smem[];  // 2 * max_cols: re then im of the source row (down moves)
  const int g = blockIdx.x, N = t.N, tid = threadIdx.x;
  const int p = t.up_grp[g];
  const int ncT = t.nc_t[p];
  if (ncT == 0) return;  // uniform over the workgroup
  const int64_t baseT = t.blk_t[p] + (int64_t)t.up_loc[g] * ncT;
  const int ps = t.msub[p * N + t.Q];
  const int64_t blk_ps = t.blk_s[ps];
  const int ncS = t.nc_s[p];
  const int64_t rowS = t.blk_s[p] + (int64_t)t.up_loc[g] * ncS;
  for (int c = tid; c < ncS; c += kBlock) {
    smem[c] = x[rowS + c];
    smem[ncS + c] = x[t.DS + rowS + c];
  }
  __syncthreads();
  const int cu = t.up_cnt[g];
  const uint16_t *__restrict__ us = t.up_src + (int64_t)g * t.mu;
  const double *__restrict__ uw = t.up_w + (int64_t)g * t.mu;
  const int dgT0 = t.dn_off_t[t.msub[t.KT * N + p]];  // target down group K_T - p
  for (int c = tid; c < ncT; c += kBlock) {
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < cu; ++i) {
      const int64_t o = blk_ps + (int64_t)us[i] * ncT + c;
      ar += uw[i] * x[o];
      ai += uw[i] * x[t.DS + o];
    }
    const int64_t h = dgT0 + c;
    const int cd = t.dn_cnt[h];
    for (int l = 0; l < cd; ++l) {
      const int ld = t.dn_src[h * t.md + l];
      const double w = t.dn_w[h * t.md + l];
      ar += w * smem[ld];
      ai += w * smem[ncS + ld];
    }
    y[baseT + c] = ar;
    y[t.DT + baseT + c] = ai;
  }
}

// planar device order <-> interleaved reference order through the sector's map
template <bool TO_DEVICE>
__global__ __launch_bounds__(kBlock) void mom_perm_kernel(const int64_t *__restrict__ ref_of_pos,
                                                          int64_t D, const double *__restrict__ src,
                                                          double *__restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = ref_of_pos[e];
    if constexpr (TO_DEVICE) {
      dst[e] = src[2 * r];
      dst[D + e] = src[2 * r + 1];
    } else {
      dst[2 * r] = src[e];
      dst[2 * r + 1] = src[D + e];
    }
  }
}

// seeded N(0,1) per reference index, the generator of fill_gaussian_kernel
__global__ __launch_bounds__(kBlock) void mom_fill_gaussian_kernel(const int64_t *__restrict__ ref_of_pos,
                                                                   int64_t D, uint64_t seed,
                                                                   double *__restrict__ x) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const uint64_t i = (uint64_t)ref_of_pos[e];
    const double u1 = unit_open(mix64(seed, 2 * i));
    const double u2 = unit_half(mix64(seed, 2 * i + 1));
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    x[e] = rad * cos(th);
    x[D + e] = rad * sin(th);
  }
}

// ------------------------------------------------------------ compiled expressions

static_assert(kExprChunk == kBlock, "one thread per column of an expression work unit");

// colex rank of a string: sum over its set bits (ascending, i = 0, 1, ...) of
// C(position, i + 1), the rank Sector::colex_rank gives
__device__ __forceinline__ int64_t colex_rank_dev(const BinomLds &B, uint32_t bits) {
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

__device__ __forceinline__ void cfma(double ar, double ai, double xr, double xi, double &yr,
                                     double &yi) {
  yr += ar * xr - ai * xi;
  yi += ar * xi + ai * xr;
}

// Up-only part of a compiled expression on a real-space sector, strip by strip
// (the layout of up_strip_kernel, both planes at once because the amplitudes
// are complex): a strip is 16 columns of all D_up rows of both planes (2.9 MB
// at D_up = 11,440), resident in one XCD's L2 while all its row chunks are
// produced; 16 lanes per row, 32 rows per workgroup, the rows' source lists
// (byte offset of the source row, amplitude) staged in LDS.  Writes y = the
// up-only sum with non-temporal stores.
template <int MAXH>
__global__ __launch_bounds__(kBlock) void expr_strip_kernel(ExprTables t, const double *__restrict__ x,
                                                            double *__restrict__ y, int nJ, int nC,
                                                            const int *flags) {
  constexpr int LPR = 16, RPW = kWave / LPR, RC = 32;
  __shared__ uint32_t s_off[RC][MAXH];
  __shared__ double s_ar[RC][MAXH], s_ai[RC][MAXH];
  __shared__ int s_n[RC];
  if (flags && flags[0]) return;
  const int i = blockIdx.x;
  const int xcd = i & 7, q = i >> 3;
  const int s = xcd + 8 * (q / nC);
  const int c = q % nC;
  if (s >= nJ) return;
  const int u_begin = c * RC;
  const int u_end = (int)min((int64_t)u_begin + RC, t.d_up);
  for (int e = threadIdx.x; e < RC * MAXH; e += kBlock) {
    const int r = e / MAXH, k = e % MAXH;
    const int u = u_begin + r;
    if (u < u_end) {
      const int64_t b = t.uo_ptr[u];
      const int n = (int)(t.uo_ptr[u + 1] - b);
      if (k == 0) s_n[r] = n;
      if (k < n) {
        s_off[r][k] = (uint32_t)(t.uo_base[b + k] * (int64_t)sizeof(double));
        s_ar[r][k] = t.uo_amp[2 * (b + k)];
        s_ai[r][k] = t.uo_amp[2 * (b + k) + 1];
      }
    } else if (k == 0) {
      s_n[r] = 0;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int64_t Ddn = t.d_dn, D = t.D;
  const int64_t col = (int64_t)s * LPR + (lane % LPR);
  const bool col_ok = col < Ddn;
  const char *__restrict__ xr = reinterpret_cast<const char *>(x);
  const char *__restrict__ xi = reinterpret_cast<const char *>(x + D);
  const uint32_t colb = (uint32_t)(col * (int64_t)sizeof(double));
  const int sub = lane / LPR;
  for (int rb = wv * RPW; rb < RC; rb += (kBlock / kWave) * RPW) {
    const int r = rb + sub;
    const int u = u_begin + r;
    const bool ok = col_ok && u < u_end;
    const int n = ok ? s_n[r] : 0;
    double ar = 0.0, ai = 0.0;
#pragma unroll 4
    for (int k = 0; k < n; ++k) {
      const uint32_t o = s_off[r][k] + colb;
      const double vr = *reinterpret_cast<const double *>(xr + o);
      const double vi = *reinterpret_cast<const double *>(xi + o);
      cfma(s_ar[r][k], s_ai[r][k], vr, vi, ar, ai);
    }
    if (ok) {
      const int64_t e = (int64_t)u * Ddn + col;
      __builtin_nontemporal_store(ar, y + e);
      __builtin_nontemporal_store(ai, y + D + e);
    }
  }
}

// H.v of a compiled expression (expression.cpp).  One workgroup per (row,
// chunk of kExprChunk columns); row = up string u, column = down string d of
// the row's group.  Work units are spread so that the chunks of one row run on
// one XCD (unit = xcd * ceil(nwg / 8) + blockIdx / 8) and its x row stays in
// that L2.  Per element: the diagonal (per-row + per-column values), the
// up-only sources (row bases shared by the workgroup, same column; or, with
// STRIP, the sum expr_strip_kernel left in y), the down-only sources (columns
// of the same row), and the mixed terms: for each up part the row meets
// (uniform), its down parts tested on d.  One partial sum per work unit and
// plane for the Lanczos epilogue.
template <Epi E, bool STRIP>
__global__ __launch_bounds__(kBlock) void expr_hv_kernel(ExprTables t, const double *__restrict__ x,
                                                         double *__restrict__ y, EpiArgs a) {
  __shared__ BinomLds B;
  __shared__ double red[2][kBlock / kWave];
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const int per = (int)((t.nwg + 7) / 8);
  const int w = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (w >= t.nwg) return;  // uniform over the workgroup
  if (!t.momentum) load_binom(B, t.binom);
  const int tid = threadIdx.x;
  const int g = t.wg_row[w];
  const int c = t.wg_c0[w] + tid;
  const int nc = t.row_ncols[g];
  const int64_t D = t.D, base = t.row_base[g];
  const EpiK<double> bp = epi_consts<double, E>(a);
  Partial2 accr, acci;
  if (c < nc) {
    const int h = t.row_dg0[g] + c;
    const uint32_t d = t.dn_bits[h];
    const int64_t i = base + c;
    const double xr = x[i], xi = x[D + i];
    double yr = 0.0, yi = 0.0;
    if constexpr (STRIP) {
      yr = y[i];
      yi = y[D + i];
    }
    cfma(t.diag_up[2 * g] + t.diag_dn[2 * h], t.diag_up[2 * g + 1] + t.diag_dn[2 * h + 1], xr,
         xi, yr, yi);
    if constexpr (!STRIP) {
      for (int64_t e = t.uo_ptr[g]; e < t.uo_ptr[g + 1]; ++e) {
        const int64_t o = t.uo_base[e] + c;
        cfma(t.uo_amp[2 * e], t.uo_amp[2 * e + 1], x[o], x[D + o], yr, yi);
      }
    }
    if (t.do_pack) {  // column-major packed lists: one coalesced word an entry
      const int n = t.do_cnt[h];
#pragma unroll 4  // four entries' loads in flight (9.56 -> 9.47 ms on the 4 x 4 Hubbard expression)
      for (int k = 0; k < n; ++k) {
        const uint32_t w2 = t.do_pack[(int64_t)k * t.n_dstr + h];
        const int64_t o = base + (w2 >> 12);
        const uint32_t id = w2 & 0xFFFu;
        cfma(t.do_amps[2 * id], t.do_amps[2 * id + 1], x[o], x[D + o], yr, yi);
      }
    } else if (t.do_ell_col) {  // column-major lists: coalesced over the lanes
      const int n = t.do_cnt[h];
      for (int k = 0; k < n; ++k) {
        const int64_t q = (int64_t)k * t.n_dstr + h;
        const int64_t o = base + t.do_ell_col[q];
        cfma(t.do_ell_amp[2 * q], t.do_ell_amp[2 * q + 1], x[o], x[D + o], yr, yi);
      }
    } else {
      for (int64_t e = t.do_ptr[h]; e < t.do_ptr[h + 1]; ++e) {
        const int64_t o = base + t.do_col[e];
        cfma(t.do_amp[2 * e], t.do_amp[2 * e + 1], x[o], x[D + o], yr, yi);
      }
    }
    double mr = 0.0, mi = 0.0;  // mixed terms that map the element onto itself
    for (int64_t e = t.mx_ptr[g]; e < t.mx_ptr[g + 1]; ++e) {
      const int k = t.mx_grp[e];
      const int64_t b = t.mx_base[e];
      const int grp = t.mx_dgrp[e];
      const double su = t.mx_sign[e];
      const bool same_row = b == base;
      for (int l = t.dp_ptr[k]; l < t.dp_ptr[k + 1]; ++l) {
        if ((d & t.dp_touched[l]) != t.dp_out[l]) continue;
        const uint32_t fl = t.dp_flip[l];
        const double sg = (popc(d & t.dp_parity[l]) & 1) ? -su : su;
        if (fl == 0 && same_row) {
          mr += sg * t.dp_c[2 * l];
          mi += sg * t.dp_c[2 * l + 1];
          continue;
        }
        const uint32_t d2 = d ^ fl;
        int64_t col;
        if (t.momentum) {
          if (t.dn_mom[d2] != grp) continue;
          col = t.dn_loc[d2];
        } else {
          col = fl ? colex_rank_dev(B, d2) : c;  // an unchanged down string keeps its column
        }
        const int64_t o = b + col;
        cfma(sg * t.dp_c[2 * l], sg * t.dp_c[2 * l + 1], x[o], x[D + o], yr, yi);
      }
    }
    cfma(mr, mi, xr, xi, yr, yi);
    epilogue<double, E>(yr, xr, i, y, a, bp, accr);
    epilogue<double, E>(yi, xi, D + i, y, a, bp, acci);
  }
  if constexpr (E == Epi::Lanczos) {
    const double sr = wg_sum(accr.s0, red[0]);
    const double si = wg_sum(acci.s0, red[1]);
    if (tid == 0) {
      a.partial[w] = sr;
      a.partial[t.nwg + w] = si;
    }
  }
}

// reference index of every device position of a real-space sector
__global__ __launch_bounds__(kBlock) void ref_of_pos_kernel(PermTables p, int64_t *__restrict__ out) {
  __shared__ BinomLds B;
  load_binom(B, p.binom);
  const int64_t D = p.d_up * p.d_dn;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < D; e += (int64_t)gridDim.x * kBlock) {
    const int64_t u = e / p.d_dn, d = e - u * p.d_dn;
    out[e] = ref_rank(B, p.L, p.n_up, p.n_dn, p.s_up[u], p.s_dn[d]);
  }
}

bool force_generic() {
  const char *e = std::getenv("MANYBODY_AMD_HV_PATH");
  return e && std::strcmp(e, "generic") == 0;
}

template <typename T>
bool use_two_pass(const HvTables &t) {
  // the strip kernel addresses a plane with 32-bit byte offsets
  const bool plane_fits = t.d_up * t.d_dn * (int64_t)sizeof(T) < ((int64_t)1 << 32);
  return t.ell_ok && t.max_up_row <= 64 && (t.d_dn + 1) * (int64_t)sizeof(T) <= kMaxRowLds &&
         plane_fits &&
         !force_generic();
}

int row_block(int64_t Ddn) {
  if (Ddn >= 8192) return 1024;
  if (Ddn >= 2048) return 512;
  if (Ddn >= 256) return 256;
  return (int)((Ddn + kWave - 1) / kWave) * kWave;
}

bool force_compact_row() {
  const char *e = std::getenv("MANYBODY_AMD_ROW_KERNEL");
  return e && std::strcmp(e, "compact") == 0;
}

// Compute units of the current device (cached per device).
int device_cus() {
  static std::atomic<int> cus[64];  // zero-initialised; host threads may race to fill
  int dev = 0;
  MB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  int v = cus[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    int n = 0;
    MB_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    v = n > 0 ? n : 1;
    cus[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// LDS of the sorted row kernel: the row with its zero slot, then half a row.
template <typename T>
size_t sorted_row_lds(const HvTables &t) {
  return (size_t)(t.d_dn + 1 + kSortedHalf) * sizeof(T);
}

// Planes per workgroup of the multi-plane row kernel: the most of 8, 4, 2 up
// to the cap that divides P and keeps the interleaved row within kMpRowLds;
// 1 = the one-plane kernels.  The default cap is 8 bytes per string (two FP32
// planes): measured on the L = 14, 7 + 7 ring, the FP32 Chebyshev step over 16
// vectors takes 4.03 ms with one plane per workgroup, 3.22 ms with two and
// 3.73 ms with four (fewer workgroups fit a CU); FP64 gains nothing (6.08 ms
// with one plane, 6.16 ms with two).  MANYBODY_AMD_ROW_PLANES=n sets the cap
// (up to 32 bytes per string; 1 switches the multi-plane kernel off).
#ifndef MB_MP_ROW_LDS
#define MB_MP_ROW_LDS (64 * 1024)
#endif
constexpr size_t kMpRowLds = MB_MP_ROW_LDS;

template <typename T>
int row_planes(const HvTables &t, int P) {
  int cap = 8 / (int)sizeof(T);
  if (const char *e = std::getenv("MANYBODY_AMD_ROW_PLANES"))
    cap = std::max(1, std::min(32 / (int)sizeof(T), std::atoi(e)));
  for (int np = cap; np >= 2; np /= 2)
    if (P % np == 0 && (size_t)(t.d_dn + 1) * np * sizeof(T) <= kMpRowLds) return np;
  return 1;
}

template <typename T, Epi E, int NP>
void launch_mp(dim3 g, dim3 b, size_t lds, hipStream_t s, const HvTables &t, const T *x, T *y,
               const EpiArgs &a) {
  if constexpr (NP * sizeof(T) <= 32) {
    hipLaunchKernelGGL((row_kernel_mp<T, E, NP>), g, b, lds, s, t, x, y, a);
    MB_HIP(hipGetLastError());
  } else {
    throw std::logic_error("multi-plane row kernel: pack wider than 32 bytes");
  }
}

// The sorted row kernel: slots of at most kSortedRounds rounds, 2 halves of
// kSortedHalf strings; for rows below 8192 strings the per-row kernel, several
// workgroups per CU, is faster.
template <typename T>
bool use_sorted_row(const HvTables &t) {
  return t.srt_ok && t.d_dn >= 8192 && t.srt_np <= (int64_t)kSortedRound * kSortedRounds &&
         t.d_dn <= 2 * (int64_t)kSortedHalf && sorted_row_lds<T>(t) <= (size_t)kMaxRowLds &&
         !force_compact_row();
}

template <typename T, Epi E>
void launch_row(const HvTables &t, int P, const T *x, T *y, const EpiArgs &a, hipStream_t s) {
  const size_t lds = (size_t)(t.d_dn + 1) * sizeof(T);  // + the zero slot of the compact table
  static std::atomic<bool> attr_set{false};  // once per template instance (idempotent)
  const void *kc = reinterpret_cast<const void *>(&row_kernel_compact<T, E>);
  const void *ke = reinterpret_cast<const void *>(&row_kernel_ell<T, E>);
  const void *ks = reinterpret_cast<const void *>(&row_kernel_sorted<T, E>);
  if (!attr_set) {
    MB_HIP(hipFuncSetAttribute(kc, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRowLds));
    MB_HIP(hipFuncSetAttribute(ke, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRowLds));
    MB_HIP(hipFuncSetAttribute(ks, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRowLds));
    attr_set = true;
  }
  const int64_t nrows = (int64_t)P * t.d_up;
  if (use_sorted_row<T>(t)) {
    const int64_t cus = device_cus();
    const dim3 g((unsigned)(nrows < cus ? nrows : cus)), b(kSortedRound);
    hipLaunchKernelGGL((row_kernel_sorted<T, E>), g, b, sorted_row_lds<T>(t), s, t, x, y, a, nrows);
    MB_HIP(hipGetLastError());
    return;
  }
  const dim3 b(row_block(t.d_dn));
  const int np = t.comp_ok ? row_planes<T>(t, P) : 1;
  if (np > 1) {
    const dim3 g((unsigned)(nrows / np));
    const size_t lm = (size_t)(t.d_dn + 1) * np * sizeof(T);
    switch (np) {
      case 2: launch_mp<T, E, 2>(g, b, lm, s, t, x, y, a); break;
      case 4: launch_mp<T, E, 4>(g, b, lm, s, t, x, y, a); break;
      default: launch_mp<T, E, 8>(g, b, lm, s, t, x, y, a); break;
    }
    return;
  }
  const dim3 g((unsigned)nrows);
  if (t.comp_ok)
    hipLaunchKernelGGL((row_kernel_compact<T, E>), g, b, lds, s, t, x, y, a);
  else
    hipLaunchKernelGGL((row_kernel_ell<T, E>), g, b, lds, s, t, x, y, a);
  MB_HIP(hipGetLastError());
}

template <typename T, int VEC>
void launch_strip(const HvTables &t, int P, const T *x, T *y, const int *flags, hipStream_t s) {
  constexpr int W = kStripLanes * VEC, RC = 32;
  const int nJ = (int)((t.d_dn + W - 1) / W);
  const int nC = (int)((t.d_up + RC - 1) / RC);
  const int64_t S8 = ((int64_t)P * nJ + 7) / 8;
  const int64_t grid = 8 * (int64_t)nC * S8;
  if (grid >= (1ll << 31)) throw std::invalid_argument("sector too large for the strip kernel");
  if (t.max_up_row <= 32)
    hipLaunchKernelGGL((up_strip_kernel<T, VEC, 32>), dim3((unsigned)grid), dim3(kBlock), 0, s, t, P,
                       x, y, nJ, nC, flags);
  else
    hipLaunchKernelGGL((up_strip_kernel<T, VEC, 64>), dim3((unsigned)grid), dim3(kBlock), 0, s, t, P,
                       x, y, nJ, nC, flags);
  MB_HIP(hipGetLastError());
}

// The block-staged up kernel runs where its tables exist and the strip rows
// are whole 16-byte packs; MANYBODY_AMD_UP_KERNEL=strip keeps the per-row
// strip kernel (the tests compare the two).
template <typename T>
bool use_up_block(const HvTables &t) {
  if (!t.ub_ok || !t.up_c16_ok || t.d_dn % (int64_t)(16 / sizeof(T)) != 0) return false;
  const char *e = std::getenv("MANYBODY_AMD_UP_KERNEL");
  return !(e && std::strcmp(e, "strip") == 0);
}

template <typename T>
void launch_up_block(const HvTables &t, int P, const T *x, T *y, const int *flags, hipStream_t s) {
  constexpr int W = kUbLanes * (16 / (int)sizeof(T));
  const int nJ = (int)((t.d_dn + W - 1) / W);
  const int64_t S8 = ((int64_t)P * nJ + 7) / 8;
  const int64_t grid = 8 * (int64_t)t.ub_nb * S8;
  if (grid >= (1ll << 31)) throw std::invalid_argument("sector too large for the up kernel");
  const size_t lds = (size_t)t.ub_rows * (kUbLanes * 16 + 4) + (size_t)t.ub_ent * 2;
  if (lds > (size_t)kUpBlockLds) throw std::logic_error("up kernel: block tables exceed the LDS budget");
  hipLaunchKernelGGL((up_block_kernel<T>), dim3((unsigned)grid), dim3(kUbThreads), lds, s, t, P, x, y,
                     nJ, flags, nullptr, EpiArgs());
  MB_HIP(hipGetLastError());
}

// Operator algebra: h = alpha z1 + beta z2 + gamma x per complex element
// (terms with a null operand left out), then the epilogue on both planes; one
// partial sum per workgroup (fixed-order reductions).
template <typename T, Epi E>
__global__ __launch_bounds__(kBlock) void combine_kernel(int64_t D, const T *__restrict__ z1,
                                                         const T *__restrict__ z2,
                                                         const T *__restrict__ x,
                                                         T *__restrict__ y, CombineCoef c,
                                                         EpiArgs a) {
  __shared__ double red[2][1024 / kWave];
  if constexpr (E == Epi::Lanczos) {
    if (a.flags[0]) return;
  }
  const EpiK<T> bp = epi_consts<T, E>(a);
  const T ar = (T)c.ar, ai = (T)c.ai, br = (T)c.br, bi = (T)c.bi, gr = (T)c.gr, gi = (T)c.gi;
  Partial2 acc;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < D;
       i += (int64_t)gridDim.x * kBlock) {
    const T xr = x[i], xi = x[D + i];
    T hr = T(0), hi = T(0);
    if (z1) {
      const T zr = z1[i], zi = z1[D + i];
      hr = ar * zr - ai * zi;
      hi = ar * zi + ai * zr;
    }
    if (z2) {
      const T zr = z2[i], zi = z2[D + i];
      hr += br * zr - bi * zi;
      hi += br * zi + bi * zr;
    }
    if (c.with_x) {
      hr += gr * xr - gi * xi;
      hi += gr * xi + gi * xr;
    }
    epilogue<T, E>(hr, xr, i, y, a, bp, acc);
    epilogue<T, E>(hi, xi, D + i, y, a, bp, acc);
  }
  row_reduce<T, E>(acc, a, blockIdx.x, red);
}

}  // namespace

// ------------------------------------------------------------ launchers

int64_t combine_slots(int64_t D) { return (int64_t)grid_for((D + 1) / 2, 512); }

template <typename T>
int64_t launch_combine(int64_t D, const T *z1, const T *z2, const T *x, T *y, const CombineCoef &c,
                       Epi epi, const EpiArgs &a, hipStream_t s) {
  const int64_t slots = combine_slots(D);
  const dim3 g((unsigned)(2 * slots));
  switch (epi) {
    case Epi::Plain:
      hipLaunchKernelGGL((combine_kernel<T, Epi::Plain>), g, dim3(kBlock), 0, s, D, z1, z2, x, y, c, a);
      break;
    case Epi::Lanczos:
      hipLaunchKernelGGL((combine_kernel<T, Epi::Lanczos>), g, dim3(kBlock), 0, s, D, z1, z2, x, y, c, a);
      break;
    case Epi::Recon:
      hipLaunchKernelGGL((combine_kernel<T, Epi::Recon>), g, dim3(kBlock), 0, s, D, z1, z2, x, y, c, a);
      break;
    case Epi::Cheb:
      hipLaunchKernelGGL((combine_kernel<T, Epi::Cheb>), g, dim3(kBlock), 0, s, D, z1, z2, x, y, c, a);
      break;
  }
  MB_HIP(hipGetLastError());
  return slots;  // partial slots per plane: the launch wrote 2 slots
}
template int64_t launch_combine<double>(int64_t, const double *, const double *, const double *,
                                        double *, const CombineCoef &, Epi, const EpiArgs &,
                                        hipStream_t);
template int64_t launch_combine<float>(int64_t, const float *, const float *, const float *,
                                       float *, const CombineCoef &, Epi, const EpiArgs &,
                                       hipStream_t);

int64_t hv_partials(const HvTables &t, int P) {
  const int64_t two = (int64_t)P * t.d_up, gen = (int64_t)P * kGenericGrid;
  return two > gen ? two : gen;
}

// Planes per workgroup of the one-pass kernel, or 0 for the two passes.  The
// one pass needs the compact down table, at most kFusedMaxHops up hops per row,
// a pack of 8 to 32 bytes per string, and a plane group (the planes one
// workgroup covers) of at most kFusedGroupBytes, so that the up-source rows of
// the group being worked on stay in the 256 MB Infinity Cache.  Two planes per
// workgroup (re and im of a vector) where P allows: measured on the L = 14,
// 7 + 7 ring, one Chebyshev step over 16 vectors takes 2.62 ms in FP32 with
// two planes (2.72 with four) and 4.50 ms in FP64 with two (4.83 with one),
// against 3.24 and 6.11 ms for the two passes.  MANYBODY_AMD_ROW_PLANES sets the
// planes as for the two-pass row kernel; MANYBODY_AMD_FUSED=0 keeps the two
// passes, =1 takes the one pass wherever the tables allow it, whatever the group
// size.
constexpr int64_t kFusedGroupBytes = int64_t(200) << 20;

template <typename T>
int fused_planes(const HvTables &t, int P) {
  const char *e = std::getenv("MANYBODY_AMD_FUSED");
  if (e && std::strcmp(e, "0") == 0) return 0;
  const bool force = e && std::strcmp(e, "1") == 0;
  if (!t.comp_ok || t.max_up_row > kFusedMaxHops) return 0;
  if (t.d_up * t.d_dn * (int64_t)sizeof(T) > (int64_t)UINT32_MAX) return 0;  // 32-bit row offsets
  int np = row_planes<T>(t, P);
  if (!std::getenv("MANYBODY_AMD_ROW_PLANES") && P % 2 == 0 &&
      (size_t)(t.d_dn + 1) * 2 * sizeof(T) <= (size_t)kMpRowLds)
    np = 2;
  if (np * sizeof(T) < 8) return 0;  // packs of 8, 16 or 32 bytes
  if ((size_t)(t.d_dn + 1) * np * sizeof(T) > (size_t)kMpRowLds) return 0;
  if (!force && (int64_t)np * t.d_up * t.d_dn * (int64_t)sizeof(T) > kFusedGroupBytes) return 0;
  return np;
}

template <typename T, Epi E, int NP>
void launch_fused_np(const HvTables &t, int P, const T *x, T *y, const EpiArgs &a, hipStream_t s) {
  // 1024 threads from 2048 strings on (L = 14 ring, FP64: 4.17 ms per Chebyshev
  // step over 16 vectors against 4.56 with 512; FP32 the same either way)
  const dim3 g((unsigned)((int64_t)P / NP * t.d_up)), b(t.d_dn >= 2048 ? 1024 : row_block(t.d_dn));
  const size_t lds = (size_t)(t.d_dn + 1) * NP * sizeof(T);
  hipLaunchKernelGGL((row_kernel_fused<T, E, NP>), g, b, lds, s, t, x, y, a);
  MB_HIP(hipGetLastError());
}

template <typename T, Epi E>
void launch_fused(const HvTables &t, int P, int np, const T *x, T *y, const EpiArgs &a, hipStream_t s) {
  switch (np) {
    case 1:
      if constexpr (sizeof(T) == 8) {
        launch_fused_np<T, E, 1>(t, P, x, y, a, s);
        break;
      }
      throw std::logic_error("one-pass H.v: pack narrower than 8 bytes");
    case 2: launch_fused_np<T, E, 2>(t, P, x, y, a, s); break;
    case 4: launch_fused_np<T, E, 4>(t, P, x, y, a, s); break;
    default:
      if constexpr (sizeof(T) == 4) {
        launch_fused_np<T, E, 8>(t, P, x, y, a, s);
        break;
      }
      throw std::logic_error("one-pass H.v: pack wider than 32 bytes");
  }
}

template <typename T>
int64_t launch_hv(const HvTables &t, int P, const T *x, T *y, Epi epi, const EpiArgs &a,
                  hipStream_t s) {
  if (use_two_pass<T>(t) && (int64_t)P * t.d_up < (1ll << 31)) {
    if (const int np = fused_planes<T>(t, P)) {
      switch (epi) {
        case Epi::Plain: launch_fused<T, Epi::Plain>(t, P, np, x, y, a, s); break;
        case Epi::Lanczos: launch_fused<T, Epi::Lanczos>(t, P, np, x, y, a, s); break;
        case Epi::Recon: launch_fused<T, Epi::Recon>(t, P, np, x, y, a, s); break;
        case Epi::Cheb: launch_fused<T, Epi::Cheb>(t, P, np, x, y, a, s); break;
      }
      return t.d_up;  // partial slots per plane
    }
  }
  if (use_two_pass<T>(t)) {
    if ((int64_t)P * t.d_up >= (1ll << 31))
      throw std::invalid_argument("sector too large for the two-pass H.v");
    const int *flags = epi == Epi::Lanczos ? a.flags : nullptr;
    constexpr int V = 16 / sizeof(T);
    if (use_up_block<T>(t))
      launch_up_block<T>(t, P, x, y, flags, s);
    else if (t.d_dn % V == 0)
      launch_strip<T, V>(t, P, x, y, flags, s);
    else
      launch_strip<T, 1>(t, P, x, y, flags, s);
    switch (epi) {
      case Epi::Plain: launch_row<T, Epi::Plain>(t, P, x, y, a, s); break;
      case Epi::Lanczos: launch_row<T, Epi::Lanczos>(t, P, x, y, a, s); break;
      case Epi::Recon: launch_row<T, Epi::Recon>(t, P, x, y, a, s); break;
      case Epi::Cheb: launch_row<T, Epi::Cheb>(t, P, x, y, a, s); break;
    }
    return t.d_up;  // partial slots per plane
  }
  const dim3 g(grid_for(t.d_up * t.d_dn, kGenericGrid), (unsigned)P);
  switch (epi) {
    case Epi::Plain: hipLaunchKernelGGL((generic_kernel<T, Epi::Plain>), g, dim3(kBlock), 0, s, t, x, y, a); break;
    case Epi::Lanczos: hipLaunchKernelGGL((generic_kernel<T, Epi::Lanczos>), g, dim3(kBlock), 0, s, t, x, y, a); break;
    case Epi::Recon: hipLaunchKernelGGL((generic_kernel<T, Epi::Recon>), g, dim3(kBlock), 0, s, t, x, y, a); break;
    case Epi::Cheb: hipLaunchKernelGGL((generic_kernel<T, Epi::Cheb>), g, dim3(kBlock), 0, s, t, x, y, a); break;
  }
  MB_HIP(hipGetLastError());
  return g.x;
}

// The row-first Lanczos step (see mb_internal.hpp): FP64 two-pass sectors
// whose up hops run in the block-staged kernel and whose rows run in the
// sorted row kernel.  MANYBODY_AMD_LANCZOS_FOLD=0 keeps the up-then-row H.v
// with a separate axpy pass.
bool lanczos_fold_ok(const HvTables &t) {
  const char *e = std::getenv("MANYBODY_AMD_LANCZOS_FOLD");
  if (e && std::strcmp(e, "0") == 0) return false;
  return use_two_pass<double>(t) && 2 * t.d_up < (1ll << 31) && fused_planes<double>(t, 2) == 0 &&
         use_up_block<double>(t) && use_sorted_row<double>(t);
}

int64_t up_block_grid(const HvTables &t, int P, int *nJ_out) {
  constexpr int W = kUbLanes * 2;  // FP64: 16 lanes x 2 doubles
  const int nJ = (int)((t.d_dn + W - 1) / W);
  if (nJ_out) *nJ_out = nJ;
  return 8 * (int64_t)t.ub_nb * (((int64_t)P * nJ + 7) / 8);
}

int64_t lanczos_fold_partials(const HvTables &t) {
  return kRedBlocks + std::max<int64_t>(2 * t.d_up, up_block_grid(t, 2, nullptr));
}

void launch_row_down(const HvTables &t, const double *x, double *z, const EpiArgs &a, hipStream_t s) {
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    MB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&row_kernel_sorted<double, Epi::Down>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRowLds));
    attr_set = true;
  }
  const int64_t nrows = 2 * t.d_up, cus = device_cus();
  const dim3 g((unsigned)(nrows < cus ? nrows : cus)), b(kSortedRound);
  hipLaunchKernelGGL((row_kernel_sorted<double, Epi::Down>), g, b, sorted_row_lds<double>(t), s, t, x, z,
                     a, nrows);
  MB_HIP(hipGetLastError());
}

int64_t launch_up_lanczos(const HvTables &t, const double *x, const double *z, double *w,
                          const EpiArgs &a, hipStream_t s) {
  int nJ = 0;
  const int64_t grid = up_block_grid(t, 2, &nJ);
  if (grid >= (1ll << 31)) throw std::invalid_argument("sector too large for the up kernel");
  const size_t lds = (size_t)t.ub_rows * (kUbLanes * 16 + 4) + (size_t)t.ub_ent * 2;
  // one partial sum per workgroup after kRedBlocks slots, then summed in
  // chunks into the first kRedBlocks slots (one workgroup would take ~0.3 ms
  // over the ~180 K partial sums of the 4 x 4, 7 + 7 sector)
  EpiArgs ua = a;
  ua.partial = a.partial + kRedBlocks;
  hipLaunchKernelGGL((up_block_kernel<double, kUbLanes, kUbThreads, true>), dim3((unsigned)grid),
                     dim3(kUbThreads), lds, s, t, 2, x, w, nJ, a.flags, z, ua);
  MB_HIP(hipGetLastError());
  hipLaunchKernelGGL(chunk_sum_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, ua.partial, grid, a.partial);
  MB_HIP(hipGetLastError());
  return kRedBlocks;
}

template <Epi E, int CPT>
void launch_mom_cpt(const MomTables &t, const double *x, double *y, const EpiArgs &a, hipStream_t s) {
  const size_t lds = (size_t)2 * t.max_cols * sizeof(double);
  const unsigned grid = 8u * (unsigned)((t.nrows + 7) / 8);
  hipLaunchKernelGGL((mom_hv_kernel<E, CPT>), dim3(grid), dim3(kBlock), lds, s, t, x, y, a);
}

// Workgroups per XCD of the persistent momentum kernel (one per CU), or 0 when
// the sector does not fit it: more than kLockCols columns per thread of a group,
// rows beyond kLockRows per workgroup, more than kLockQ momenta, source lists
// wider than kLockSrc, or more than kMomLockMaxW workgroups per XCD (the
// partial slots the operator provides).
// Sectors below 2^20 states keep the per-row kernel unless
// MANYBODY_AMD_MOM_KERNEL=lock; MANYBODY_AMD_MOM_KERNEL=rows always keeps it.
int mom_lock_wgs(const MomTables &t) {
  const char *e = std::getenv("MANYBODY_AMD_MOM_KERNEL");
  if (e && std::strcmp(e, "rows") == 0) return 0;
  const bool force = e && std::strcmp(e, "lock") == 0;
  const int W = device_cus() / 8;
  if (W < 1 || W > kMomLockMaxW || t.N > kLockQ || (t.D < (1 << 20) && !force)) return 0;
  if (t.max_cols > kLockCols * kLockGroup || t.max_cols > kLockS) return 0;
  if (t.mu != kLockSrc || t.md != kLockSrc) return 0;
  if ((t.max_rows + W - 1) / W > kLockRows) return 0;
  return W;
}

template <Epi E>
void launch_mom_lock(const MomTables &t, int W, const double *x, double *y, const EpiArgs &a,
                     hipStream_t s) {
  const size_t lds = lock_lds_bytes();
  static std::atomic<bool> attr_set{false};  // once per template instance (idempotent)
  const void *k = reinterpret_cast<const void *>(&mom_hv_lock_kernel<E>);
  if (!attr_set) {
    MB_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  hipLaunchKernelGGL((mom_hv_lock_kernel<E>), dim3(8 * W), dim3(kLockBlock), lds, s, t, x, y, a);
}

template <Epi E>
void launch_mom_e(const MomTables &t, const double *x, double *y, const EpiArgs &a, hipStream_t s) {
  if (const int W = mom_lock_wgs(t)) {
    launch_mom_lock<E>(t, W, x, y, a, s);
    return;
  }
  const int cpt = (t.max_cols + kBlock - 1) / kBlock;
  if (cpt <= 1)
    launch_mom_cpt<E, 1>(t, x, y, a, s);
  else if (cpt <= 2)
    launch_mom_cpt<E, 2>(t, x, y, a, s);
  else if (cpt <= 4)
    launch_mom_cpt<E, 4>(t, x, y, a, s);
  else if (cpt <= 8)
    launch_mom_cpt<E, 8>(t, x, y, a, s);
  else if (cpt <= 16)
    launch_mom_cpt<E, 16>(t, x, y, a, s);
  else
    throw std::invalid_argument("momentum group too large for the H.v kernel (more than 4096 strings)");
}

int64_t launch_mom_hv(const MomTables &t, const double *x, double *y, Epi epi, const EpiArgs &a,
                      hipStream_t s) {
  if (t.nrows >= (1ll << 31)) throw std::invalid_argument("sector too large");
  switch (epi) {
    case Epi::Plain: launch_mom_e<Epi::Plain>(t, x, y, a, s); break;
    case Epi::Lanczos: launch_mom_e<Epi::Lanczos>(t, x, y, a, s); break;
    case Epi::Recon: launch_mom_e<Epi::Recon>(t, x, y, a, s); break;
    case Epi::Cheb: throw std::invalid_argument("KPM is not available for momentum-sector operators");
  }
  MB_HIP(hipGetLastError());
  const int W = mom_lock_wgs(t);
  return W ? (int64_t)t.N * W : t.nrows;  // partial slots per plane
}

void launch_transfer(const TransferTables &t, const double *x_src, double *y_tgt, hipStream_t s) {
  if (t.rows >= (1ll << 31)) throw std::invalid_argument("sector too large");
  const size_t lds = (size_t)2 * t.max_cols * sizeof(double);
  hipLaunchKernelGGL(transfer_kernel, dim3((unsigned)t.rows), dim3(kBlock), lds, s, t, x_src, y_tgt);
  MB_HIP(hipGetLastError());
}

template <Epi E>
void launch_expr_e(const ExprTables &t, const double *x, double *y, const EpiArgs &a, hipStream_t s) {
  // the strip pass: real space, up-only terms present, lists that fit the LDS
  // staging, planes addressable with 32-bit byte offsets
  const bool strip = !t.momentum && t.max_uo > 0 && t.max_uo <= 64 &&
                     t.D * (int64_t)sizeof(double) < ((int64_t)1 << 32) && !force_generic();
  const int per = (int)((t.nwg + 7) / 8);
  const dim3 g((unsigned)(8 * per));
  if (strip) {
    const int nJ = (int)((t.d_dn + 15) / 16), nC = (int)((t.d_up + 31) / 32);
    const int64_t grid = (int64_t)8 * ((nJ + 7) / 8) * nC;
    if (grid >= ((int64_t)1 << 31)) throw std::invalid_argument("sector too large");
    const int *flags = (E == Epi::Lanczos) ? a.flags : nullptr;
    if (t.max_uo <= 32)
      hipLaunchKernelGGL((expr_strip_kernel<32>), dim3((unsigned)grid), dim3(kBlock), 0, s, t, x, y,
                         nJ, nC, flags);
    else
      hipLaunchKernelGGL((expr_strip_kernel<64>), dim3((unsigned)grid), dim3(kBlock), 0, s, t, x, y,
                         nJ, nC, flags);
    hipLaunchKernelGGL((expr_hv_kernel<E, true>), g, dim3(kBlock), 0, s, t, x, y, a);
  } else {
    hipLaunchKernelGGL((expr_hv_kernel<E, false>), g, dim3(kBlock), 0, s, t, x, y, a);
  }
}

int64_t launch_expr_hv(const ExprTables &t, const double *x, double *y, Epi epi, const EpiArgs &a,
                       hipStream_t s) {
  if (t.nwg == 0) return 0;
  switch (epi) {
    case Epi::Plain: launch_expr_e<Epi::Plain>(t, x, y, a, s); break;
    case Epi::Lanczos: launch_expr_e<Epi::Lanczos>(t, x, y, a, s); break;
    case Epi::Recon: launch_expr_e<Epi::Recon>(t, x, y, a, s); break;
    case Epi::Cheb: throw std::invalid_argument("KPM is not available for expression operators");
  }
  MB_HIP(hipGetLastError());
  return t.nwg;  // partial slots per plane
}

void launch_ref_of_pos(const PermTables &p, int64_t *out, hipStream_t s) {
  hipLaunchKernelGGL(ref_of_pos_kernel, dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock), 0,
                     s, p, out);
  MB_HIP(hipGetLastError());
}

void launch_mom_perm(const int64_t *map, int64_t D, const double *src, double *dst, bool to_device,
                     hipStream_t s) {
  if (to_device)
    hipLaunchKernelGGL((mom_perm_kernel<true>), dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s, map, D,
                       src, dst);
  else
    hipLaunchKernelGGL((mom_perm_kernel<false>), dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s, map, D,
                       src, dst);
  MB_HIP(hipGetLastError());
}

void launch_mom_fill_gaussian(const int64_t *ref_of_pos, int64_t D, uint64_t seed, double *x,
                              hipStream_t s) {
  hipLaunchKernelGGL(mom_fill_gaussian_kernel, dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s,
                     ref_of_pos, D, seed, x);
  MB_HIP(hipGetLastError());
}

template int64_t launch_hv<double>(const HvTables &, int, const double *, double *, Epi,
                                   const EpiArgs &, hipStream_t);
template int64_t launch_hv<float>(const HvTables &, int, const float *, float *, Epi,
                                  const EpiArgs &, hipStream_t);

void launch_gather_c128(const PermTables &p, const void *x_ref, void *x_dev, hipStream_t s) {
  hipLaunchKernelGGL((perm_kernel<true>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock), 0,
                     s, p, (const double *)x_ref, (double *)x_dev);
  MB_HIP(hipGetLastError());
}

void launch_scatter_c128(const PermTables &p, const void *x_dev, void *x_ref, hipStream_t s) {
  hipLaunchKernelGGL((perm_kernel<false>), dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock),
                     0, s, p, (const double *)x_dev, (double *)x_ref);
  MB_HIP(hipGetLastError());
}

void launch_fill_gaussian_c128(const PermTables &p, uint64_t seed, void *x, hipStream_t s) {
  hipLaunchKernelGGL(fill_gaussian_kernel, dim3(grid_for(p.d_up * p.d_dn, 1 << 16)), dim3(kBlock),
                     0, s, p, seed, (double *)x);
  MB_HIP(hipGetLastError());
}

void launch_fill_phase(const PermTables &p, uint64_t seed, int nvec, bool fp32, void *x,
                       hipStream_t s) {
  const dim3 g(grid_for(p.d_up * p.d_dn, 1 << 16)), b(kBlock);
  if (fp32)
    hipLaunchKernelGGL((fill_phase_kernel<float>), g, b, 0, s, p, seed, nvec, (float *)x);
  else
    hipLaunchKernelGGL((fill_phase_kernel<double>), g, b, 0, s, p, seed, nvec, (double *)x);
  MB_HIP(hipGetLastError());
}

void launch_fill_phase_ref(const int64_t *ref, int64_t D, uint64_t seed, int nvec, int r,
                           bool fp32, void *x, hipStream_t s) {
  const dim3 g(grid_for(D, 1 << 16)), b(kBlock);
  if (fp32)
    hipLaunchKernelGGL((fill_phase_ref_kernel<float>), g, b, 0, s, ref, D, seed, nvec, r, (float *)x);
  else
    hipLaunchKernelGGL((fill_phase_ref_kernel<double>), g, b, 0, s, ref, D, seed, nvec, r, (double *)x);
  MB_HIP(hipGetLastError());
}

void launch_reduce_pairs(const double *partial, int64_t n, double *out0, double *out1,
                         hipStream_t s) {
  hipLaunchKernelGGL(reduce_pairs_kernel, dim3(1), dim3(kBlock), 0, s, partial, n, out0, out1);
  MB_HIP(hipGetLastError());
}

void launch_sumsq(const void *x, bool fp32, int64_t n, double *partial, double *out,
                  bool take_sqrt, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL((sumsq_partial_kernel<float>), dim3(kRedBlocks), dim3(kBlock), 0, s,
                       (const float *)x, n, partial);
  else
    hipLaunchKernelGGL((sumsq_partial_kernel<double>), dim3(kRedBlocks), dim3(kBlock), 0, s,
                       (const double *)x, n, partial);
  hipLaunchKernelGGL(final_kernel, dim3(1), dim3(kBlock), 0, s, (const double *)partial,
                     (int64_t)kRedBlocks, out, take_sqrt ? 1 : 0);
  MB_HIP(hipGetLastError());
}

void launch_cdot(const double *a, const double *b, int64_t D, double *partial, double *out,
                 hipStream_t s) {
  hipLaunchKernelGGL(cdot_partial_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, a, b, D, partial);
  hipLaunchKernelGGL(final_kernel, dim3(1), dim3(kBlock), 0, s, (const double *)partial,
                     (int64_t)kRedBlocks, out, 0);
  hipLaunchKernelGGL(final_kernel, dim3(1), dim3(kBlock), 0, s, (const double *)partial + kRedBlocks,
                     (int64_t)kRedBlocks, out + 1, 0);
  MB_HIP(hipGetLastError());
}

void launch_sumsq_partial(const double *x, int64_t n, double *partial, hipStream_t s) {
  hipLaunchKernelGGL((sumsq_partial_kernel<double>), dim3(kRedBlocks), dim3(kBlock), 0, s, x, n,
                     partial);
  MB_HIP(hipGetLastError());
}

void launch_reorth(const double *V, int m, double *w, int64_t D, double *partial, double *c,
                   const int *flags, hipStream_t s) {
  constexpr int nb = kProjBlocks;
  hipLaunchKernelGGL(proj_partial_kernel, dim3(nb, (unsigned)m), dim3(kBlock), 0, s, V,
                     (const double *)w, D, partial, flags);
  hipLaunchKernelGGL(proj_final_kernel, dim3((unsigned)m), dim3(kBlock), 0, s, (const double *)partial,
                     nb, c, flags);
  hipLaunchKernelGGL(proj_sub_kernel, dim3(grid_for(D, 4096)), dim3(kBlock), 0, s, V, m,
                     (const double *)c, w, D, flags);
  MB_HIP(hipGetLastError());
}

void launch_reorth_sub(const double *V, int m, const double *c, double *w, int64_t D,
                       const int *flags, hipStream_t s) {
  hipLaunchKernelGGL(proj_sub_kernel, dim3(grid_for(D, 4096)), dim3(kBlock), 0, s, V, m, c, w, D,
                     flags);
  MB_HIP(hipGetLastError());
}

// Clears n 32-bit words (a kernel rather than hipMemsetAsync, so a captured
// stream records no memset node).
__global__ __launch_bounds__(kBlock) void zero_words_kernel(uint32_t *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = 0u;
}

void launch_zero(void *p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (bytes % 4 != 0) throw std::invalid_argument("launch_zero: byte count not a multiple of 4");
  const int64_t n = (int64_t)(bytes / 4);
  hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(n, 1 << 14)), dim3(kBlock), 0, s,
                     static_cast<uint32_t *>(p), n);
  MB_HIP(hipGetLastError());
}

void launch_norm(const double *x, int64_t n, double *partial, double *out, hipStream_t s) {
  launch_sumsq(x, false, n, partial, out, true, s);
}

template <typename T>
void launch_scale_div(const double *x, const double *den, T *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL((scale_div_kernel<T>), dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s, x, den, y,
                     n);
  MB_HIP(hipGetLastError());
}
template void launch_scale_div<double>(const double *, const double *, double *, int64_t, hipStream_t);
template void launch_scale_div<float>(const double *, const double *, float *, int64_t, hipStream_t);

void launch_lanczos_finish_alpha(const LanczosState &st, int step, int64_t nparts, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_finish_alpha_kernel, dim3(1), dim3(kBlock), 0, s, st, step, nparts);
  MB_HIP(hipGetLastError());
}

template <typename T>
void launch_lanczos_axpy_norm(const T *v, T *w, int64_t n, const LanczosState &st, int step,
                              hipStream_t s, bool lazy) {
  hipLaunchKernelGGL((lanczos_axpy_norm_kernel<T>), dim3(kRedBlocks), dim3(kBlock), 0, s, v, w, n, st,
                     step, lazy ? 1 : 0);
  MB_HIP(hipGetLastError());
}
template void launch_lanczos_axpy_norm<double>(const double *, double *, int64_t, const LanczosState &,
                                               int, hipStream_t, bool);
template void launch_lanczos_axpy_norm<float>(const float *, float *, int64_t, const LanczosState &,
                                              int, hipStream_t, bool);

void launch_lanczos_finish_beta(const LanczosState &st, int step, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_finish_beta_kernel, dim3(1), dim3(kBlock), 0, s, st, step, (int64_t)kRedBlocks);
  MB_HIP(hipGetLastError());
}

void launch_lanczos_finish_beta_n(const LanczosState &st, int step, int64_t nparts, hipStream_t s) {
  // two levels, as for the up pass's alpha partials: one workgroup over the
  // 22,880 row sums of the 4 x 4, 7 + 7 sector took 24 us
  hipLaunchKernelGGL(chunk_sum_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, st.partial + kRedBlocks,
                     nparts, st.partial);
  MB_HIP(hipGetLastError());
  hipLaunchKernelGGL(lanczos_finish_beta_kernel, dim3(1), dim3(kBlock), 0, s, st, step,
                     (int64_t)kRedBlocks);
  MB_HIP(hipGetLastError());
}

template <typename T>
void launch_lanczos_normalize(T *w, int64_t n, const LanczosState &st, int step, hipStream_t s) {
  hipLaunchKernelGGL((lanczos_normalize_kernel<T>), dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s, w,
                     n, st, step);
  MB_HIP(hipGetLastError());
}
template void launch_lanczos_normalize<double>(double *, int64_t, const LanczosState &, int, hipStream_t);
template void launch_lanczos_normalize<float>(float *, int64_t, const LanczosState &, int, hipStream_t);

template <typename T>
void launch_recon_update(T *w, double *x, double beta, double yr, double yi, int64_t D,
                         hipStream_t s) {
  hipLaunchKernelGGL((recon_update_kernel<T>), dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s, w, x,
                     beta, yr, yi, D);
  MB_HIP(hipGetLastError());
}
template <typename T>
void launch_basis_combine(const T *V, int m, int64_t D, const double *y, double *x, hipStream_t s) {
  hipLaunchKernelGGL((basis_combine_kernel<T>), dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s, V, m,
                     D, y, x);
  MB_HIP(hipGetLastError());
}
template void launch_basis_combine<double>(const double *, int, int64_t, const double *, double *,
                                           hipStream_t);
template void launch_basis_combine<float>(const float *, int, int64_t, const double *, double *,
                                          hipStream_t);

template void launch_recon_update<double>(double *, double *, double, double, double, int64_t,
                                          hipStream_t);
template void launch_recon_update<float>(float *, double *, double, double, double, int64_t,
                                         hipStream_t);

void launch_cheb_accum(const double *y, const double *t, double *tp, double *res, int64_t n,
                       double sc, double b, bool first, double c, hipStream_t s) {
  hipLaunchKernelGGL(cheb_accum_kernel, dim3(grid_for(n, 1 << 16)), dim3(kBlock), 0, s, y, t, tp,
                     res, n, sc, b, first ? 1 : 0, c);
  MB_HIP(hipGetLastError());
}

void launch_power_step(const double *v, double *w, int64_t D, double sign, double shift,
                       double *partial, hipStream_t s) {
  hipLaunchKernelGGL(power_step_kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, v, w, D, sign, shift,
                     partial);
  MB_HIP(hipGetLastError());
}

template <typename T>
void launch_scale_complex(const T *x, double yr, double yi, double *out, int64_t D, hipStream_t s) {
  hipLaunchKernelGGL((scale_complex_kernel<T>), dim3(grid_for(D, 1 << 16)), dim3(kBlock), 0, s, x, yr,
                     yi, out, D);
  MB_HIP(hipGetLastError());
}
template void launch_scale_complex<double>(const double *, double, double, double *, int64_t,
                                           hipStream_t);
template void launch_scale_complex<float>(const float *, double, double, double *, int64_t,
                                          hipStream_t);

void launch_reduce_cheb(const double *partial, int64_t rows_per_plane, int nvec, double *out,
                        hipStream_t s) {
  hipLaunchKernelGGL(reduce_cheb_kernel, dim3(nvec), dim3(kBlock), 0, s, partial, rows_per_plane,
                     nvec, out);
  MB_HIP(hipGetLastError());
}

}  // namespace mb
