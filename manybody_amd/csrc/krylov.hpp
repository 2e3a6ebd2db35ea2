// Host drivers of the device Krylov layer (see krylov.cpp).
#pragma once

#include <complex>
#include <functional>
#include <memory>
#include <utility>
#include <vector>

#include "mb_internal.hpp"

namespace mb {

// The operator the Krylov layer runs on (the reference's LinearOperator,
// linear_operator.h:34-43): one complex H.v over device-order planes with a
// fused epilogue; returns the partial-sum slots it wrote per plane.
struct HvOp {
  virtual ~HvOp() = default;
  virtual int64_t dim() const = 0;
  virtual int64_t partials() const = 0;  // partial slots one H.v may write (both planes)
  virtual int64_t hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const = 0;
  virtual int64_t hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const = 0;
  // the real-space Hubbard tables, for the one-workgroup kernels of small
  // sectors; nullptr for other operators
  virtual const HvTables *hubbard_tables() const { return nullptr; }
};

// The real-space Hubbard operator (two-pass H.v)
struct RealHvOp final : HvOp {
  explicit RealHvOp(const Hubbard &h) : D_(h.dim()), t_(tables_of(h)) {}
  int64_t dim() const override { return D_; }
  int64_t partials() const override { return hv_partials(t_, 2); }
  int64_t hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const override {
    return launch_hv<double>(t_, 2, x, y, e, a, s);
  }
  int64_t hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const override {
    return launch_hv<float>(t_, 2, x, y, e, a, s);
  }
  const HvTables *hubbard_tables() const override { return &t_; }
  int64_t D_;
  HvTables t_;
};

// LanczosDecomposition (lanczos.h:60-68)
struct LanczosResult {
  std::vector<double> alphas, betas;
  int steps = 0;
  double b_norm = 0.0;
};

// Device workspace of one Lanczos run: three complex128 vectors and the
// device-resident scalars (alpha, beta, partial sums, flags).
struct LanczosWork {
  int64_t D;
  int k;
  size_t esz;  // bytes per vector element: 8 (FP64) or 4 (FP32 Krylov vectors)
  DevBuf v, vp, w, scal;
  LanczosState st;
  // the row-first step of real-space operators: the row pass's result and
  // its own partial-sum slots (allocated on first use)
  DevBuf z, fold_part;
  // vectors = false: scalars only (the recurrence then runs in a stored basis)
  LanczosWork(int64_t D, int k, int64_t nparts, size_t esz = sizeof(double), bool vectors = true);
};

// Whether k + 1 Krylov vectors of `bytes` should be kept in HBM: under a
// request region always (the allocation itself decides), otherwise when they
// take at most half of the free device memory.
bool keep_krylov_basis(size_t bytes);

LanczosResult lanczos_pass_one(const HvOp &op, const void *b, int k, hipStream_t s,
                               LanczosWork *work = nullptr);
void lanczos_pass_two(const HvOp &op, const void *b, const LanczosResult &dec,
                      const std::vector<std::complex<double>> &y, void *x, hipStream_t s,
                      LanczosWork *work = nullptr);
double tridiag_min_eig(int m, const double *a, const double *b);
// lanczos_pass_one with every device buffer carved from a static per-device
// scratch array (noalloc.hip): no hipMalloc / hipFree; std::invalid_argument
// when the run needs more scratch than the array holds.
LanczosResult lanczos_pass_one_noalloc(const HvOp &op, const void *b, int k, hipStream_t s);
size_t lanczos_noalloc_bytes(int64_t D, int k, int64_t nparts);
size_t lanczos_noalloc_scratch_bytes();
double find_min_eigenpair(const HvOp &op, const void *b, int k, void *x, hipStream_t s);
// Reorthogonalised variants (all Lanczos vectors kept in HBM; see krylov.cpp).
LanczosResult lanczos_pass_one_reorth(const HvOp &op, const void *b, int k, hipStream_t s,
                                      DevBuf *basis = nullptr);
double find_min_eigenpair_reorth(const HvOp &op, const void *b, int k, void *x, hipStream_t s);
std::vector<std::complex<double>> expm_first_column(const LanczosResult &dec, double t);
// fp32: Krylov vectors and H.v in FP32; the state, norms, alpha/beta
// reductions and the exponential of T_k stay FP64.
void time_evolve(const HvOp &op, const void *psi, double t, int krylov_steps, void *out,
                 hipStream_t s, bool fp32 = false);
// callback(t, state) at t0 and after every step (time_evolve_state_steps's
// callback, evolve_state.h:199-231), the stream synchronised before each call
using StepCallback = std::function<void(double, const void *)>;
void time_evolve_steps(const HvOp &op, const void *psi, double t0, double t1, double dt,
                       int krylov_steps, void *out, hipStream_t s, bool fp32 = false,
                       const StepCallback &callback = nullptr);
// FP64 exp(-i t H) psi, or exp(-t H) psi with imag (imaginary time), with an
// optional cache holding the stored Krylov basis between calls
void time_evolve_kind(const HvOp &op, const void *psi, double t, int krylov_steps, void *out,
                      hipStream_t s, DevBuf *cache, bool imag);
std::vector<std::complex<double>> expm_first_column_imag(const LanczosResult &dec, double tau);
// the same on the real-space Hubbard operator
inline LanczosResult lanczos_pass_one(const Hubbard &h, const void *b, int k, hipStream_t s,
                                      LanczosWork *work = nullptr) {
  return lanczos_pass_one(RealHvOp(h), b, k, s, work);
}
inline double find_min_eigenpair(const Hubbard &h, const void *b, int k, void *x, hipStream_t s) {
  return find_min_eigenpair(RealHvOp(h), b, k, x, s);
}
inline void time_evolve(const Hubbard &h, const void *psi, double t, int krylov_steps, void *out,
                        hipStream_t s, bool fp32 = false) {
  time_evolve(RealHvOp(h), psi, t, krylov_steps, out, s, fp32);
}
inline void time_evolve_steps(const Hubbard &h, const void *psi, double t0, double t1, double dt,
                              int krylov_steps, void *out, hipStream_t s, bool fp32 = false) {
  time_evolve_steps(RealHvOp(h), psi, t0, t1, dt, krylov_steps, out, s, fp32);
}
void kpm_dos_moments(const Hubbard &h, const Sector &sec, double a, double b, int M, int nvec,
                     uint64_t seed, bool fp32, double *mu, hipStream_t s);
// The same for any operator (FP64); ref: device position -> reference index
void kpm_dos_moments_op(const HvOp &op, const int64_t *ref, double a, double b, int M, int nvec,
                        uint64_t seed, double *mu, hipStream_t s);
void kpm_accumulate(const double *raw, int M, int nvec, int64_t D, double *mu, hipStream_t s);

// out[0] = sum x_i^2 over n values (FP64 accumulation), or its square root
// solve(op, b, k, solver) (lanczos.h:220-237) with a host solver giving the
// first-column coefficients y (size = steps taken) from (alphas, betas)
using TridiagSolver = std::function<std::vector<std::complex<double>>(
    const std::vector<double> &, const std::vector<double> &)>;
void krylov_solve(const HvOp &op, const void *b, int k, const TridiagSolver &solver, void *x,
                  hipStream_t s);
// find_max_eigenpair (lanczos.h:246-281) with an explicit start vector; x may be null
double find_max_eigenpair(const HvOp &op, const void *b, int k, void *x, hipStream_t s);
// T_next = sc (y - b T) [- T_prev] over T_prev, result += c T_next (n doubles)
void launch_cheb_accum(const double *y, const double *t, double *tp, double *res, int64_t n,
                       double sc, double b, bool first, double c, hipStream_t s);
// kpm::jackson_kernel (kpm.h:19-30), kpm::projector_moments (:46-59),
// kpm::lorentz_kernel (:32-43)
std::vector<double> jackson_kernel(int N);
std::vector<double> projector_moments(int M, double epsilon);
std::vector<double> lorentz_kernel(int N, double lambda);
// SparseKPMProjector(H, M, E_min, E_max, E_fermi, padding).apply(v) (kpm.h:184-241):
// sum_m g_m mu_m(eps) T_m(H~) v with H~ = (H - b) / a, Jackson g
void kpm_projector_apply(const HvOp &op, int M, double e_min, double e_max, double e_fermi,
                         double padding, const void *v, void *out, hipStream_t s);
// w <- sign w + shift v; partial[3 b + {0, 1, 2}] = Re sum v w, Re sum v v (no
// conjugation), sum |w|^2 per block b < kRedBlocks
void launch_power_step(const double *v, double *w, int64_t D, double sign, double shift,
                       double *partial, hipStream_t s);
// power_method(op, v, iterations) (linear_operator_utils.h:18-50) on
// sign H + shift I from seed (device order, not modified)
double power_method(const HvOp &op, const void *seed, int iterations, double sign, double shift,
                    hipStream_t s);
// estimate_bounds(op, power_iterations, spectral_padding) (:52-78): {alpha, beta}
std::pair<double, double> estimate_bounds(const HvOp &op, const void *seed, int power_iterations,
                                          double spectral_padding, hipStream_t s);
void launch_sumsq(const void *x, bool fp32, int64_t n, double *partial, double *out,
                  bool take_sqrt, hipStream_t s);

}  // namespace mb
