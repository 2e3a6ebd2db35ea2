// manybody_amd.hpp — C++ host API over the C ABI (manybody_hip.h), shaped like
// keyehzy/manybody's own API for the Hubbard H.v path so that a reference user
// can switch with few edits.  Header-only; link libmanybody_hip.so and
// libamdhip64.
//
//   reference                                       here
//   FermionBasis::with_fixed_particle_number_and_    mbamd::FermionBasis::with_fixed_...
//     spin            (src/algebra/basis.h:25-26)
//   HubbardModel{,2D,3D}(t, u, sizes...)             mbamd::HubbardModel{,2D,3D}
//     (src/algebra/fermion/model/hubbard_model.h)
//   LinearOperator<cx_vec>::apply / dimension        mbamd::HubbardOperator::apply / dimension
//     (src/numerics/linear_operator.h:34-43)
//   LanczosDecomposition, lanczos_pass_one           mbamd::LanczosDecomposition, lanczos_pass_one
//     (src/numerics/lanczos.h:60-68, 130-175)
//   EigenPair, find_min_eigenpair (:239-290)         mbamd::EigenPair, find_min_eigenpair
//   EvolutionOptions, time_evolve_state[_steps]      mbamd::EvolutionOptions, time_evolve_state[_steps]
//     (src/numerics/evolve_state.h:110-231)
//   kpm::jackson_kernel, rescaling::from_bounds      mbamd::kpm::jackson_kernel, mbamd::rescaling::from_bounds
//   CurrentOperatorFixedK (src/numerics/              mbamd::CurrentOperatorFixedK
//     current_operator_fixed_k.h:28-69)
//   imaginary_time_evolve_state (evolve_state.h:      mbamd::imaginary_time_evolve_state
//     184-197)
//   compute_optical_conductivity (src/algorithms/     mbamd::compute_optical_conductivity
//     optical_conductivity.h:6-13)
//   compute_matrix_elements(basis, FermionExpression)  mbamd::expression_operator (device
//     (src/algebra/matrix_elements.h:74-105)          operator), mbamd::compute_matrix_elements
//   save/load of arma::sp_cx_mat with cereal         mbamd::SparseMatrix::save / load
//     (src/utils/arma_cereal.h:9-30)
//
// Vectors are device buffers (DeviceVector, complex128, device product order);
// to_device_order / to_reference_order convert from and to the reference's
// basis index order.  Errors throw the reference's exception types.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "manybody_hip.h"

namespace mbamd {

inline void check(int status) {
  if (status == MB_OK) return;
  const std::string msg = mb_last_error();
  switch (status) {
    case MB_ERR_INVALID: throw std::invalid_argument(msg);
    case MB_ERR_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}

// Owning device buffer of n complex128 values.  Device-order vectors are
// planar (n real parts, then n imaginary parts); reference-order vectors (the
// input of to_device_order, the output of to_reference_order) are interleaved.
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(size_t n) : n_(n) {
    if (hipMalloc(&p_, n * sizeof(std::complex<double>) + 1) != hipSuccess)
      throw std::runtime_error("hipMalloc failed");
  }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(o.n_) {}
  DeviceVector &operator=(DeviceVector &&o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~DeviceVector() {
    if (p_) (void)hipFree(p_);
  }
  void *data() const { return p_; }
  size_t size() const { return n_; }
  // device order (planar) -> host complex values
  std::vector<std::complex<double>> to_host() const {
    std::vector<double> raw(2 * n_);
    copy_out(raw.data());
    std::vector<std::complex<double>> h(n_);
    for (size_t i = 0; i < n_; ++i) h[i] = {raw[i], raw[n_ + i]};
    return h;
  }
  // reference order (interleaved) -> host complex values
  std::vector<std::complex<double>> to_host_interleaved() const {
    std::vector<std::complex<double>> h(n_);
    copy_out(h.data());
    return h;
  }
  static DeviceVector from_host(const std::vector<std::complex<double>> &h) {
    std::vector<double> raw(2 * h.size());
    for (size_t i = 0; i < h.size(); ++i) {
      raw[i] = h[i].real();
      raw[h.size() + i] = h[i].imag();
    }
    DeviceVector d(h.size());
    if (hipMemcpy(d.p_, raw.data(), raw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
    return d;
  }
  static DeviceVector from_host_interleaved(const std::vector<std::complex<double>> &h) {
    DeviceVector d(h.size());
    if (hipMemcpy(d.p_, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
    return d;
  }

 private:
  void copy_out(void *dst) const {
    if (hipMemcpy(dst, p_, n_ * sizeof(std::complex<double>), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
  }
  void *p_ = nullptr;
  size_t n_ = 0;
};

class FermionBasis {
 public:
  static FermionBasis with_fixed_particle_number_and_spin(size_t orbitals, size_t particles,
                                                          int spin_projection) {
    mb_basis *b = nullptr;
    check(mb_basis_create((int)orbitals, (int)particles, spin_projection, &b));
    return FermionBasis(b, orbitals, particles);
  }
  // Basis::with_fixed_particle_number (basis.h:24): every spin split of N particles
  static FermionBasis with_fixed_particle_number(size_t orbitals, size_t particles) {
    mb_basis *b = nullptr;
    check(mb_fixed_n_basis_create((int)orbitals, (int)particles, &b));
    return FermionBasis(b, orbitals, particles);
  }
  // Basis::with_fixed_particle_number_spin_momentum (basis.h:27-29): total momentum
  // `momentum` on the periodic lattice `lattice_size` (the reference's Index)
  static FermionBasis with_fixed_particle_number_spin_momentum(
      size_t orbitals, size_t particles, int spin_projection, const std::vector<int> &lattice_size,
      const std::vector<int> &momentum) {
    mb_basis *b = nullptr;
    if (momentum.size() != lattice_size.size())
      throw std::invalid_argument("momentum and lattice dimensionality differ");
    check(mb_momentum_basis_create((int)orbitals, (int)particles, spin_projection,
                                   (int)lattice_size.size(), lattice_size.data(), momentum.data(),
                                   &b));
    return FermionBasis(b, orbitals, particles);
  }
  FermionBasis(FermionBasis &&o) noexcept
      : h_(std::exchange(o.h_, nullptr)), orbitals(o.orbitals), particles(o.particles) {}
  FermionBasis(const FermionBasis &) = delete;
  ~FermionBasis() { mb_basis_destroy(h_); }

  size_t size() const { return (size_t)mb_basis_size(h_); }
  // basis.set[index] as (up, down) occupation bitmasks
  std::pair<uint64_t, uint64_t> state(size_t index) const {
    uint64_t u, d;
    check(mb_basis_state(h_, (int64_t)index, &u, &d));
    return {u, d};
  }
  size_t index_of(uint64_t up, uint64_t dn) const {
    int64_t i;
    check(mb_basis_index_of(h_, up, dn, &i));
    return (size_t)i;
  }
  bool contains(uint64_t up, uint64_t dn) const {
    int64_t i;
    return mb_basis_index_of(h_, up, dn, &i) == MB_OK;
  }
  const mb_basis *handle() const { return h_; }

 private:
  FermionBasis(mb_basis *h, size_t o, size_t p) : h_(h), orbitals(o), particles(p) {}
  mb_basis *h_;

 public:
  size_t orbitals;
  size_t particles;
};

class HubbardOperator {
 public:
  using ScalarType = std::complex<double>;
  HubbardOperator(const FermionBasis &basis, const std::vector<int> &sizes, double t, double u)
      : sites_(basis.orbitals) {
    check(mb_hubbard_create(basis.handle(), (int)sizes.size(), sizes.data(), t, u, &h_));
  }
  HubbardOperator(HubbardOperator &&o) noexcept
      : h_(std::exchange(o.h_, nullptr)), sites_(o.sites_) {}
  // HubbardMomentumFactorizedFixedK(basis, lattice_size, t, U, target_momentum)
  // (hubbard_momentum_factorized_fixed_k.h:41-90) on a fixed-momentum basis
  static HubbardOperator momentum_factorized_fixed_k(const FermionBasis &basis,
                                                     const std::vector<int> &lattice_size,
                                                     double t, double u,
                                                     const std::vector<int> &target_momentum) {
    if (target_momentum.size() != lattice_size.size())
      throw std::invalid_argument(
          "Target momentum dimensionality must match lattice dimensionality.");
    mb_hubbard *h = nullptr;
    check(mb_momentum_hubbard_create(basis.handle(), (int)lattice_size.size(), lattice_size.data(),
                                     t, u, target_momentum.data(), &h));
    return HubbardOperator(h, basis.orbitals);
  }
  HubbardOperator(const HubbardOperator &) = delete;
  ~HubbardOperator() { mb_hubbard_destroy(h_); }

  size_t dimension() const { return (size_t)mb_hubbard_dimension(h_); }
  // LinearOperator::apply: a new vector (device order)
  DeviceVector apply(const DeviceVector &v, hipStream_t s = nullptr) const {
    DeviceVector y(dimension());
    check(mb_hv_c128(h_, v.data(), y.data(), s));
    check(hipStreamSynchronize(s) == hipSuccess ? MB_OK : MB_ERR_DEVICE);
    return y;
  }
  void apply_into(const DeviceVector &v, DeviceVector &y, hipStream_t s = nullptr) const {
    check(mb_hv_c128(h_, v.data(), y.data(), s));
  }
  DeviceVector to_device_order(const DeviceVector &x_ref, hipStream_t s = nullptr) const {
    DeviceVector y(dimension());
    check(mb_to_device_order_c128(h_, x_ref.data(), y.data(), s));
    return y;
  }
  DeviceVector to_reference_order(const DeviceVector &x_dev, hipStream_t s = nullptr) const {
    DeviceVector y(dimension());
    check(mb_to_reference_order_c128(h_, x_dev.data(), y.data(), s));
    return y;
  }
  DeviceVector gaussian(uint64_t seed, hipStream_t s = nullptr) const {
    DeviceVector y(dimension());
    check(mb_fill_gaussian_c128(h_, seed, y.data(), s));
    return y;
  }
  // HubbardModel*::opposite_spin_correlation (hubbard_model.h:40-53) for every
  // displacement r at once: g[r] = <x|G_updn(r)|x> / <x|x>, x in device order
  std::vector<double> opposite_spin_correlation(const DeviceVector &x,
                                                hipStream_t s = nullptr) const {
    std::vector<double> g(sites_);
    check(mb_opposite_spin_correlation(h_, x.data(), g.data(), s));
    return g;
  }
  // (<x|H|x> / <x|x>, |Hx - E x| / |x|)
  std::pair<double, double> energy_residual(const DeviceVector &x, hipStream_t s = nullptr) const {
    double e = 0.0, r = 0.0;
    check(mb_energy_residual(h_, x.data(), &e, &r, s));
    return {e, r};
  }
  const mb_hubbard *handle() const { return h_; }
  // takes ownership of a handle made by another factory entry point
  static HubbardOperator adopt(mb_hubbard *h, size_t sites) { return HubbardOperator(h, sites); }
  size_t sites() const { return sites_; }

 private:
  HubbardOperator(mb_hubbard *h, size_t sites) : h_(h), sites_(sites) {}
  mb_hubbard *h_ = nullptr;
  size_t sites_ = 0;
};

// HubbardMomentumFactorizedFixedK: a HubbardOperator on a fixed-momentum basis
inline HubbardOperator HubbardMomentumFactorizedFixedK(const FermionBasis &basis,
                                                       const std::vector<int> &lattice_size,
                                                       double t, double u,
                                                       const std::vector<int> &target_momentum) {
  return HubbardOperator::momentum_factorized_fixed_k(basis, lattice_size, t, u, target_momentum);
}

struct HubbardModel {
  HubbardModel(double t_val, double u_val, size_t size_val) : t(t_val), u(u_val), size(size_val) {}
  HubbardOperator op(const FermionBasis &b) const { return {b, {(int)size}, t, u}; }
  double t, u;
  size_t size;
};
struct HubbardModel2D {
  HubbardModel2D(double t_val, double u_val, size_t sx, size_t sy)
      : t(t_val), u(u_val), size_x(sx), size_y(sy) {}
  HubbardOperator op(const FermionBasis &b) const { return {b, {(int)size_x, (int)size_y}, t, u}; }
  double t, u;
  size_t size_x, size_y;
};
struct HubbardModel3D {
  HubbardModel3D(double t_val, double u_val, size_t sx, size_t sy, size_t sz)
      : t(t_val), u(u_val), size_x(sx), size_y(sy), size_z(sz) {}
  HubbardOperator op(const FermionBasis &b) const {
    return {b, {(int)size_x, (int)size_y, (int)size_z}, t, u};
  }
  double t, u;
  size_t size_x, size_y, size_z;
};

struct LanczosDecomposition {
  std::vector<double> alphas{};
  std::vector<double> betas{};
  size_t steps_taken = 0;
  double b_norm = 0.0;
};

// reorthogonalize: full reorthogonalisation (an option; all k + 1 vectors kept in HBM).
inline LanczosDecomposition lanczos_pass_one(const HubbardOperator &op, const DeviceVector &b,
                                             size_t k, hipStream_t s = nullptr,
                                             bool reorthogonalize = false) {
  if (op.dimension() != b.size())
    throw std::runtime_error("Operator dimension does not match input vector");
  std::vector<double> al(k ? k : 1), be(k ? k : 1);
  int steps = 0;
  double bn = 0.0;
  check(mb_lanczos_pass_one_ex(op.handle(), b.data(), (int)k,
                               reorthogonalize ? MB_LANCZOS_REORTHOGONALIZE : 0, al.data(), be.data(),
                               &steps, &bn, s));
  LanczosDecomposition d;
  d.alphas.assign(al.begin(), al.begin() + steps);
  d.betas.assign(be.begin(), be.begin() + (steps > 0 ? steps - 1 : 0));
  d.steps_taken = (size_t)steps;
  d.b_norm = bn;
  return d;
}

struct EigenPair {
  double value;
  DeviceVector vector;
};

// find_min_eigenpair(op, k) with an explicit start vector (the reference seeds
// it with arma::randu internally).
inline EigenPair find_min_eigenpair(const HubbardOperator &op, size_t k, const DeviceVector &b,
                                    hipStream_t s = nullptr, bool reorthogonalize = false) {
  EigenPair ep{0.0, DeviceVector(op.dimension())};
  check(mb_find_min_eigenpair_ex(op.handle(), b.data(), (int)k,
                                 reorthogonalize ? MB_LANCZOS_REORTHOGONALIZE : 0, &ep.value,
                                 ep.vector.data(), s));
  return ep;
}

// EvolutionOptions (evolve_state.h:110-112); fp32 is an option of this library:
// Krylov vectors and H.v in FP32, state and reductions in FP64.
struct EvolutionOptions {
  size_t krylov_steps = 30;
  bool fp32 = false;
};

inline DeviceVector time_evolve_state(const HubbardOperator &H, const DeviceVector &psi, double t,
                                      EvolutionOptions opts = {}, hipStream_t s = nullptr) {
  DeviceVector out(H.dimension());
  check(mb_time_evolve_ex(H.handle(), psi.data(), t, (int)opts.krylov_steps,
                          opts.fp32 ? MB_EVOLVE_FP32 : 0, out.data(), s));
  return out;
}

inline DeviceVector time_evolve_state_steps(const HubbardOperator &H, const DeviceVector &psi,
                                            double t0, double t1, double dt,
                                            EvolutionOptions opts = {}, hipStream_t s = nullptr) {
  DeviceVector out(H.dimension());
  check(mb_time_evolve_steps_ex(H.handle(), psi.data(), t0, t1, dt, (int)opts.krylov_steps,
                                opts.fp32 ? MB_EVOLVE_FP32 : 0, out.data(), s));
  return out;
}

// with callback(t, state_dev) at t0 and after every step (evolve_state.h:199-231);
// state_dev is the device-order vector being evolved, valid during the call
inline DeviceVector time_evolve_state_steps(
    const HubbardOperator &H, const DeviceVector &psi, double t0, double t1, double dt,
    EvolutionOptions opts, const std::function<void(double, const void *)> &callback,
    hipStream_t s = nullptr) {
  DeviceVector out(H.dimension());
  auto tramp = [](double t, const void *state, void *user) -> int {
    try {
      (*static_cast<const std::function<void(double, const void *)> *>(user))(t, state);
      return 0;
    } catch (...) {
      return 1;
    }
  };
  check(mb_time_evolve_steps_cb(H.handle(), psi.data(), t0, t1, dt, (int)opts.krylov_steps,
                                opts.fp32 ? MB_EVOLVE_FP32 : 0, tramp,
                                const_cast<std::function<void(double, const void *)> *>(&callback),
                                out.data(), s));
  return out;
}

// power_method(op, v, iterations) and estimate_bounds(op, power_iterations,
// spectral_padding) (linear_operator_utils.h:10-78); the seed is explicit.
// solve(op, b, k, solver) (lanczos.h:220-237): solver(alphas, betas) gives the
// first-column coefficients; returns |b| sum_j y_j v_j (device order)
using TridiagSolver = std::function<std::vector<std::complex<double>>(
    const std::vector<double> &, const std::vector<double> &)>;
inline DeviceVector solve(const HubbardOperator &H, const DeviceVector &b, size_t k,
                          const TridiagSolver &solver, hipStream_t s = nullptr) {
  struct Ctx {
    const TridiagSolver *f;
  } ctx{&solver};
  auto tramp = [](int m, const double *al, const double *be, double *y, void *user) -> int {
    try {
      const auto &f = *static_cast<Ctx *>(user)->f;
      const auto r = f(std::vector<double>(al, al + m),
                       std::vector<double>(be, be + (m > 0 ? m - 1 : 0)));
      if ((int)r.size() != m) return 1;
      for (int i = 0; i < m; ++i) {
        y[2 * i] = r[i].real();
        y[2 * i + 1] = r[i].imag();
      }
      return 0;
    } catch (...) {
      return 1;
    }
  };
  DeviceVector x(H.dimension());
  check(mb_lanczos_solve(H.handle(), b.data(), (int)k, tramp, &ctx, x.data(), s));
  return x;
}

// find_max_eigenpair(op, k) (lanczos.h:246-281) with an explicit start vector
inline EigenPair find_max_eigenpair(const HubbardOperator &H, size_t k, const DeviceVector &b,
                                    hipStream_t s = nullptr) {
  EigenPair ep{0.0, DeviceVector(H.dimension())};
  check(mb_find_max_eigenpair(H.handle(), b.data(), (int)k, &ep.value, ep.vector.data(), s));
  return ep;
}

// Exp(op, ExpOptions{krylov_steps}) (exp_operator.h:46-89): apply(v) = exp(H) v
class Exp {
 public:
  explicit Exp(const HubbardOperator &op, size_t krylov_steps = 30)
      : op_(op), k_(krylov_steps) {}
  DeviceVector apply(const DeviceVector &v, hipStream_t s = nullptr) const {
    DeviceVector out(op_.dimension());
    check(mb_exp_apply(op_.handle(), v.data(), (int)k_, out.data(), s));
    return out;
  }
  size_t dimension() const { return op_.dimension(); }

 private:
  const HubbardOperator &op_;
  size_t k_;
};

// Bounds (linear_operator_utils.h:10-16) for the complex operators: alpha <= beta
struct Bounds {
  double alpha = 0.0, beta = 0.0;
};
inline double power_method(const HubbardOperator &H, const DeviceVector &seed, size_t iterations,
                           hipStream_t s = nullptr) {
  double v = 0.0;
  check(mb_power_method(H.handle(), seed.data(), (int)iterations, &v, s));
  return v;
}
inline Bounds estimate_bounds(const HubbardOperator &H, const DeviceVector &seed,
                                size_t power_iterations, double spectral_padding,
                                hipStream_t s = nullptr) {
  Bounds b;
  check(mb_estimate_bounds(H.handle(), seed.data(), (int)power_iterations, spectral_padding,
                           &b.alpha, &b.beta, s));
  return b;
}

// exp(-tau H) psi (imaginary_time_evolve_state, evolve_state.h:184-197)
inline DeviceVector imaginary_time_evolve_state(const HubbardOperator &H, const DeviceVector &psi,
                                                double tau, EvolutionOptions opts = {},
                                                hipStream_t s = nullptr) {
  DeviceVector out(H.dimension());
  check(mb_imaginary_time_evolve(H.handle(), psi.data(), tau, (int)opts.krylov_steps, out.data(),
                                 s));
  return out;
}

// CurrentOperatorFixedK(source_basis, target_basis, lattice_size, t, Q, direction)
// (current_operator_fixed_k.h:28-69): sector K -> K + Q, Q a flat momentum index.
class CurrentOperatorFixedK {
 public:
  CurrentOperatorFixedK(const FermionBasis &source_basis, const FermionBasis &target_basis,
                        const std::vector<int> &lattice_size, double t, size_t Q,
                        size_t direction)
      : ns_(source_basis.size()), nt_(target_basis.size()) {
    if (lattice_size.empty())
      throw std::invalid_argument("CurrentOperatorFixedK requires at least one dimension.");
    check(mb_current_create(source_basis.handle(), target_basis.handle(), (int)lattice_size.size(),
                            lattice_size.data(), t, (int64_t)Q, (int)direction, &j_));
  }
  CurrentOperatorFixedK(const CurrentOperatorFixedK &) = delete;
  ~CurrentOperatorFixedK() { mb_current_destroy(j_); }
  size_t source_dimension() const { return ns_; }
  size_t target_dimension() const { return nt_; }
  DeviceVector apply(const DeviceVector &v, hipStream_t s = nullptr) const {
    DeviceVector w(nt_);
    check(mb_current_apply(j_, v.data(), w.data(), s));
    return w;
  }
  DeviceVector adjoint_apply(const DeviceVector &w, hipStream_t s = nullptr) const {
    DeviceVector v(ns_);
    check(mb_current_adjoint_apply(j_, w.data(), v.data(), s));
    return v;
  }
  const mb_current *handle() const { return j_; }

 private:
  mb_current *j_ = nullptr;
  size_t ns_ = 0, nt_ = 0;
};

// The stochastic current-current correlator of
// hubbard_momentum_optical_conductivity_factorized.cpp:226-283 (sample n starts
// from H_K.gaussian(seed + n)).
inline std::vector<std::complex<double>> optical_correlator(
    const HubbardOperator &H_K, const HubbardOperator &H_KQ, const CurrentOperatorFixedK &J,
    size_t samples, double beta, double dt, size_t steps, uint64_t seed,
    EvolutionOptions opts = {}, hipStream_t s = nullptr) {
  std::vector<std::complex<double>> c(steps);
  check(mb_optical_correlator(H_K.handle(), H_KQ.handle(), J.handle(), (int)samples, beta, dt,
                              (int)steps, (int)opts.krylov_steps, seed,
                              reinterpret_cast<double *>(c.data()), s));
  return c;
}

// OpticalConductivityResult, compute_optical_conductivity (optical_conductivity.h:6-13)
struct OpticalConductivityResult {
  std::vector<double> frequencies{};
  std::vector<std::complex<double>> sigma{};
  double sum_rule = 0.0;
};

inline OpticalConductivityResult compute_optical_conductivity(
    const std::vector<std::complex<double>> &correlation, double dt, double beta, double volume) {
  const double *c = reinterpret_cast<const double *>(correlation.data());
  int nf = 0;
  check(mb_optical_conductivity(c, (int)correlation.size(), dt, beta, volume, nullptr, nullptr, 0,
                                &nf, nullptr));
  OpticalConductivityResult r;
  r.frequencies.resize(nf);
  r.sigma.resize(nf);
  check(mb_optical_conductivity(c, (int)correlation.size(), dt, beta, volume, r.frequencies.data(),
                                reinterpret_cast<double *>(r.sigma.data()), nf, &nf, &r.sum_rule));
  return r;
}

// One term of a FermionExpression: coefficient and operator string in the
// reference's OperatorStorage encoding (bit 15 annihilation, bit 14 spin down,
// bits 0-13 orbital), o_1 o_2 ... o_n with o_n acting first.
struct ExpressionTerm {
  std::complex<double> c;
  std::vector<uint16_t> ops;
};

// The reference's operator algebra (linear_operator.h:93-281) on device
// operators: -A, A + B, A - B, A * B (B applied first), A * c, c * A,
// Shifted(A, c) = A + c I, NegShift(A, c) = -A + c I, Identity(basis).  The
// result references its operands (the reference's wrappers copy them): keep
// them alive while it is used.
inline HubbardOperator operator-(const HubbardOperator &a) {
  mb_hubbard *h = nullptr;
  check(mb_operator_negated(a.handle(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator operator*(const HubbardOperator &a, std::complex<double> c) {
  mb_hubbard *h = nullptr;
  check(mb_operator_scaled(a.handle(), c.real(), c.imag(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator operator*(std::complex<double> c, const HubbardOperator &a) { return a * c; }
inline HubbardOperator Shifted(const HubbardOperator &a, std::complex<double> c) {
  mb_hubbard *h = nullptr;
  check(mb_operator_shifted(a.handle(), c.real(), c.imag(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator NegShift(const HubbardOperator &a, std::complex<double> c) {
  mb_hubbard *h = nullptr;
  check(mb_operator_neg_shifted(a.handle(), c.real(), c.imag(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator operator+(const HubbardOperator &a, const HubbardOperator &b) {
  mb_hubbard *h = nullptr;
  check(mb_operator_sum(a.handle(), b.handle(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator operator-(const HubbardOperator &a, const HubbardOperator &b) {
  mb_hubbard *h = nullptr;
  check(mb_operator_difference(a.handle(), b.handle(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator operator*(const HubbardOperator &a, const HubbardOperator &b) {
  mb_hubbard *h = nullptr;
  check(mb_operator_composed(a.handle(), b.handle(), &h));
  return HubbardOperator::adopt(h, a.sites());
}
inline HubbardOperator Identity(const FermionBasis &basis) {
  mb_hubbard *h = nullptr;
  check(mb_operator_identity(basis.handle(), &h));
  return HubbardOperator::adopt(h, basis.orbitals);
}

// compute_matrix_elements(basis, A) as a device operator (H.v, Lanczos,
// eigenpair, time evolution, energy/residual); terms merged by key as the
// reference's FermionExpression does.
inline HubbardOperator expression_operator(const FermionBasis &basis,
                                           const std::vector<ExpressionTerm> &terms) {
  std::vector<double> coeffs;
  std::vector<int64_t> offsets{0};
  std::vector<uint16_t> ops;
  for (const auto &t : terms) {
    coeffs.push_back(t.c.real());
    coeffs.push_back(t.c.imag());
    ops.insert(ops.end(), t.ops.begin(), t.ops.end());
    offsets.push_back((int64_t)ops.size());
  }
  mb_hubbard *h = nullptr;
  check(mb_expression_create(basis.handle(), (int64_t)terms.size(), coeffs.data(), offsets.data(),
                             ops.data(), &h));
  return HubbardOperator::adopt(h, basis.orbitals);
}

// arma::sp_cx_mat as compute_matrix_elements returns it (entries column-major,
// rows ascending), with the reference's cereal save / load.
class SparseMatrix {
 public:
  explicit SparseMatrix(mb_sparse *s) : s_(s) {}
  SparseMatrix(SparseMatrix &&o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
  SparseMatrix(const SparseMatrix &) = delete;
  ~SparseMatrix() { mb_sparse_destroy(s_); }
  uint64_t n_rows() const { return shape()[0]; }
  uint64_t n_cols() const { return shape()[1]; }
  int64_t n_nonzero() const {
    uint64_t r, c;
    int64_t n;
    check(mb_sparse_shape(s_, &r, &c, &n));
    return n;
  }
  void entries(std::vector<uint64_t> &rows, std::vector<uint64_t> &cols,
               std::vector<std::complex<double>> &vals) const {
    const int64_t n = n_nonzero();
    rows.resize(n);
    cols.resize(n);
    vals.resize(n);
    check(mb_sparse_entries(s_, rows.data(), cols.data(), reinterpret_cast<double *>(vals.data())));
  }
  void save(const std::string &path) const { check(mb_sparse_save_cereal(s_, path.c_str())); }
  static SparseMatrix load(const std::string &path) {
    mb_sparse *s = nullptr;
    check(mb_sparse_load_cereal(path.c_str(), &s));
    return SparseMatrix(s);
  }

 private:
  std::vector<uint64_t> shape() const {
    uint64_t r, c;
    int64_t n;
    check(mb_sparse_shape(s_, &r, &c, &n));
    return {r, c};
  }
  mb_sparse *s_ = nullptr;
};

inline SparseMatrix compute_matrix_elements(const HubbardOperator &expression_op,
                                            hipStream_t s = nullptr) {
  mb_sparse *sp = nullptr;
  check(mb_compute_matrix_elements(expression_op.handle(), &sp, s));
  return SparseMatrix(sp);
}

namespace detail {
struct TermArrays {
  std::vector<double> coeffs;
  std::vector<int64_t> offsets{0};
  std::vector<uint16_t> ops;
  explicit TermArrays(const std::vector<ExpressionTerm> &terms) {
    for (const auto &t : terms) {
      coeffs.push_back(t.c.real());
      coeffs.push_back(t.c.imag());
      ops.insert(ops.end(), t.ops.begin(), t.ops.end());
      offsets.push_back((int64_t)ops.size());
    }
    if (ops.empty()) ops.push_back(0);
  }
  int64_t n() const { return (int64_t)offsets.size() - 1; }
};
}  // namespace detail

// compute_rectangular_matrix_elements<arma::sp_cx_mat>(row_basis, col_basis, A)
// (matrix_elements.h:110-167), built on the device
inline SparseMatrix compute_rectangular_matrix_elements(const FermionBasis &row_basis,
                                                        const FermionBasis &col_basis,
                                                        const std::vector<ExpressionTerm> &terms,
                                                        hipStream_t s = nullptr) {
  detail::TermArrays a(terms);
  mb_sparse *sp = nullptr;
  check(mb_compute_rectangular_matrix_elements(row_basis.handle(), col_basis.handle(), a.n(),
                                               a.coeffs.data(), a.offsets.data(), a.ops.data(),
                                               &sp, s));
  return SparseMatrix(sp);
}

// compute_vector_elements<arma::cx_vec>(basis, A) (matrix_elements.h:31-51): <i|A|0>
inline std::vector<std::complex<double>> compute_vector_elements(
    const FermionBasis &basis, const std::vector<ExpressionTerm> &terms, hipStream_t s = nullptr) {
  detail::TermArrays a(terms);
  std::vector<std::complex<double>> v(std::max<size_t>(basis.size(), 1));
  check(mb_compute_vector_elements(basis.handle(), a.n(), a.coeffs.data(), a.offsets.data(),
                                   a.ops.data(), reinterpret_cast<double *>(v.data()), s));
  v.resize(basis.size());
  return v;
}

// Serving loop: requests admitted as they arrive and as slots and device
// memory free up (mb_server_*).
class Server {
 public:
  Server(int slots, int64_t arena_bytes, int64_t lane_bytes = 0) {
    check(mb_server_create(slots, arena_bytes, lane_bytes, &h_));
  }
  ~Server() { mb_server_destroy(h_); }
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;
  int64_t submit(const mb_request &r) {
    int64_t id = -1;
    check(mb_server_submit(h_, &r, &id));
    return id;
  }
  // Blocks until request id has finished; KPM moments go to *moments.
  mb_result wait(int64_t id, std::vector<double> *moments = nullptr) {
    mb_result r;
    std::vector<double> mu(moments ? 4096 : 0);
    check(mb_server_wait(h_, id, &r, moments ? mu.data() : nullptr, (int)mu.size()));
    if (moments) moments->assign(mu.begin(), mu.begin() + r.n_moments);
    return r;
  }
  void drain() { check(mb_server_drain(h_)); }

 private:
  mb_server *h_ = nullptr;
};

namespace kpm {
inline std::vector<double> jackson_kernel(size_t N) {
  std::vector<double> g(N + 1);
  check(mb_jackson_kernel((int)N, g.data()));
  return g;
}
// Chebyshev DOS moments mu_0..mu_{M-1} of (H - b)/a over nvec random-phase vectors.
inline std::vector<double> dos_moments(const HubbardOperator &H, double a, double b, size_t M,
                                       int nvec, uint64_t seed, bool fp32 = true,
                                       hipStream_t s = nullptr) {
  std::vector<double> mu(M);
  check(mb_kpm_dos_moments(H.handle(), a, b, (int)M, nvec, seed, fp32 ? 1 : 0, mu.data(), s));
  return mu;
}
}  // namespace kpm

namespace rescaling {
struct Rescaling {
  double a, b;
  double rescale(double E) const { return (E - b) / a; }
  double inverse(double E_scaled) const { return E_scaled * a + b; }
};
inline Rescaling from_bounds(double E_min, double E_max, double padding = 0.01) {
  Rescaling r{};
  check(mb_rescaling_from_bounds(E_min, E_max, padding, &r.a, &r.b));
  return r;
}
// estimate_rescaling (rescaling.h:21-31) for a device operator: extreme
// eigenvalues from min(k, D) Lanczos steps from a seeded start (exact up to
// rounding once k >= D) instead of a dense eig_sym
inline Rescaling estimate_rescaling(const HubbardOperator &H, double padding = 0.01,
                                    size_t k = 300, uint64_t seed = 1) {
  auto b = H.gaussian(seed);
  const size_t kk = std::min<size_t>(k, H.dimension());
  const double lo = find_min_eigenpair(H, kk, b).value;
  const double hi = find_max_eigenpair(H, kk, b).value;
  return from_bounds(lo, hi, padding);
}
// rescale_hamiltonian (rescaling.h:40-48): (H - b I) / a; references H
inline HubbardOperator rescale_hamiltonian(const HubbardOperator &H, const Rescaling &r) {
  mb_hubbard *h = nullptr;
  check(mb_operator_rescaled(H.handle(), r.a, r.b, &h));
  return HubbardOperator::adopt(h, H.sites());
}
}  // namespace rescaling

}  // namespace mbamd
