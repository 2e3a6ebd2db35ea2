// Current operator between momentum sectors, imaginary-time evolution and the
// stochastic current-current correlator of the optical conductivity (see
// conductivity.cpp).
#pragma once

#include <complex>
#include <vector>

#include "krylov.hpp"
#include "momentum.hpp"

namespace mb {

// One-body momentum transfer sum_s w(s) c+_{s+Q,sigma} c_{s,sigma} (both spins)
// from sector S to sector T = S + Q; w is indexed by the source orbital s.
struct TransferTables {
  int N, Q, mu, md;
  int64_t rows;                             // up strings (target rows)
  const uint16_t *up_src, *dn_src;          // source index inside its momentum group
  const double *up_w, *dn_w;                // signed weights
  const uint8_t *up_cnt, *dn_cnt;
  const int32_t *up_grp, *up_loc;           // momentum group of each up string, index in it
  const int32_t *dn_off_t, *msub;           // target down group offsets, p - q table
  const int64_t *blk_s, *blk_t;             // block offsets of S and T
  const int32_t *nc_s, *nc_t;               // block columns of S and T
  int KS, KT;                               // total momenta
  int64_t DS, DT;
  int max_cols;
};

class MomentumTransfer {
 public:
  MomentumTransfer(const MomentumSector &S, const MomentumSector &T, int Q,
                   const std::vector<double> &w);
  void apply(const double *x_src, double *y_tgt, hipStream_t s) const;
  int64_t source_dim() const { return tb_.DS; }
  int64_t target_dim() const { return tb_.DT; }

 private:
  TransferTables tb_{};
  DevBuf d_up_src_, d_dn_src_, d_up_w_, d_dn_w_, d_up_cnt_, d_dn_cnt_, d_up_grp_, d_up_loc_,
      d_dn_off_t_, d_msub_, d_blk_s_, d_blk_t_, d_nc_s_, d_nc_t_;
};

// CurrentOperatorFixedK (current_operator_fixed_k.h:28-293):
// J_d(Q) = sum_{k,s} v_d(k) c+_{k+Q,s} c_{k,s}, v_d(k) = 2 t (2 pi / L_d) sin(2 pi k_d / L_d),
// sector K -> K + Q; adjoint_apply is J_d(Q)^dagger, K + Q -> K.
class CurrentOperator {
 public:
  CurrentOperator(const MomentumSector &source, const MomentumSector &target, double t, int Q,
                  int direction);
  void apply(const double *x_k, double *y_kq, hipStream_t s) const { fwd_.apply(x_k, y_kq, s); }
  void adjoint_apply(const double *y_kq, double *x_k, hipStream_t s) const {
    adj_.apply(y_kq, x_k, s);
  }
  int64_t source_dim() const { return fwd_.source_dim(); }
  int64_t target_dim() const { return fwd_.target_dim(); }
  const std::vector<double> &velocities() const { return v_; }

 private:
  std::vector<double> v_;
  MomentumTransfer fwd_, adj_;
};

// psi(tau) = exp(-tau H) psi (imaginary_time_evolve_state, evolve_state.h:184-197)
void imaginary_time_evolve(const HvOp &op, const void *psi, double tau, int krylov_steps,
                           void *out, hipStream_t s);

// Stochastic current-current correlator (hubbard_momentum_optical_conductivity_
// factorized.cpp:226-283): per sample a seeded N(0, 1) vector v in sector K
// (per reference index, then normalised), v_b = exp(-beta H_K / 2) v,
// Z = <v_b|v_b>, phi1 = v_b, phi2 = J v_b; C(t_i) += <phi1|J^dagger phi2> / Z,
// then phi1 <- exp(-i H_K dt) phi1, phi2 <- exp(-i H_{K+Q} dt) phi2; averaged
// over the samples.  Sample n uses seed + n.
std::vector<std::complex<double>> optical_correlator(const HvOp &hk, const HvOp &hkq,
                                                     const CurrentOperator &J,
                                                     const int64_t *ref_of_pos_k, int samples,
                                                     double beta, double dt, int steps,
                                                     int krylov_steps, uint64_t seed,
                                                     hipStream_t s);

// compute_optical_conductivity (optical_conductivity.cpp:8-80)
struct Conductivity {
  std::vector<double> omega;
  std::vector<std::complex<double>> sigma;
  double sum_rule = 0.0;
};
Conductivity optical_conductivity(const std::vector<std::complex<double>> &corr, double dt,
                                  double beta, double volume);

// kernels (kernels.hip)
void launch_transfer(const TransferTables &t, const double *x_src, double *y_tgt, hipStream_t s);

}  // namespace mb
