// Current operator between momentum sectors, imaginary-time evolution and the
// stochastic current-current correlator of the optical conductivity.
//
// Transfer.  sum_s w(s) c+_{s+Q} c_s of one species moves a string's orbital s to
// s + Q with weight w(s) and the sign (-1)^(occupied orbitals strictly between);
// from the target side a target string's sources are the strings with one of
// its orbitals s + Q moved back to s.  With Q = 0 it is diagonal: weight
// sum_{s in string} w(s).  A target state (u', d') gathers the up moves from
// (u, d') and the down moves from (u', d) of the source sector; both source
// blocks have the target block's column or row structure (momentum groups are
// the same in every sector), see transfer_kernel.
#include "conductivity.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mb {

namespace {

constexpr double kPi = 3.14159265358979323846;

int between(uint32_t s, int a, int b) {
  const int lo = std::min(a, b), hi = std::max(a, b);
  const uint32_t mask = ((1u << hi) - 1u) & ~((2u << lo) - 1u);
  return __builtin_popcount(s & mask);
}

}  // namespace

MomentumTransfer::MomentumTransfer(const MomentumSector &S, const MomentumSector &T, int Q,
                                   const std::vector<double> &w) {
  const int N = S.L;
  if (T.L != N || T.n_up != S.n_up || T.n_dn != S.n_dn || T.ndim != S.ndim ||
      T.dims[0] != S.dims[0] || T.dims[1] != S.dims[1] || T.dims[2] != S.dims[2])
    throw std::invalid_argument("source and target bases differ in lattice or filling");
  if (T.K != S.madd(S.K, Q)) throw std::invalid_argument("target momentum must be K + Q");
  if ((int)w.size() != N) throw std::invalid_argument("one weight per orbital");
  auto local = [&](const std::vector<uint32_t> &strs, const std::vector<int32_t> &off, uint32_t x) {
    const int p = S.momentum_of(x);
    auto f = std::lower_bound(strs.begin() + off[p], strs.begin() + off[p + 1], x);
    return (int)(f - (strs.begin() + off[p]));
  };
  auto lists = [&](const std::vector<uint32_t> &strs, const std::vector<int32_t> &off, int width,
                   std::vector<uint16_t> &src, std::vector<double> &wt, std::vector<uint8_t> &cnt) {
    const size_t ns = strs.size();
    src.assign(ns * width, 0);
    wt.assign(ns * width, 0.0);
    cnt.assign(ns, 0);
    for (size_t g = 0; g < ns; ++g) {
      const uint32_t o = strs[g];
      int c = 0;
      if (Q == 0) {  // diagonal: sum of the weights of the occupied orbitals
        double sw = 0.0;
        for (int k = 0; k < N; ++k)
          if ((o >> k) & 1u) sw += w[k];
        src[g * width] = (uint16_t)local(strs, off, o);
        wt[g * width] = sw;
        cnt[g] = 1;
        continue;
      }
      for (int kt = 0; kt < N; ++kt) {  // target orbital kt = s + Q
        if (!((o >> kt) & 1u)) continue;
        const int ks = S.msub(kt, Q);
        if ((o >> ks) & 1u) continue;
        const uint32_t sstr = (o & ~(1u << kt)) | (1u << ks);
        src[g * width + c] = (uint16_t)local(strs, off, sstr);
        wt[g * width + c] = ((between(o, kt, ks) & 1) ? -1.0 : 1.0) * w[ks];
        ++c;
      }
      cnt[g] = (uint8_t)c;
    }
  };
  const int mu = std::max(1, std::min(S.n_up, N - S.n_up));
  const int md = std::max(1, std::min(S.n_dn, N - S.n_dn));
  std::vector<uint16_t> us, ds;
  std::vector<double> uw, dw;
  std::vector<uint8_t> uc, dc;
  lists(S.up_str, S.up_off, mu, us, uw, uc);
  lists(S.dn_str, S.dn_off, md, ds, dw, dc);
  std::vector<int32_t> grp(S.up_str.size()), loc(S.up_str.size());
  for (int p = 0; p < N; ++p)
    for (int g = S.up_off[p]; g < S.up_off[p + 1]; ++g) {
      grp[g] = p;
      loc[g] = g - S.up_off[p];
    }
  std::vector<int32_t> msub((size_t)N * N);
  for (int p = 0; p < N; ++p)
    for (int q = 0; q < N; ++q) msub[(size_t)p * N + q] = S.msub(p, q);
  for (int p = 0; p < N; ++p)  // an up move keeps the column group: same column count
    if (T.ncols[p] && T.ncols[p] != S.ncols[S.msub(p, Q)])
      throw std::runtime_error("inconsistent momentum blocks");
  d_up_src_ = upload(us);
  d_dn_src_ = upload(ds);
  d_up_w_ = upload(uw);
  d_dn_w_ = upload(dw);
  d_up_cnt_ = upload(uc);
  d_dn_cnt_ = upload(dc);
  d_up_grp_ = upload(grp);
  d_up_loc_ = upload(loc);
  d_dn_off_t_ = upload(T.dn_off);
  d_msub_ = upload(msub);
  d_blk_s_ = upload(S.blk_off);
  d_blk_t_ = upload(T.blk_off);
  d_nc_s_ = upload(S.ncols);
  d_nc_t_ = upload(T.ncols);
  tb_.N = N;
  tb_.Q = Q;
  tb_.mu = mu;
  tb_.md = md;
  tb_.rows = (int64_t)S.up_str.size();
  tb_.up_src = d_up_src_.as<uint16_t>();
  tb_.dn_src = d_dn_src_.as<uint16_t>();
  tb_.up_w = d_up_w_.as<double>();
  tb_.dn_w = d_dn_w_.as<double>();
  tb_.up_cnt = d_up_cnt_.as<uint8_t>();
  tb_.dn_cnt = d_dn_cnt_.as<uint8_t>();
  tb_.up_grp = d_up_grp_.as<int32_t>();
  tb_.up_loc = d_up_loc_.as<int32_t>();
  tb_.dn_off_t = d_dn_off_t_.as<int32_t>();
  tb_.msub = d_msub_.as<int32_t>();
  tb_.blk_s = d_blk_s_.as<int64_t>();
  tb_.blk_t = d_blk_t_.as<int64_t>();
  tb_.nc_s = d_nc_s_.as<int32_t>();
  tb_.nc_t = d_nc_t_.as<int32_t>();
  tb_.KS = S.K;
  tb_.KT = T.K;
  tb_.DS = S.D;
  tb_.DT = T.D;
  tb_.max_cols = std::max(*std::max_element(S.ncols.begin(), S.ncols.end()),
                          *std::max_element(T.ncols.begin(), T.ncols.end()));
  if (tb_.max_cols > 4096)
    throw std::invalid_argument("momentum group too large for the transfer kernel");
}

void MomentumTransfer::apply(const double *x_src, double *y_tgt, hipStream_t s) const {
  launch_transfer(tb_, x_src, y_tgt, s);
}

static std::vector<double> current_velocities(const MomentumSector &m, double t, int direction) {
  if (direction < 0 || direction >= m.ndim)
    throw std::invalid_argument("Direction must be less than lattice dimensionality.");
  std::vector<double> v(m.L);
  const int Ld = m.dims[direction];
  for (int k = 0; k < m.L; ++k) {
    const int c[3] = {k % m.dims[0], (k / m.dims[0]) % m.dims[1], k / (m.dims[0] * m.dims[1])};
    const double phase = 2.0 * kPi * c[direction] / Ld;
    const double vk = 2.0 * t * (2.0 * kPi / Ld) * std::sin(phase);
    v[k] = std::fabs(vk) < 1e-15 ? 0.0 : vk;  // current_operator_fixed_k.h:160-163
  }
  return v;
}

static std::vector<double> shifted(const std::vector<double> &v, const MomentumSector &m, int Q) {
  std::vector<double> w(v.size());
  for (int s = 0; s < m.L; ++s) w[s] = v[m.msub(s, Q)];  // adjoint: w(k + Q) = v(k)
  return w;
}

CurrentOperator::CurrentOperator(const MomentumSector &source, const MomentumSector &target,
                                 double t, int Q, int direction)
    : v_(current_velocities(source, t, direction)),
      fwd_(source, target, Q, v_),
      adj_(target, source, source.msub(0, Q), shifted(v_, source, Q)) {}

// ------------------------------------------------------------ imaginary time

void imaginary_time_evolve(const HvOp &op, const void *psi, double tau, int krylov_steps,
                           void *out, hipStream_t s) {
  time_evolve_kind(op, psi, tau, krylov_steps, out, s, nullptr, true);
}

// ------------------------------------------------------------ correlator

std::vector<std::complex<double>> optical_correlator(const HvOp &hk, const HvOp &hkq,
                                                     const CurrentOperator &J,
                                                     const int64_t *ref_of_pos_k, int samples,
                                                     double beta, double dt, int steps,
                                                     int krylov_steps, uint64_t seed,
                                                     hipStream_t s) {
  if (samples < 1 || steps < 1) throw std::invalid_argument("samples and steps must be positive");
  if (!(beta >= 0.0) || !(dt > 0.0)) throw std::invalid_argument("beta >= 0 and dt > 0 required");
  const int64_t DK = hk.dim(), DQ = hkq.dim();
  if (J.source_dim() != DK || J.target_dim() != DQ)
    throw std::invalid_argument("current operator does not map sector K to sector K + Q");
  std::vector<std::complex<double>> C(steps, 0.0);
  DevBuf v(16 * DK), p1(16 * DK), t1(16 * DK), p2(16 * DQ);
  DevBuf red(sizeof(double) * (2 * kRedBlocks + 4));
  double *partial = red.as<double>(), *scal = partial + 2 * kRedBlocks;
  DevBuf cache_k, cache_q;  // stored Krylov bases of the two evolutions
  for (int n = 0; n < samples; ++n) {
    launch_mom_fill_gaussian(ref_of_pos_k, DK, seed + (uint64_t)n, v.as<double>(), s);
    launch_norm(v.as<double>(), 2 * DK, partial, scal, s);
    launch_scale_div<double>(v.as<double>(), scal, v.as<double>(), 2 * DK, s);
    time_evolve_kind(hk, v.p, 0.5 * beta, krylov_steps, p1.p, s, &cache_k, true);
    launch_sumsq(p1.p, false, 2 * DK, partial, scal, false, s);
    double Z = 0.0;
    MB_HIP(hipMemcpyAsync(&Z, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    MB_HIP(hipStreamSynchronize(s));
    if (!(Z > 0.0)) throw std::runtime_error("zero partition-function estimate");
    J.apply(p1.as<double>(), p2.as<double>(), s);
    for (int i = 0; i < steps; ++i) {
      J.adjoint_apply(p2.as<double>(), t1.as<double>(), s);
      launch_cdot(p1.as<double>(), t1.as<double>(), DK, partial, scal, s);
      double c[2];
      MB_HIP(hipMemcpyAsync(c, scal, sizeof(c), hipMemcpyDeviceToHost, s));
      MB_HIP(hipStreamSynchronize(s));
      C[i] += std::complex<double>(c[0], c[1]) / Z;
      if (i + 1 < steps) {
        time_evolve_kind(hk, p1.p, dt, krylov_steps, p1.p, s, &cache_k, false);
        time_evolve_kind(hkq, p2.p, dt, krylov_steps, p2.p, s, &cache_q, false);
      }
    }
  }
  for (auto &c : C) c /= (double)samples;
  return C;
}

// ------------------------------------------------------------ sigma(omega)

Conductivity optical_conductivity(const std::vector<std::complex<double>> &corr, double dt,
                                  double beta, double volume) {
  const size_t n = corr.size();
  if (n < 2) throw std::invalid_argument("correlation must have at least 2 samples");
  if (!(dt > 0.0) || !(volume > 0.0) || !(beta >= 0.0))
    throw std::invalid_argument("dt>0, volume>0, beta>=0 required");
  const double T = ((double)n - 1.0) * dt;
  const double eta = 4.0 / T, dw = eta / 4.0;
  const size_t nf = (size_t)std::floor((kPi / dt) / dw) + 1;
  std::vector<std::complex<double>> series(n);
  const double damp_step = std::exp(-eta * dt);
  double damp = 1.0;
  for (size_t k = 0; k < n; ++k) {
    const double wk = (k == 0 || k == n - 1) ? 0.5 * dt : dt;
    series[k] = corr[k] * (wk * damp);
    damp *= damp_step;
  }
  Conductivity r;
  r.omega.reserve(nf);
  r.sigma.reserve(nf);
  double sum = 0.0;
  for (size_t i = 0; i < nf; ++i) {
    const double om = (double)i * dw;
    const std::complex<double> step = std::exp(std::complex<double>(0.0, om * dt));
    std::complex<double> ph(1.0, 0.0), acc(0.0, 0.0);
    for (size_t k = 0; k < n; ++k) {
      acc += series[k] * ph;
      ph *= step;
    }
    const double pref = std::fabs(om) < 1e-12 ? beta / volume : -std::expm1(-beta * om) / (om * volume);
    r.omega.push_back(om);
    r.sigma.push_back(pref * acc);
    sum += r.sigma.back().real() * ((i == 0 || i == nf - 1) ? 0.5 : 1.0);
  }
  r.sum_rule = sum * dw;
  return r;
}

}  // namespace mb
