// Fixed-total-momentum sectors and the factorized momentum-space Hubbard
// operator of hubbard_momentum_factorized_fixed_k.h (device side in
// kernels.hip, mom_hv_kernel).
//
// Sector.  Orbital k is the momentum k_x + L_x k_y + L_x L_y k_z; a string's
// momentum is the component-wise periodic sum of its orbitals.  The reference
// enumerates the sector with the real-space DFS and keeps the leaves of total
// momentum K (basis.cpp:304-356), so its index is the rank among the sector's
// states in the real-space reference order: computed here by sorting the
// states' real-space reference indices (closed form, Sector::ref_index).
//
// Operator.  H = t sum_{k,s} eps(k) n_{k,s} + (U/N) sum_q rho_{q,up} rho_{-q,dn}
// with eps(k) = -2 sum_d cos(2 pi k_d / L_d) and rho_{q,s} = sum_k c+_{k+q,s}
// c_{k,s} (hubbard_momentum_factorized_fixed_k.h:92-117, 201-215).  On a state
// ordered (up orbitals ascending, then down orbitals ascending) c+_{k'} c_k of
// one species has the sign (-1)^(occupied orbitals of that species strictly
// between k and k'), as in real space (the reference's (-1)^(op_idx +
// insert_pos), :315-330, reduces to it).  rho_{0,s} = N_s, so the q = 0 term
// is the constant (U/N) N_up N_dn.
#include "momentum.hpp"

#include <mutex>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace mb {

namespace {

constexpr double kPi = 3.14159265358979323846;

}  // namespace

int MomentumSector::madd(int p, int q) const {
  const int Lx = dims[0], Ly = dims[1], Lz = dims[2];
  const int px = p % Lx, py = (p / Lx) % Ly, pz = p / (Lx * Ly);
  const int qx = q % Lx, qy = (q / Lx) % Ly, qz = q / (Lx * Ly);
  return (px + qx) % Lx + Lx * ((py + qy) % Ly) + Lx * Ly * ((pz + qz) % Lz);
}

int MomentumSector::msub(int p, int q) const {
  const int Lx = dims[0], Ly = dims[1], Lz = dims[2];
  const int px = p % Lx, py = (p / Lx) % Ly, pz = p / (Lx * Ly);
  const int qx = q % Lx, qy = (q / Lx) % Ly, qz = q / (Lx * Ly);
  return (px - qx + Lx) % Lx + Lx * ((py - qy + Ly) % Ly) + Lx * Ly * ((pz - qz + Lz) % Lz);
}

int MomentumSector::momentum_of(uint32_t bits) const {
  int m = 0;
  for (int k = 0; k < L; ++k)
    if ((bits >> k) & 1u) m = madd(m, k);
  return m;
}

void MomentumSector::ref_state(int64_t i, uint32_t *up, uint32_t *dn) const {
  if (i < 0 || i >= D) throw std::out_of_range("index out of range");
  const int64_t pos = pos_of_ref[i];
  const int p = (int)(std::upper_bound(blk_off.begin(), blk_off.end(), pos) - blk_off.begin()) - 1;
  const int64_t r = (pos - blk_off[p]) / ncols[p], c = (pos - blk_off[p]) % ncols[p];
  *up = up_str[up_off[p] + r];
  *dn = dn_str[dn_off[msub(K, p)] + c];
}

int64_t MomentumSector::ref_index(uint32_t up, uint32_t dn) const {
  if (__builtin_popcount(up) != n_up || __builtin_popcount(dn) != n_dn) return -1;
  if ((L < 32 && ((up >> L) || (dn >> L)))) return -1;
  const int p = momentum_of(up);
  if (madd(p, momentum_of(dn)) != K) return -1;
  const int pd = msub(K, p);
  auto fu = std::lower_bound(up_str.begin() + up_off[p], up_str.begin() + up_off[p + 1], up);
  auto fd = std::lower_bound(dn_str.begin() + dn_off[pd], dn_str.begin() + dn_off[pd + 1], dn);
  const int64_t r = fu - (up_str.begin() + up_off[p]), c = fd - (dn_str.begin() + dn_off[pd]);
  return ref_of_pos[blk_off[p] + r * ncols[p] + c];
}

void MomentumSector::ensure_device() {
  static std::mutex mu;  // the first upload is serialised across host threads
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  MB_HIP(hipGetDevice(&dev));
  if (device == dev) return;
  d_ref_of_pos = upload(ref_of_pos);
  d_pos_of_ref = upload(pos_of_ref);
  device = dev;
}

void build_momentum_sector(MomentumSector &m, int ndim, const int *sizes, int n_up, int n_dn,
                           const int *momentum) {
  if (ndim < 1 || ndim > 3) throw std::invalid_argument("ndim must be 1, 2 or 3");
  int L = 1;
  for (int d = 0; d < ndim; ++d) {
    if (sizes[d] < 1) throw std::invalid_argument("Dimension with size zero is invalid.");
    if (momentum[d] < 0 || momentum[d] >= sizes[d])
      throw std::invalid_argument("Target momentum components must be less than lattice size.");
    L *= sizes[d];
  }
  if (L > kMaxSites) throw std::invalid_argument("at most 32 orbitals");
  if (n_up < 0 || n_dn < 0 || n_up > L || n_dn > L)
    throw std::invalid_argument("spin sector exceeds the orbital count");
  m.L = L;
  m.n_up = n_up;
  m.n_dn = n_dn;
  m.ndim = ndim;
  for (int d = 0; d < 3; ++d) m.dims[d] = d < ndim ? sizes[d] : 1;
  m.K = momentum[0] + (ndim > 1 ? m.dims[0] * momentum[1] : 0) +
        (ndim > 2 ? m.dims[0] * m.dims[1] * momentum[2] : 0);
  build_sector(m.real, L, n_up, n_dn);
  const int N = L;
  auto group = [&](const std::vector<uint32_t> &strs, std::vector<uint32_t> &out,
                   std::vector<int32_t> &off) {
    std::vector<std::vector<uint32_t>> by(N);
    for (uint32_t s : strs) by[m.momentum_of(s)].push_back(s);  // colex order kept
    out.clear();
    off.assign(N + 1, 0);
    for (int p = 0; p < N; ++p) {
      off[p] = (int32_t)out.size();
      out.insert(out.end(), by[p].begin(), by[p].end());
    }
    off[N] = (int32_t)out.size();
  };
  group(m.real.s_up, m.up_str, m.up_off);
  group(m.real.s_dn, m.dn_str, m.dn_off);
  m.blk_off.assign(N + 1, 0);
  m.ncols.assign(N, 0);
  for (int p = 0; p < N; ++p) {
    const int pd = m.msub(m.K, p);
    m.ncols[p] = m.dn_off[pd + 1] - m.dn_off[pd];
    if (m.ncols[p] > 32767 || m.up_off[p + 1] - m.up_off[p] > 32767)
      throw std::invalid_argument("momentum group too large (more than 32767 strings)");
    m.blk_off[p + 1] = m.blk_off[p] + (int64_t)(m.up_off[p + 1] - m.up_off[p]) * m.ncols[p];
  }
  m.D = m.blk_off[N];
  // reference order: sort the sector's states by their real-space reference index
  std::vector<std::pair<int64_t, int64_t>> rk((size_t)m.D);
  for (int p = 0; p < N; ++p) {
    const int pd = m.msub(m.K, p);
    const int64_t nr = m.up_off[p + 1] - m.up_off[p];
    for (int64_t r = 0; r < nr; ++r)
      for (int64_t c = 0; c < m.ncols[p]; ++c) {
        const int64_t pos = m.blk_off[p] + r * m.ncols[p] + c;
        rk[pos] = {m.real.ref_index(m.up_str[m.up_off[p] + r], m.dn_str[m.dn_off[pd] + c]), pos};
      }
  }
  std::sort(rk.begin(), rk.end());
  m.ref_of_pos.assign((size_t)m.D, 0);
  m.pos_of_ref.assign((size_t)m.D, 0);
  for (int64_t i = 0; i < m.D; ++i) {
    m.ref_of_pos[rk[i].second] = i;
    m.pos_of_ref[i] = rk[i].second;
  }
  m.device = -1;
}

// ------------------------------------------------------------ operator

MomentumHubbard::MomentumHubbard(const MomentumSector &m, double t, double U) : m_(m) {
  const int N = m.L;
  std::vector<double> eps(N);
  for (int k = 0; k < N; ++k) {
    const int c[3] = {k % m.dims[0], (k / m.dims[0]) % m.dims[1], k / (m.dims[0] * m.dims[1])};
    double e = 0.0;
    for (int d = 0; d < m.ndim; ++d) e += -2.0 * std::cos(2.0 * kPi * c[d] / m.dims[d]);
    eps[k] = e;
  }
  auto eps_of = [&](uint32_t s) {
    double e = 0.0;
    for (int k = 0; k < N; ++k)
      if ((s >> k) & 1u) e += eps[k];
    return e;
  };
  // local index of a string inside its momentum group
  auto local = [&](const std::vector<uint32_t> &strs, const std::vector<int32_t> &off, uint32_t s) {
    const int p = m.momentum_of(s);
    auto f = std::lower_bound(strs.begin() + off[p], strs.begin() + off[p + 1], s);
    return (int)(f - (strs.begin() + off[p]));
  };
  auto between = [](uint32_t s, int a, int b) {
    const int lo = std::min(a, b), hi = std::max(a, b);
    const uint32_t mask = ((hi >= 32 ? 0xFFFFFFFFu : ((1u << hi) - 1u))) & ~((2u << lo) - 1u);
    return __builtin_popcount(s & mask);
  };
  // sources of an output string under c+_{k} c_{k - shift} summed over k, i.e.
  // the output string with one orbital k moved to k - shift (shift = q for
  // rho_{q,up}, -q for rho_{-q,dn})
  auto sources = [&](const std::vector<uint32_t> &strs, const std::vector<int32_t> &off, int width,
                     std::vector<uint16_t> &src, std::vector<uint8_t> &cnt, bool up) {
    const size_t ns = strs.size();
    src.assign(ns * N * width, 0);
    cnt.assign(ns * N, 0);
    for (size_t g = 0; g < ns; ++g) {
      const uint32_t o = strs[g];
      for (int q = 1; q < N; ++q) {
        int c = 0;
        for (int k = 0; k < N; ++k) {
          if (!((o >> k) & 1u)) continue;
          const int ks = up ? m.msub(k, q) : m.madd(k, q);  // orbital of the source
          if ((o >> ks) & 1u) continue;
          const uint32_t sstr = (o & ~(1u << k)) | (1u << ks);
          const int sign = between(o, k, ks) & 1;
          src[(g * N + q) * width + c] = (uint16_t)(local(strs, off, sstr) | (sign ? 0x8000 : 0));
          ++c;
        }
        cnt[g * N + q] = (uint8_t)c;
      }
    }
  };
  auto stride = [](int width) { return width <= 8 ? 8 : width; };
  const int mu = stride(std::max(1, std::min(m.n_up, N - m.n_up)));
  const int md = stride(std::max(1, std::min(m.n_dn, N - m.n_dn)));
  std::vector<uint16_t> up_src, dn_src;
  std::vector<uint8_t> up_cnt, dn_cnt;
  sources(m.up_str, m.up_off, mu, up_src, up_cnt, true);
  sources(m.dn_str, m.dn_off, md, dn_src, dn_cnt, false);
  std::vector<double> eu(m.up_str.size()), ed(m.dn_str.size());
  for (size_t g = 0; g < eu.size(); ++g) eu[g] = eps_of(m.up_str[g]);
  for (size_t g = 0; g < ed.size(); ++g) ed[g] = eps_of(m.dn_str[g]);
  std::vector<int32_t> grp(m.up_str.size()), loc(m.up_str.size());
  for (int p = 0; p < N; ++p)
    for (int g = m.up_off[p]; g < m.up_off[p + 1]; ++g) {
      grp[g] = p;
      loc[g] = g - m.up_off[p];
    }
  std::vector<int32_t> msub((size_t)N * N);
  for (int p = 0; p < N; ++p)
    for (int q = 0; q < N; ++q) msub[(size_t)p * N + q] = m.msub(p, q);
  d_eps_up_ = upload(eu);
  d_eps_dn_ = upload(ed);
  d_up_src_ = upload(up_src);
  d_dn_src_ = upload(dn_src);
  d_up_cnt_ = upload(up_cnt);
  d_dn_cnt_ = upload(dn_cnt);
  d_up_grp_ = upload(grp);
  d_up_loc_ = upload(loc);
  d_up_off_ = upload(m.up_off);
  d_dn_off_ = upload(m.dn_off);
  d_ncols_ = upload(m.ncols);
  d_msub_ = upload(msub);
  d_blk_off_ = upload(m.blk_off);
  tb_.N = N;
  tb_.K = m.K;
  tb_.mu = mu;
  tb_.md = md;
  tb_.D = m.D;
  tb_.nrows = (int64_t)m.up_str.size();
  tb_.t = t;
  tb_.U = U;
  tb_.q0 = U / N * (double)m.n_up * (double)m.n_dn;
  tb_.eps_up = d_eps_up_.as<double>();
  tb_.eps_dn = d_eps_dn_.as<double>();
  tb_.up_src = d_up_src_.as<uint16_t>();
  tb_.dn_src = d_dn_src_.as<uint16_t>();
  tb_.up_cnt = d_up_cnt_.as<uint8_t>();
  tb_.dn_cnt = d_dn_cnt_.as<uint8_t>();
  tb_.up_grp = d_up_grp_.as<int32_t>();
  tb_.up_loc = d_up_loc_.as<int32_t>();
  tb_.up_off = d_up_off_.as<int32_t>();
  tb_.dn_off = d_dn_off_.as<int32_t>();
  tb_.ncols = d_ncols_.as<int32_t>();
  tb_.msub = d_msub_.as<int32_t>();
  tb_.blk_off = d_blk_off_.as<int64_t>();
  tb_.max_cols = *std::max_element(m.ncols.begin(), m.ncols.end());
  tb_.max_rows = 0;
  for (int p = 0; p < N; ++p) tb_.max_rows = std::max(tb_.max_rows, m.up_off[p + 1] - m.up_off[p]);
}

int64_t MomentumHubbard::hv(const double *x, double *y, Epi e, const EpiArgs &a,
                            hipStream_t s) const {
  if (e == Epi::Cheb) throw std::invalid_argument("KPM is not available for momentum-sector operators");
  return launch_mom_hv(tb_, x, y, e, a, s);
}

int64_t MomentumHubbard::hv(const float *, float *, Epi, const EpiArgs &, hipStream_t) const {
  throw std::invalid_argument("FP32 H.v is not available for momentum-sector operators");
}

}  // namespace mb
