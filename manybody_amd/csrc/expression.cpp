// Generic fermionic expressions (FermionExpression, fermion/expression.h) compiled
// to bit operations on the basis words, for H.v on the device.
//
// A basis state is the word s = up | dn << L: mode m = spin L + orbital, the
// state c+_{m1} c+_{m2} ... |0> with m1 < m2 < ... (the reference's key order:
// up creators ascending, then down, basis.cpp:273-276).  c_m or c+_m on s gives
// (-1)^popcount(s & (2^m - 1)) and flips bit m, or zero.  Applied operator by
// operator, a monomial's action on any s depends only on the occupancy of the
// modes it touches, which it requires (in) and leaves (out); its sign is a
// constant times the parity of the occupied untouched modes below each of its
// operators, i.e. (-1)^popcount(s & parity) with parity the XOR of those masks.
// This is the same algebra as the reference's canonicalisation of A |s>
// (canonicalize.h:18-94, matrix_elements.h:74-105) with the output states
// outside the basis dropped.
//
// Compiled terms are split by the species they touch:
//   up only    H_up (x) 1: per up string, its source up strings and amplitudes
//              (merged); diagonal ones summed into a per-row value;
//   down only  1 (x) H_dn: per down string likewise;
//   mixed      grouped by their up part: per row the up parts whose condition
//              the row's up string meets (source row and sign), and per up part
//              its down parts, tested per column on the device.
// Terms that change N_up or N_dn (or, in a momentum sector, the momentum of a
// species part) map the basis outside itself and are dropped, as the reference
// drops output states not in basis.set.
#include "expression.hpp"

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>

namespace mb {

namespace {

constexpr double kTol = 1000.0 * 2.220446049250313e-16;  // tolerances::tolerance<double>()

bool tiny(std::complex<double> c) { return std::norm(c) < kTol * kTol; }

// MANYBODY_AMD_EXPR_NATIVE=0 keeps expressions of Hubbard form on the generic
// expression kernels
bool native_enabled() {
  const char *e = std::getenv("MANYBODY_AMD_EXPR_NATIVE");
  return !(e && std::strcmp(e, "0") == 0);
}

int parity64(uint64_t v) { return __builtin_popcountll(v) & 1; }

// ExpressionBase::add_to (expression_base.h:85-97): drop below 1000 eps, merge by
// key, erase a merged sum below 1000 eps
template <typename K>
void add_to(std::map<K, std::complex<double>> &m, const K &k, std::complex<double> c) {
  if (tiny(c)) return;
  auto it = m.find(k);
  if (it == m.end()) {
    m.emplace(k, c);
  } else {
    it->second += c;
    if (tiny(it->second)) m.erase(it);
  }
}

}  // namespace

bool compile_monomial(const uint16_t *ops, int n, int L, std::complex<double> c,
                      MonomialAction &a) {
  uint64_t touched = 0;
  for (int k = 0; k < n; ++k) {
    const int orb = ops[k] & kOpOrbital;
    if (orb >= L) throw std::invalid_argument("operator orbital outside the basis orbitals");
    touched |= uint64_t{1} << ((ops[k] & kOpDown ? L : 0) + orb);
  }
  // the first operator to act on a mode (the rightmost) fixes its required
  // occupancy: occupied for c, empty for c+
  uint64_t known = 0, in = 0;
  for (int k = n - 1; k >= 0; --k) {
    const uint64_t bit = uint64_t{1} << ((ops[k] & kOpDown ? L : 0) + (ops[k] & kOpOrbital));
    if (!(known & bit)) {
      known |= bit;
      if (ops[k] & kOpAnnihilation) in |= bit;
    }
  }
  uint64_t cur = in, par = 0;  // occupancy of the touched modes as the operators act
  bool neg = false;
  for (int k = n - 1; k >= 0; --k) {  // rightmost operator acts first
    const int m = (ops[k] & kOpDown ? L : 0) + (ops[k] & kOpOrbital);
    const uint64_t bit = uint64_t{1} << m, low = bit - 1;
    const bool ann = ops[k] & kOpAnnihilation;
    const bool occ = cur & bit;
    if (ann != occ) return false;  // c on an empty mode or c+ on an occupied one
    if (parity64(cur & low)) neg = !neg;  // touched modes below m: known occupancy
    par ^= low & ~touched;                // untouched modes below m: from the state
    cur ^= bit;
  }
  a.touched = touched;
  a.in = in;
  a.out = cur;
  a.parity = par;
  a.c = neg ? -c : c;
  return true;
}

template <typename V>
const typename V::value_type *ExpressionOperator::keep(const V &v) {
  bufs_.push_back(upload(v));
  return bufs_.back().template as<typename V::value_type>();
}

namespace {

// A layout in groups: rows are the up strings of group p (momentum, or particle
// number for a fixed-N basis), their columns the down strings of group
// col_group(p); the element offset of row r of group p is blk_off[p] + r ncols[p].
struct Grouped {
  int ngroups = 0;
  const std::vector<uint32_t> *up_str = nullptr, *dn_str = nullptr;
  const std::vector<int32_t> *up_off = nullptr, *dn_off = nullptr;
  const std::vector<int64_t> *blk_off = nullptr;
  const std::vector<int32_t> *ncols = nullptr;
  std::function<int(uint32_t)> group_of;
  std::function<int(int)> col_group;  // -1: no columns
};

}  // namespace

std::vector<MonomialAction> compile_expression(int64_t n_terms, const double *coeffs,
                                               const int64_t *offsets, const uint16_t *ops,
                                               int L) {
  if (n_terms < 0) throw std::invalid_argument("negative term count");
  if (n_terms > 0 && (!coeffs || !offsets || (offsets[n_terms] > 0 && !ops)))
    throw std::invalid_argument("null term arrays");
  std::map<std::vector<uint16_t>, std::complex<double>> expr;
  for (int64_t t = 0; t < n_terms; ++t) {
    if (offsets[t + 1] < offsets[t]) throw std::invalid_argument("term offsets must not decrease");
    std::vector<uint16_t> key(ops + offsets[t], ops + offsets[t + 1]);
    add_to(expr, key, {coeffs[2 * t], coeffs[2 * t + 1]});
  }
  std::vector<MonomialAction> out;
  for (const auto &[key, c] : expr) {
    MonomialAction a;
    if (compile_monomial(key.data(), (int)key.size(), L, c, a)) out.push_back(a);
  }
  return out;
}

ExpressionOperator::ExpressionOperator(const Sector *real, const MomentumSector *mom,
                                       int64_t n_terms, const double *coeffs,
                                       const int64_t *offsets, const uint16_t *ops,
                                       const FixedNBasis *fn)
    : real_(real), mom_(mom), fn_(fn) {
  if ((real != nullptr) + (mom != nullptr) + (fn != nullptr) != 1)
    throw std::invalid_argument("one basis required");
  if (n_terms < 0) throw std::invalid_argument("negative term count");
  const int L = real ? real->L : mom ? mom->L : fn->L;
  const int n_dn = real ? real->n_dn : mom ? mom->n_dn : 0;
  const uint64_t U = (uint64_t{1} << L) - 1;
  Grouped G;
  if (mom) {
    G.ngroups = mom->L;
    G.up_str = &mom->up_str;
    G.dn_str = &mom->dn_str;
    G.up_off = &mom->up_off;
    G.dn_off = &mom->dn_off;
    G.blk_off = &mom->blk_off;
    G.ncols = &mom->ncols;
    G.group_of = [mom](uint32_t b) { return mom->momentum_of(b); };
    G.col_group = [mom](int p) { return mom->msub(mom->K, p); };
  } else if (fn) {
    G.ngroups = fn->L + 1;
    G.up_str = &fn->up_str;
    G.dn_str = &fn->dn_str;
    G.up_off = &fn->up_off;
    G.dn_off = &fn->dn_off;
    G.blk_off = &fn->blk_off;
    G.ncols = &fn->ncols;
    G.group_of = [](uint32_t b) { return __builtin_popcount(b); };
    const int N = fn->N, Lf = fn->L;
    G.col_group = [N, Lf](int k) { return (N - k >= 0 && N - k <= Lf) ? N - k : -1; };
  }

  // the expression's own terms, merged by operator key (FermionExpression's map)
  std::map<std::vector<uint16_t>, std::complex<double>> expr;
  for (int64_t t = 0; t < n_terms; ++t) {
    if (offsets[t + 1] < offsets[t]) throw std::invalid_argument("term offsets must not decrease");
    std::vector<uint16_t> key(ops + offsets[t], ops + offsets[t + 1]);
    add_to(expr, key, {coeffs[2 * t], coeffs[2 * t + 1]});
  }

  // compiled actions by class
  std::vector<MonomialAction> up_diag, up_off, dn_diag, dn_off, mixed;
  for (const auto &[key, c] : expr) {
    MonomialAction a;
    if (!compile_monomial(key.data(), (int)key.size(), L, c, a)) continue;
    const uint64_t D_ = ~U;
    if (fn) {  // a fixed-N basis holds every N_up: only N must be kept
      if (__builtin_popcountll(a.in) != __builtin_popcountll(a.out)) continue;
    } else if (__builtin_popcountll(a.in & U) != __builtin_popcountll(a.out & U) ||
               __builtin_popcountll(a.in & D_) != __builtin_popcountll(a.out & D_)) {
      continue;  // changes N_up or N_dn: no output in the basis
    }
    ++kept_;
    const bool diag = a.in == a.out;
    if ((a.touched & ~U) == 0)
      (diag ? up_diag : up_off).push_back(a);
    else if ((a.touched & U) == 0)
      (diag ? dn_diag : dn_off).push_back(a);
    else
      mixed.push_back(a);
  }

  // rows (up strings) and columns (down strings) of the layout
  std::vector<uint32_t> rows_bits, dn_bits;
  std::vector<int64_t> row_base;
  std::vector<int32_t> row_nc, row_dg0;
  int64_t D = 0;
  if (real) {
    rows_bits = real->s_up;
    dn_bits = real->s_dn;
    D = real->dim();
    for (int64_t g = 0; g < real->d_up; ++g) {
      row_base.push_back(g * real->d_dn);
      row_nc.push_back((int32_t)real->d_dn);
      row_dg0.push_back(0);
    }
  } else {
    rows_bits = *G.up_str;
    dn_bits = *G.dn_str;
    D = mom ? mom->D : fn->D;
    for (int p = 0; p < G.ngroups; ++p)
      for (int g = (*G.up_off)[p]; g < (*G.up_off)[p + 1]; ++g) {
        const int nc = (*G.ncols)[p];
        const int cg = G.col_group(p);
        row_base.push_back((*G.blk_off)[p] + (int64_t)(g - (*G.up_off)[p]) * nc);
        row_nc.push_back(nc);
        row_dg0.push_back(cg >= 0 ? (*G.dn_off)[cg] : 0);
      }
  }
  const int64_t nrows = (int64_t)rows_bits.size(), ndn = (int64_t)dn_bits.size();

  // string -> (row base | column) lookups of the source side
  // (group id 0xFF: a string outside the down table, which matches no group)
  std::vector<int32_t> dn_loc, up_row;
  std::vector<uint8_t> dn_mom;
  if (!real) {
    if (L > 24) throw std::invalid_argument("grouped-basis expressions support at most 24 orbitals");
    const size_t n = size_t{1} << L;
    dn_loc.assign(n, -1);
    dn_mom.assign(n, 0xFF);
    up_row.assign(n, -1);
    for (int p = 0; p < G.ngroups; ++p) {
      for (int h = (*G.dn_off)[p]; h < (*G.dn_off)[p + 1]; ++h) {
        dn_loc[(*G.dn_str)[h]] = h - (*G.dn_off)[p];
        dn_mom[(*G.dn_str)[h]] = (uint8_t)p;
      }
    }
    for (int64_t g = 0; g < nrows; ++g) up_row[rows_bits[g]] = (int32_t)g;
  }
  auto source_row = [&](uint32_t u2, int32_t *grp) -> int64_t {  // row index or -1
    if (real) {
      *grp = -1;
      return Sector::colex_rank(u2);
    }
    const int32_t g2 = up_row[u2];
    if (g2 < 0) return -1;
    *grp = G.col_group(G.group_of(u2));  // group of the row's columns
    return (row_nc[g2] > 0 && *grp >= 0) ? g2 : -1;
  };
  auto source_col = [&](uint32_t d2, int32_t grp) -> int64_t {
    if (real) return Sector::colex_rank(d2);
    return (dn_mom[d2] == grp && dn_loc[d2] >= 0) ? dn_loc[d2] : -1;
  };

  // diagonal parts of one species
  std::vector<double> diag_up(2 * nrows, 0.0), diag_dn(2 * ndn, 0.0);
  for (int64_t g = 0; g < nrows; ++g) {
    std::complex<double> v = 0.0;
    for (const auto &a : up_diag)
      if (((uint64_t)rows_bits[g] & a.touched) == a.in) v += a.c;
    diag_up[2 * g] = v.real();
    diag_up[2 * g + 1] = v.imag();
  }
  for (int64_t h = 0; h < ndn; ++h) {
    std::complex<double> v = 0.0;
    const uint64_t w = (uint64_t)dn_bits[h] << L;
    for (const auto &a : dn_diag)
      if ((w & a.touched) == a.in) v += a.c;
    diag_dn[2 * h] = v.real();
    diag_dn[2 * h + 1] = v.imag();
  }

  // up-only off-diagonal: per row, source row base -> amplitude (gather form:
  // the output row must hold `out`, its source is row ^ (in ^ out))
  std::vector<int64_t> uo_ptr{0}, uo_base;
  std::vector<double> uo_amp;
  for (int64_t g = 0; g < nrows; ++g) {
    std::map<int64_t, std::complex<double>> acc;
    const uint64_t u = rows_bits[g];
    for (const auto &a : up_off) {
      if ((u & a.touched) != a.out) continue;
      const uint32_t u2 = (uint32_t)(u ^ (a.in ^ a.out));
      int32_t grp;
      const int64_t r2 = source_row(u2, &grp);
      if (r2 < 0) continue;
      if (!real && grp != G.col_group(G.group_of((uint32_t)u))) continue;
      acc[row_base[r2]] += parity64(u & a.parity) ? -a.c : a.c;
    }
    for (const auto &[b, c] : acc) {
      if (c == 0.0) continue;
      uo_base.push_back(b);
      uo_amp.push_back(c.real());
      uo_amp.push_back(c.imag());
    }
    uo_ptr.push_back((int64_t)uo_base.size());
  }
  // down-only off-diagonal: per down string, source column -> amplitude
  std::vector<int64_t> do_ptr{0};
  std::vector<int32_t> do_col;
  std::vector<double> do_amp;
  for (int64_t h = 0; h < ndn; ++h) {
    std::map<int64_t, std::complex<double>> acc;
    const uint64_t w = (uint64_t)dn_bits[h] << L;
    const int32_t grp = real ? -1 : G.group_of(dn_bits[h]);
    for (const auto &a : dn_off) {
      if ((w & a.touched) != a.out) continue;
      const uint32_t d2 = (uint32_t)((w ^ (a.in ^ a.out)) >> L);
      const int64_t c2 = source_col(d2, grp);
      if (c2 < 0) continue;
      acc[c2] += parity64(w & a.parity) ? -a.c : a.c;
    }
    for (const auto &[col, c] : acc) {
      if (c == 0.0) continue;
      do_col.push_back((int32_t)col);
      do_amp.push_back(c.real());
      do_amp.push_back(c.imag());
    }
    do_ptr.push_back((int64_t)do_col.size());
  }

  // mixed terms grouped by their up part; per group its down parts (merged)
  using UpKey = std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>;  // touched, in, out, parity
  using DnKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>;
  std::map<UpKey, std::map<DnKey, std::complex<double>>> groups;
  for (const auto &a : mixed) {
    const UpKey uk{a.touched & U, a.in & U, a.out & U, a.parity & U};
    const DnKey dk{(uint32_t)(a.touched >> L), (uint32_t)(a.in >> L), (uint32_t)(a.out >> L),
                   (uint32_t)(a.parity >> L)};
    groups[uk][dk] += a.c;
  }
  std::vector<UpKey> gkeys;
  std::vector<int32_t> dp_ptr{0};
  std::vector<uint32_t> dp_t, dp_o, dp_f, dp_p;
  std::vector<double> dp_c;
  for (const auto &[uk, parts] : groups) {
    gkeys.push_back(uk);
    for (const auto &[dk, c] : parts) {
      if (c == 0.0) continue;
      dp_t.push_back(std::get<0>(dk));
      dp_o.push_back(std::get<2>(dk));
      dp_f.push_back(std::get<1>(dk) ^ std::get<2>(dk));
      dp_p.push_back(std::get<3>(dk));
      dp_c.push_back(c.real());
      dp_c.push_back(c.imag());
    }
    dp_ptr.push_back((int32_t)dp_t.size());
  }
  std::vector<int64_t> mx_ptr{0}, mx_base;
  std::vector<int32_t> mx_grp, mx_dgrp;
  std::vector<double> mx_sign;
  for (int64_t g = 0; g < nrows; ++g) {
    const uint64_t u = rows_bits[g];
    for (size_t k = 0; k < gkeys.size(); ++k) {
      const auto &[t, in, out, par] = gkeys[k];
      if ((u & t) != out || dp_ptr[k + 1] == dp_ptr[k]) continue;
      int32_t grp;
      const int64_t r2 = source_row((uint32_t)(u ^ (in ^ out)), &grp);
      if (r2 < 0) continue;
      mx_grp.push_back((int32_t)k);
      mx_base.push_back(row_base[r2]);
      mx_dgrp.push_back(grp);
      mx_sign.push_back(parity64(u & par) ? -1.0 : 1.0);
    }
    mx_ptr.push_back((int64_t)mx_base.size());
  }

  // Hubbard form on a real-space sector: every hop, on-site value and mixed
  // amplitude real, and the mixed terms U n_{i up} n_{i dn} with one U on every
  // site.  Such an expression runs on the Hubbard tables and kernels (block or
  // strip up pass, sorted row pass; FP32 too) built from its own hop lists.
  if (real && native_enabled()) {
    bool ok = true;
    for (size_t k = 1; ok && k < uo_amp.size(); k += 2) ok = uo_amp[k] == 0.0;
    for (size_t k = 1; ok && k < do_amp.size(); k += 2) ok = do_amp[k] == 0.0;
    for (size_t k = 1; ok && k < diag_up.size(); k += 2) ok = diag_up[k] == 0.0;
    for (size_t k = 1; ok && k < diag_dn.size(); k += 2) ok = diag_dn[k] == 0.0;
    std::vector<double> u_site(L, 0.0);
    for (const auto &a : mixed) {
      if (!ok) break;
      const uint64_t up = a.touched & U, dn = a.touched >> L;
      ok = a.in == a.out && a.in == a.touched && a.parity == 0 && a.c.imag() == 0.0 &&
           __builtin_popcountll(up) == 1 && up == dn;
      if (ok) u_site[__builtin_ctzll(up)] += a.c.real();
    }
    for (int i = 1; ok && i < L; ++i) ok = u_site[i] == u_site[0];
    if (ok) {
      hub_ = std::make_unique<Hubbard>();
      Hubbard &h = *hub_;
      h.sec = real;
      h.U = L > 0 ? u_site[0] : 0.0;
      auto lists = [](SpeciesHops &sp, int64_t n, const std::vector<int64_t> &ptr,
                      const std::vector<double> &amp, const std::vector<double> &diag,
                      auto tgt_of) {
        sp.row_ptr.assign(1, 0);
        for (int64_t r = 0; r < n; ++r) {
          for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e) {
            sp.tgt.push_back(tgt_of(e));
            sp.amp.push_back(amp[2 * e]);
          }
          sp.row_ptr.push_back((int32_t)sp.tgt.size());
          sp.max_row = std::max(sp.max_row, (int)(ptr[r + 1] - ptr[r]));
          sp.onsite.push_back(diag[2 * r]);
        }
      };
      const int64_t ddn = real->d_dn;
      lists(h.up, nrows, uo_ptr, uo_amp, diag_up,
            [&](int64_t e) { return (int32_t)(uo_base[e] / ddn); });
      lists(h.dn, ndn, do_ptr, do_amp, diag_dn, [&](int64_t e) { return do_col[e]; });
      const_cast<Sector &>(*real).ensure_device();
      finish_hubbard(h);
      htb_ = tables_of(h);
    }
  }

  // work units: (row, chunk of columns)
  std::vector<int32_t> wg_row, wg_c0;
  for (int64_t g = 0; g < nrows; ++g)
    for (int c0 = 0; c0 < row_nc[g]; c0 += kExprChunk) {
      wg_row.push_back((int32_t)g);
      wg_c0.push_back(c0);
    }
  if (wg_row.size() >= (size_t{1} << 31)) throw std::invalid_argument("sector too large");

  auto nonempty = [](auto v) {
    if (v.empty()) v.resize(1);
    return v;
  };
  tb_.D = D;
  tb_.nrows = nrows;
  tb_.nwg = (int64_t)wg_row.size();
  tb_.L = L;
  tb_.n_dn = n_dn;
  tb_.momentum = real ? 0 : 1;  // grouped layout: source columns through the tables
  for (int64_t g = 0; g < nrows; ++g)
    tb_.max_uo = std::max<int>(tb_.max_uo, (int)(uo_ptr[g + 1] - uo_ptr[g]));
  if (real) {
    tb_.d_up = real->d_up;
    tb_.d_dn = real->d_dn;
  }
  tb_.wg_row = keep(nonempty(wg_row));
  tb_.wg_c0 = keep(nonempty(wg_c0));
  tb_.row_base = keep(nonempty(row_base));
  tb_.row_ncols = keep(nonempty(row_nc));
  tb_.row_dg0 = keep(nonempty(row_dg0));
  tb_.diag_up = keep(nonempty(diag_up));
  tb_.dn_bits = keep(nonempty(dn_bits));
  tb_.diag_dn = keep(nonempty(diag_dn));
  tb_.uo_ptr = keep(uo_ptr);
  tb_.uo_base = keep(nonempty(uo_base));
  tb_.uo_amp = keep(nonempty(uo_amp));
  // column-major copy of the down-only lists (padding bounded at 4x the entries
  // plus one column, and at most 255 entries a string)
  int max_do = 0;
  for (int64_t h = 0; h < ndn; ++h) max_do = std::max<int>(max_do, (int)(do_ptr[h + 1] - do_ptr[h]));
  tb_.max_do = max_do;
  tb_.n_dstr = ndn;
  if (max_do > 0 && max_do <= 255 &&
      (int64_t)max_do * ndn <= 4 * (int64_t)do_col.size() + ndn) {
    std::vector<int32_t> ec((size_t)max_do * ndn, 0);
    std::vector<double> ea((size_t)2 * max_do * ndn, 0.0);
    std::vector<uint8_t> cnt(ndn);
    for (int64_t h = 0; h < ndn; ++h) {
      const int64_t b = do_ptr[h];
      cnt[h] = (uint8_t)(do_ptr[h + 1] - b);
      for (int k = 0; k < cnt[h]; ++k) {
        ec[(size_t)k * ndn + h] = do_col[b + k];
        ea[2 * ((size_t)k * ndn + h)] = do_amp[2 * (b + k)];
        ea[2 * ((size_t)k * ndn + h) + 1] = do_amp[2 * (b + k) + 1];
      }
    }
    tb_.do_cnt = keep(cnt);
    std::map<std::pair<double, double>, uint32_t> amp_index;
    bool packable = ndn <= (int64_t{1} << 20);
    for (size_t e = 0; packable && e < do_col.size(); ++e) {
      const auto key = std::make_pair(do_amp[2 * e], do_amp[2 * e + 1]);
      if (!amp_index.count(key)) {
        const uint32_t id = (uint32_t)amp_index.size();
        if (id >= 4096) packable = false;
        amp_index.emplace(key, id);
      }
    }
    if (packable) {
      std::vector<uint32_t> pk((size_t)max_do * ndn, 0);
      std::vector<double> amps(2 * std::max<size_t>(amp_index.size(), 1), 0.0);
      for (const auto &[key, id] : amp_index) {
        amps[2 * id] = key.first;
        amps[2 * id + 1] = key.second;
      }
      for (size_t q = 0; q < pk.size(); ++q)
        pk[q] = ((uint32_t)ec[q] << 12) | amp_index[std::make_pair(ea[2 * q], ea[2 * q + 1])];
      tb_.do_pack = keep(pk);
      tb_.do_amps = keep(amps);
    } else {
      tb_.do_ell_col = keep(ec);
      tb_.do_ell_amp = keep(ea);
    }
  }
  tb_.do_ptr = keep(do_ptr);
  tb_.do_col = keep(nonempty(do_col));
  tb_.do_amp = keep(nonempty(do_amp));
  tb_.mx_ptr = keep(mx_ptr);
  tb_.mx_base = keep(nonempty(mx_base));
  tb_.mx_grp = keep(nonempty(mx_grp));
  tb_.mx_dgrp = keep(nonempty(mx_dgrp));
  tb_.mx_sign = keep(nonempty(mx_sign));
  tb_.dp_ptr = keep(dp_ptr);
  tb_.dp_touched = keep(nonempty(dp_t));
  tb_.dp_out = keep(nonempty(dp_o));
  tb_.dp_flip = keep(nonempty(dp_f));
  tb_.dp_parity = keep(nonempty(dp_p));
  tb_.dp_c = keep(nonempty(dp_c));
  if (real) {
    const Binom &B = binom();
    std::vector<uint64_t> bt((kMaxSites + 1) * (kMaxSites + 1));
    for (int n = 0; n <= kMaxSites; ++n)
      for (int k = 0; k <= kMaxSites; ++k) bt[n * (kMaxSites + 1) + k] = B.c[n][k];
    tb_.binom = keep(bt);
  } else {
    tb_.dn_loc = keep(dn_loc);
    tb_.dn_mom = keep(dn_mom);
  }
}

int64_t ExpressionOperator::hv(const double *x, double *y, Epi e, const EpiArgs &a,
                               hipStream_t s) const {
  if (hub_) return launch_hv<double>(htb_, 2, x, y, e, a, s);
  return launch_expr_hv(tb_, x, y, e, a, s);
}

int64_t ExpressionOperator::hv(const float *x, float *y, Epi e, const EpiArgs &a,
                               hipStream_t s) const {
  if (hub_) return launch_hv<float>(htb_, 2, x, y, e, a, s);
  throw std::invalid_argument(
      "FP32 H.v is available for expression operators of Hubbard form on real-space sectors only");
}

void ExpressionOperator::export_coo(std::vector<uint64_t> &rows, std::vector<uint64_t> &cols,
                                    std::vector<std::complex<double>> &vals,
                                    hipStream_t s) const {
  const int64_t D = tb_.D;
  DevBuf refmap((size_t)std::max<int64_t>(D, 1) * sizeof(int64_t));
  if (real_) {
    auto &sec = const_cast<Sector &>(*real_);
    sec.ensure_device();
    launch_ref_of_pos(perm_of(sec), refmap.as<int64_t>(), s);
  } else if (mom_) {
    auto &m = const_cast<MomentumSector &>(*mom_);
    m.ensure_device();
    MB_HIP(hipMemcpyAsync(refmap.p, m.d_ref_of_pos.p, (size_t)D * sizeof(int64_t),
                          hipMemcpyDeviceToDevice, s));
  } else {
    auto &f = const_cast<FixedNBasis &>(*fn_);
    f.ensure_device();
    MB_HIP(hipMemcpyAsync(refmap.p, f.d_ref_of_pos.p, (size_t)D * sizeof(int64_t),
                          hipMemcpyDeviceToDevice, s));
  }
  DevBuf keys, vbuf;
  const int64_t n = expr_export(tb_, refmap.as<int64_t>(), keys, vbuf, s);
  std::vector<uint64_t> k(n);
  std::vector<double> v(2 * n);
  if (n > 0) {
    MB_HIP(hipMemcpyAsync(k.data(), keys.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    MB_HIP(hipMemcpyAsync(v.data(), vbuf.p, (size_t)n * 16, hipMemcpyDeviceToHost, s));
  }
  MB_HIP(hipStreamSynchronize(s));
  rows.resize(n);
  cols.resize(n);
  vals.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    cols[i] = k[i] / (uint64_t)D;
    rows[i] = k[i] % (uint64_t)D;
    vals[i] = {v[2 * i], v[2 * i + 1]};
  }
}

}  // namespace mb
