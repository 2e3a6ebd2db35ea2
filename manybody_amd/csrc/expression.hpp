// Generic fermionic expressions compiled to bit operations on the basis words
// (see expression.cpp).
#pragma once

#include <complex>
#include <memory>
#include <vector>

#include "fixed_n.hpp"
#include "krylov.hpp"
#include "mb_internal.hpp"
#include "momentum.hpp"

namespace mb {

// Reference operator storage (operator.h:14-40): bit 15 = annihilation, bit 14 =
// spin down, bits 0-13 = orbital.
constexpr uint16_t kOpAnnihilation = 0x8000u, kOpDown = 0x4000u, kOpOrbital = 0x3FFFu;

// The action of one monomial o_1 o_2 ... o_n (o_n applied first) on a basis word
// s = up | dn << L, the state c+_{up i1} ... c+_{dn j1} ... |0> with modes ascending
// (up before down, basis.cpp:273-276).  On s with (s & touched) == in it gives
// s ^ (in ^ out) times c (-1)^popcount(s & parity); otherwise zero.
struct MonomialAction {
  uint64_t touched = 0, in = 0, out = 0, parity = 0;
  std::complex<double> c;
};

// false if the monomial is identically zero (two equal creators or annihilators
// with nothing in between that restores the mode, ...)
bool compile_monomial(const uint16_t *ops, int n, int L, std::complex<double> c,
                      MonomialAction &a);

// Device tables of a compiled expression on a fixed-N basis (real-space product
// order or a fixed-momentum sector).  Work unit: (row, chunk of kExprChunk
// columns); row = up string, its columns = the down strings of its group.
struct ExprTables {
  int64_t D = 0, nrows = 0, nwg = 0;
  int L = 0, n_dn = 0, momentum = 0;  // momentum: a grouped layout (fixed K or fixed N)
  int max_uo = 0;                                     // most up-only sources of a row
  int64_t d_up = 0, d_dn = 0;                         // real space: the product shape
  const int32_t *wg_row = nullptr, *wg_c0 = nullptr;  // per workgroup
  const int64_t *row_base = nullptr;                 // element offset of column 0
  const int32_t *row_ncols = nullptr, *row_dg0 = nullptr;
  const double *diag_up = nullptr;                   // (re, im) per row
  const uint32_t *dn_bits = nullptr;                 // per down string (grouped)
  const double *diag_dn = nullptr;                   // (re, im) per down string
  const int64_t *uo_ptr = nullptr, *uo_base = nullptr;  // up-only: source row base
  const double *uo_amp = nullptr;
  const int64_t *do_ptr = nullptr;                   // down-only: source column
  const int32_t *do_col = nullptr;
  const double *do_amp = nullptr;
  // the down-only lists again, column-major (entry k of string h at k * n_dstr + h,
  // amplitudes interleaved re, im) so that lanes of consecutive strings load
  // consecutive words; null when padding would exceed the limit
  int max_do = 0;
  int64_t n_dstr = 0;
  const int32_t *do_ell_col = nullptr;
  const double *do_ell_amp = nullptr;
  const uint8_t *do_cnt = nullptr;
  // packed form of the same lists when the columns fit 20 bits and the
  // amplitudes take at most 4096 values: column << 12 | amplitude index
  const uint32_t *do_pack = nullptr;
  const double *do_amps = nullptr;  // (re, im) per amplitude index
  const int64_t *mx_ptr = nullptr, *mx_base = nullptr;  // mixed: per row, applicable up parts
  const int32_t *mx_grp = nullptr, *mx_dgrp = nullptr;
  const double *mx_sign = nullptr;
  const int32_t *dp_ptr = nullptr;                   // mixed: down parts of each up part
  const uint32_t *dp_touched = nullptr, *dp_out = nullptr, *dp_flip = nullptr,
                 *dp_parity = nullptr;
  const double *dp_c = nullptr;
  const uint64_t *binom = nullptr;                   // real space: colex ranks of down strings
  const int32_t *dn_loc = nullptr;                   // grouped: index in its group, by bits
  const uint8_t *dn_mom = nullptr;                   // grouped: group (momentum or particle
                                                     // number) by bits, 0xFF outside
};

constexpr int kExprChunk = 256;

// compute_matrix_elements(basis, A) (matrix_elements.h:74-105) as an operator:
// H.v, the Krylov layer and a sparse export run on the compiled form.
class ExpressionOperator final : public HvOp {
 public:
  // terms: n_terms coefficients (re, im interleaved), offsets[n_terms + 1] into
  // ops (reference operator storage); exactly one of real / mom / fn is set
  ExpressionOperator(const Sector *real, const MomentumSector *mom, int64_t n_terms,
                     const double *coeffs, const int64_t *offsets, const uint16_t *ops,
                     const FixedNBasis *fn = nullptr);
  int64_t dim() const override { return tb_.D; }
  int64_t partials() const override {
    return hub_ ? std::max<int64_t>(hv_partials(htb_, 2), 2 * tb_.nwg + 16) : 2 * tb_.nwg + 16;
  }
  int64_t hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const override;
  int64_t hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const override;
  const ExprTables &tables() const { return tb_; }
  // the Hubbard tables the operator runs on when its terms have Hubbard form
  // (real one-body hops per species, real on-site values, U n_up n_dn with
  // one U on every site; real-space sectors); nullptr otherwise
  const HvTables *hubbard_tables() const override { return hub_ ? &htb_ : nullptr; }
  const Hubbard *hubbard() const { return hub_.get(); }
  int64_t terms_kept() const { return kept_; }
  // the nonzeros in the reference's basis order, column-major with rows
  // ascending (arma::SpMat iteration order), entries below 1000 eps dropped
  void export_coo(std::vector<uint64_t> &rows, std::vector<uint64_t> &cols,
                  std::vector<std::complex<double>> &vals, hipStream_t s) const;

 private:
  const Sector *real_;
  const MomentumSector *mom_;
  const FixedNBasis *fn_;
  ExprTables tb_{};
  std::unique_ptr<Hubbard> hub_;
  HvTables htb_{};
  int64_t kept_ = 0;
  std::vector<DevBuf> bufs_;
  template <typename V>
  const typename V::value_type *keep(const V &v);
};

// The terms merged by operator key as FermionExpression does (add_to) and
// compiled on L orbitals; identically zero monomials are left out.
std::vector<MonomialAction> compile_expression(int64_t n_terms, const double *coeffs,
                                               const int64_t *offsets, const uint16_t *ops, int L);

// Device view of one basis for the rectangular export: rows of up strings, row
// g holding row_nc[g] elements from row_base[g] whose down strings are
// dn_bits[row_dg0[g] ...], all of group row_cgrp[g]; bit lookups up_row (row of
// an up string), dn_loc and dn_grp (index in its group and group of a down
// string; -1 outside); ref maps a device position to the reference index.
struct RectSide {
  int L = 0;
  int64_t D = 0, nrows = 0;
  int max_nc = 0;
  const uint32_t *row_up = nullptr;
  const int64_t *row_base = nullptr;
  const int32_t *row_nc = nullptr, *row_dg0 = nullptr, *row_cgrp = nullptr;
  const uint32_t *dn_bits = nullptr;
  const int32_t *up_row = nullptr, *dn_loc = nullptr, *dn_grp = nullptr;
  const int64_t *ref = nullptr;
};
struct RectMono {
  uint64_t touched, in, out, parity;
  double re, im;
};
// every (column element, monomial) whose output lies in the row basis emits
// key col_ref * row.D + row_ref; sorted and merged as expr_export
int64_t rect_export(const RectSide &row, const RectSide &col, const RectMono *monos, int n_mono,
                    DevBuf &keys, DevBuf &vals, hipStream_t s);

// kernels (kernels.hip, expression.hip)
int64_t launch_expr_hv(const ExprTables &t, const double *x, double *y, Epi epi, const EpiArgs &a,
                       hipStream_t s);
void launch_ref_of_pos(const PermTables &p, int64_t *out, hipStream_t s);
// raw (key, value) contributions of every element, key = col_ref * D + row_ref;
// sorted and merged by key; returns the number of merged entries kept
int64_t expr_export(const ExprTables &t, const int64_t *ref_of_pos, DevBuf &keys, DevBuf &vals,
                    hipStream_t s);

}  // namespace mb
