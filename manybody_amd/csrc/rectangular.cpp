// Matrix elements of an operator between two bases, and vector elements.
//
// Reference: compute_rectangular_matrix_elements (matrix_elements.h:110-167)
// canonicalises A |j> for every column state j of the column basis and keeps
// the terms that are states of the row basis; compute_vector_elements
// (matrix_elements.h:31-51) takes, for every basis state i, the constant term
// of canonicalize(adjoint(<i|) A), which is <i|A|0>: the same export with the
// vacuum as the one column state.  Fermion state normalisation is 1
// (basis.h:32-34).
//
// Here the expression is compiled to bit operations (compile_expression) and
// both bases are described as rows of up strings with their down strings and
// bit lookups (RectSide); the device applies every monomial to every column
// state, looks the output up in the row basis and emits (column, row, value)
// keyed in the reference order of both bases, then sorts and merges as the
// square export does.  Any pair of the supported kinds works: a fixed
// (N_up, N_dn) sector, a fixed-momentum sector, a fixed-N basis, and operators
// that change the particle numbers (c+, c, pair creation).
#include "rectangular.hpp"

#include <stdexcept>

#include "expression.hpp"

namespace mb {

namespace {

// Host form of RectSide plus the device buffers it points at.
struct Layout {
  int L = 0;
  int64_t D = 0;
  std::vector<uint32_t> row_up, dn_bits;
  std::vector<int64_t> row_base;
  std::vector<int32_t> row_nc, row_dg0, row_cgrp, up_row, dn_loc, dn_grp;
  DevBuf d_row_up, d_row_base, d_row_nc, d_row_dg0, d_row_cgrp, d_dn_bits, d_up_row, d_dn_loc,
      d_dn_grp, d_ref;
  RectSide side{};
};

void tables(Layout &y, int L) {
  if (L > 24) throw std::invalid_argument("rectangular matrix elements support at most 24 orbitals");
  y.L = L;
  const size_t n = size_t{1} << L;
  y.up_row.assign(n, -1);
  y.dn_loc.assign(n, -1);
  y.dn_grp.assign(n, -1);
}

// a fixed (N_up, N_dn) sector: one group; rows = up strings by colex rank
void layout_real(const Sector &sec, Layout &y) {
  tables(y, sec.L);
  y.D = sec.dim();
  y.dn_bits = sec.s_dn;
  for (int64_t h = 0; h < sec.d_dn; ++h) {
    y.dn_loc[sec.s_dn[h]] = (int32_t)h;
    y.dn_grp[sec.s_dn[h]] = 0;
  }
  for (int64_t g = 0; g < sec.d_up; ++g) {
    y.up_row[sec.s_up[g]] = (int32_t)g;
    y.row_up.push_back(sec.s_up[g]);
    y.row_base.push_back(g * sec.d_dn);
    y.row_nc.push_back((int32_t)sec.d_dn);
    y.row_dg0.push_back(0);
    y.row_cgrp.push_back(0);
  }
}

// a grouped basis: rows of group p hold the down strings of group col_group(p)
template <typename ColGroup>
void layout_grouped(int L, int64_t D, int ngroups, const std::vector<uint32_t> &up_str,
                    const std::vector<uint32_t> &dn_str, const std::vector<int32_t> &up_off,
                    const std::vector<int32_t> &dn_off, const std::vector<int64_t> &blk_off,
                    const std::vector<int32_t> &ncols, ColGroup col_group, Layout &y) {
  tables(y, L);
  y.D = D;
  y.dn_bits = dn_str;
  for (int p = 0; p < ngroups; ++p)
    for (int h = dn_off[p]; h < dn_off[p + 1]; ++h) {
      y.dn_loc[dn_str[h]] = h - dn_off[p];
      y.dn_grp[dn_str[h]] = p;
    }
  for (int p = 0; p < ngroups; ++p) {
    const int cg = col_group(p);
    for (int g = up_off[p]; g < up_off[p + 1]; ++g) {
      if (cg < 0 || ncols[p] == 0) continue;
      y.up_row[up_str[g]] = (int32_t)y.row_up.size();
      y.row_up.push_back(up_str[g]);
      y.row_base.push_back(blk_off[p] + (int64_t)(g - up_off[p]) * ncols[p]);
      y.row_nc.push_back(ncols[p]);
      y.row_dg0.push_back(dn_off[cg]);
      y.row_cgrp.push_back(cg);
    }
  }
}

void layout_of(const BasisRef &b, Layout &y, hipStream_t s) {
  if ((b.real != nullptr) + (b.mom != nullptr) + (b.fn != nullptr) != 1)
    throw std::invalid_argument("one basis required");
  if (b.real) {
    layout_real(*b.real, y);
  } else if (b.mom) {
    const MomentumSector &m = *b.mom;
    layout_grouped(m.L, m.D, m.L, m.up_str, m.dn_str, m.up_off, m.dn_off, m.blk_off, m.ncols,
                   [&m](int p) { return m.msub(m.K, p); }, y);
  } else {
    const FixedNBasis &f = *b.fn;
    const int N = f.N, L = f.L;
    layout_grouped(f.L, f.D, f.L + 1, f.up_str, f.dn_str, f.up_off, f.dn_off, f.blk_off,
                   f.ncols, [N, L](int k) { return (N - k >= 0 && N - k <= L) ? N - k : -1; }, y);
  }
  y.d_row_up = upload(y.row_up);
  y.d_row_base = upload(y.row_base);
  y.d_row_nc = upload(y.row_nc);
  y.d_row_dg0 = upload(y.row_dg0);
  y.d_row_cgrp = upload(y.row_cgrp);
  y.d_dn_bits = upload(y.dn_bits);
  y.d_up_row = upload(y.up_row);
  y.d_dn_loc = upload(y.dn_loc);
  y.d_dn_grp = upload(y.dn_grp);
  y.d_ref.alloc((size_t)std::max<int64_t>(y.D, 1) * sizeof(int64_t));
  if (b.real) {
    auto &sec = const_cast<Sector &>(*b.real);
    sec.ensure_device();
    launch_ref_of_pos(perm_of(sec), y.d_ref.as<int64_t>(), s);
  } else if (b.mom) {
    auto &m = const_cast<MomentumSector &>(*b.mom);
    m.ensure_device();
    MB_HIP(hipMemcpyAsync(y.d_ref.p, m.d_ref_of_pos.p, (size_t)y.D * sizeof(int64_t),
                          hipMemcpyDeviceToDevice, s));
  } else {
    auto &f = const_cast<FixedNBasis &>(*b.fn);
    f.ensure_device();
    MB_HIP(hipMemcpyAsync(y.d_ref.p, f.d_ref_of_pos.p, (size_t)y.D * sizeof(int64_t),
                          hipMemcpyDeviceToDevice, s));
  }
  RectSide &r = y.side;
  r.L = y.L;
  r.D = y.D;
  r.nrows = (int64_t)y.row_up.size();
  for (int32_t nc : y.row_nc) r.max_nc = std::max(r.max_nc, nc);
  r.row_up = y.d_row_up.as<uint32_t>();
  r.row_base = y.d_row_base.as<int64_t>();
  r.row_nc = y.d_row_nc.as<int32_t>();
  r.row_dg0 = y.d_row_dg0.as<int32_t>();
  r.row_cgrp = y.d_row_cgrp.as<int32_t>();
  r.dn_bits = y.d_dn_bits.as<uint32_t>();
  r.up_row = y.d_up_row.as<int32_t>();
  r.dn_loc = y.d_dn_loc.as<int32_t>();
  r.dn_grp = y.d_dn_grp.as<int32_t>();
  r.ref = y.d_ref.as<int64_t>();
}

int orbitals_of(const BasisRef &b) {
  return b.real ? b.real->L : b.mom ? b.mom->L : b.fn ? b.fn->L : 0;
}

}  // namespace

int64_t basis_size(const BasisRef &b) {
  return b.real ? b.real->dim() : b.mom ? b.mom->D : b.fn ? b.fn->D : 0;
}

void rectangular_matrix_elements(const BasisRef &row, const BasisRef &col, int64_t n_terms,
                                 const double *coeffs, const int64_t *offsets,
                                 const uint16_t *ops, std::vector<uint64_t> &rows,
                                 std::vector<uint64_t> &cols,
                                 std::vector<std::complex<double>> &vals, hipStream_t s) {
  const int L = orbitals_of(row);
  if (L != orbitals_of(col)) throw std::invalid_argument("the two bases have different orbitals");
  const std::vector<MonomialAction> acts = compile_expression(n_terms, coeffs, offsets, ops, L);
  rows.clear();
  cols.clear();
  vals.clear();
  const int64_t Dr = basis_size(row), Dc = basis_size(col);
  if (acts.empty() || Dr == 0 || Dc == 0) return;
  if ((double)Dr * (double)Dc >= 9.2e18) throw std::invalid_argument("bases too large to export");
  std::vector<RectMono> mono(acts.size());
  for (size_t k = 0; k < acts.size(); ++k)
    mono[k] = {acts[k].touched, acts[k].in, acts[k].out, acts[k].parity, acts[k].c.real(),
               acts[k].c.imag()};
  Layout yr, yc;
  layout_of(row, yr, s);
  layout_of(col, yc, s);
  DevBuf d_mono = upload(mono);
  DevBuf keys, vbuf;
  const int64_t n = rect_export(yr.side, yc.side, d_mono.as<RectMono>(), (int)mono.size(), keys,
                                vbuf, s);
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
    cols[i] = k[i] / (uint64_t)Dr;
    rows[i] = k[i] % (uint64_t)Dr;
    vals[i] = {v[2 * i], v[2 * i + 1]};
  }
}

void vector_elements(const BasisRef &basis, int64_t n_terms, const double *coeffs,
                     const int64_t *offsets, const uint16_t *ops, std::complex<double> *out,
                     hipStream_t s) {
  Sector vacuum;
  build_sector(vacuum, orbitals_of(basis), 0, 0);
  BasisRef col;
  col.real = &vacuum;
  std::vector<uint64_t> r, c;
  std::vector<std::complex<double>> v;
  rectangular_matrix_elements(basis, col, n_terms, coeffs, offsets, ops, r, c, v, s);
  const int64_t D = basis_size(basis);
  std::fill(out, out + D, std::complex<double>(0.0, 0.0));
  for (size_t i = 0; i < v.size(); ++i) out[r[i]] = v[i];
}

}  // namespace mb
