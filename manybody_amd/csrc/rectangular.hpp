// Matrix elements between two bases (compute_rectangular_matrix_elements,
// matrix_elements.h:110-167) and vector elements (compute_vector_elements,
// matrix_elements.h:31-51), built on the device; see rectangular.cpp.
#pragma once

#include <complex>
#include <vector>

#include "fixed_n.hpp"
#include "mb_internal.hpp"
#include "momentum.hpp"

namespace mb {

// exactly one of the three is set
struct BasisRef {
  const Sector *real = nullptr;
  const MomentumSector *mom = nullptr;
  const FixedNBasis *fn = nullptr;
};

int64_t basis_size(const BasisRef &b);

// The nonzeros of <i|A|j>, i in `row`, j in `col` (reference indices), in
// arma::SpMat iteration order (column-major, rows ascending); contributions to
// one element merged and dropped below 1000 eps as add_to does.  Terms as for
// the expression operator (reference operator storage).
void rectangular_matrix_elements(const BasisRef &row, const BasisRef &col, int64_t n_terms,
                                 const double *coeffs, const int64_t *offsets,
                                 const uint16_t *ops, std::vector<uint64_t> &rows,
                                 std::vector<uint64_t> &cols,
                                 std::vector<std::complex<double>> &vals, hipStream_t s);

// <i|A|0> for every state i of the basis, in reference order, into out[size].
void vector_elements(const BasisRef &basis, int64_t n_terms, const double *coeffs,
                     const int64_t *offsets, const uint16_t *ops, std::complex<double> *out,
                     hipStream_t s);

}  // namespace mb
