// The reference's LinearOperator algebra (linear_operator.h:93-281) on device
// operators, see composite.cpp.
#pragma once

#include <complex>

#include "krylov.hpp"

namespace mb {

class CompositeOp final : public HvOp {
 public:
  // Negated -A, Scaled c A, Shifted A + c I, NegShift -A + c I, Sum A + B,
  // Difference A - B, Composed A B (B first), Identity I, Rescaled (A - b I) / a
  enum class Kind {
    Negated, Scaled, Shifted, NegShift, Sum, Difference, Composed, Identity, Rescaled
  };
  // a, b: operands (not owned; b only for Sum, Difference, Composed, neither for
  // Identity); c: the scalar of Scaled, Shifted, NegShift, (a, b) of Rescaled as
  // (real, imaginary) parts; dim: the dimension
  CompositeOp(Kind k, const HvOp *a, const HvOp *b, std::complex<double> c, int64_t dim);
  int64_t dim() const override { return dim_; }
  int64_t partials() const override { return 2 * combine_slots(dim_); }
  int64_t hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const override;
  int64_t hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const override;

 private:
  template <typename T>
  int64_t run(const T *x, T *y, Epi e, const EpiArgs &ea, hipStream_t s) const;
  Kind k_;
  const HvOp *a_, *b_;
  std::complex<double> c_;
  int64_t dim_;
};

}  // namespace mb
