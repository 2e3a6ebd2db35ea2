// Operator algebra on device operators.  Rescaled is rescale_hamiltonian
// (rescaling.h:40-48) on an operator: (A - b I) / a.
//
// Reference: linear_operator.h:93-281 wraps LinearOperators into Negated,
// Scaled, Shifted, NegShift, Sum, Difference, Composed and Identity (with the
// unary -, binary +, -, * and scalar * overloads); every consumer (Lanczos,
// eigenpairs, time evolution, Exp, power method, bounds) takes any of them.
// Here a composite calls its operands' H.v with the plain epilogue into
// stream-ordered scratch vectors and finishes with one combine kernel,
// h = alpha A(x) + beta B(x) + gamma x (or A(B(x)) for Composed), which also
// applies the caller's epilogue (Lanczos, reconstruction), so the Krylov layer
// runs on composites unchanged.  Scratch comes from hipMallocAsync on the
// caller's stream, so concurrent calls on one composite from several host
// threads (each on its own stream) stay independent.
#include "composite.hpp"

#include <stdexcept>

namespace mb {

namespace {

// stream-ordered scratch of n elements of T, freed on the same stream
template <typename T>
struct Scratch {
  T *p = nullptr;
  hipStream_t s;
  Scratch(size_t n, hipStream_t st) : s(st) {
    MB_HIP(hipMallocAsync(reinterpret_cast<void **>(&p), std::max<size_t>(n, 1) * sizeof(T), s));
  }
  ~Scratch() {
    if (p) (void)hipFreeAsync(p, s);
  }
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
};

}  // namespace

CompositeOp::CompositeOp(Kind k, const HvOp *a, const HvOp *b, std::complex<double> c,
                         int64_t dim)
    : k_(k), a_(a), b_(b), c_(c), dim_(dim) {
  const bool binary = k == Kind::Sum || k == Kind::Difference || k == Kind::Composed;
  if (k != Kind::Identity && !a) throw std::invalid_argument("null operand");
  if (binary && !b) throw std::invalid_argument("null operand");
  if (a && a->dim() != dim) throw std::invalid_argument("operand dimensions differ");
  if (binary && b->dim() != dim) throw std::invalid_argument("operand dimensions differ");
  if (dim < 0) throw std::invalid_argument("negative dimension");
}

template <typename T>
int64_t CompositeOp::run(const T *x, T *y, Epi e, const EpiArgs &ea, hipStream_t s) const {
  const size_t n = 2 * (size_t)dim_;  // planar complex
  CombineCoef c;
  if (k_ == Kind::Identity) {
    c.gr = 1.0;
    c.with_x = 1;
    return launch_combine<T>(dim_, nullptr, nullptr, x, y, c, e, ea, s);
  }
  Scratch<T> z1(n, s);
  if (k_ != Kind::Composed) a_->hv(x, z1.p, Epi::Plain, EpiArgs{}, s);
  switch (k_) {
    case Kind::Negated: c.ar = -1.0; break;
    case Kind::Scaled: c.ar = c_.real(); c.ai = c_.imag(); break;
    case Kind::Shifted: c.ar = 1.0; c.gr = c_.real(); c.gi = c_.imag(); c.with_x = 1; break;
    case Kind::NegShift: c.ar = -1.0; c.gr = c_.real(); c.gi = c_.imag(); c.with_x = 1; break;
    case Kind::Rescaled:  // (A x - b x) / a
      c.ar = 1.0 / c_.real();
      c.gr = -c_.imag() / c_.real();
      c.with_x = 1;
      break;
    case Kind::Sum:
    case Kind::Difference: {
      Scratch<T> z2(n, s);
      b_->hv(x, z2.p, Epi::Plain, EpiArgs{}, s);
      c.ar = 1.0;
      c.br = k_ == Kind::Sum ? 1.0 : -1.0;
      return launch_combine<T>(dim_, z1.p, z2.p, x, y, c, e, ea, s);
    }
    case Kind::Composed: {  // A (B x): B into z2, then A into z1
      Scratch<T> z2(n, s);
      b_->hv(x, z2.p, Epi::Plain, EpiArgs{}, s);
      a_->hv(z2.p, z1.p, Epi::Plain, EpiArgs{}, s);
      c.ar = 1.0;
      return launch_combine<T>(dim_, z1.p, nullptr, x, y, c, e, ea, s);
    }
    case Kind::Identity: break;
  }
  return launch_combine<T>(dim_, z1.p, nullptr, x, y, c, e, ea, s);
}

int64_t CompositeOp::hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const {
  return run<double>(x, y, e, a, s);
}
int64_t CompositeOp::hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const {
  return run<float>(x, y, e, a, s);
}

}  // namespace mb
