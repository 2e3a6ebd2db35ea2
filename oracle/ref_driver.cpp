// Checker-only driver around the reference's own algebra core (compiled from the
// reference sources where they lie; see oracle/Makefile).  It is NOT part of the
// product; it is built only where the reference tree exists, to write the small
// golden fixtures under tests/golden/ that pin the oracle's basis order, bond
// lists, fermionic signs and diagonal against the reference itself.
//
// Usage (all output is plain text on stdout):
//   ref_driver basis  L  Nup Ndn            -> "S up_bits dn_bits" per state, in index order
//   ref_driver h1d    L  Nup Ndn t U        -> states, then "E i j re im" per nonzero
//   ref_driver h2d    Lx Ly Nup Ndn t U
//   ref_driver h3d    Lx Ly Lz Nup Ndn t U
//   ref_driver mbasis ndim L.. Nup Ndn K..          -> fixed-momentum basis states
//   ref_driver mh     ndim L.. Nup Ndn K.. t U      -> states, then the momentum-space
//                                                      Hamiltonian's nonzeros
//   ref_driver mj     ndim L.. Nup Ndn Q.. dir t    -> states of the full (all-momentum)
//                                                      basis, then the current operator's
//                                                      nonzeros
//   ref_driver sigma  dt beta volume < "re im" lines -> "W omega re im" per frequency,
//                                                      then "R sum_rule"
//   ref_driver expr   L Nup Ndn < terms             -> states, then the nonzeros of the
//                                                      FermionExpression read from stdin
//   ref_driver mexpr  ndim L.. Nup Ndn K.. < terms  -> the same on a fixed-momentum basis
//   ref_driver nbasis L N                         -> states of with_fixed_particle_number
//   ref_driver nexpr  L N < terms                 -> the expression on that basis
//        terms: one per line, "re im n  type spin orbital  ..." (n operators,
//        type 0 creation / 1 annihilation, spin 0 up / 1 down), added in order
//   ref_driver cereal path < "rows cols" then "row col re im" lines
//                                                   -> the cereal binary archive of a
//                                                      sparse matrix, written to path
//
// Reference entry points exercised:
//   FermionBasis::with_fixed_particle_number_and_spin   src/algebra/basis.cpp:91-126
//   FermionBasis::with_fixed_particle_number            src/algebra/basis.cpp:70-89, 240-265
//   HubbardModel / HubbardModel2D / HubbardModel3D      src/algebra/fermion/model/hubbard_model.h:10-212
//   compute_matrix_elements<MatrixType>                  src/algebra/matrix_elements.h:74-105
//   FermionBasis::with_fixed_particle_number_spin_momentum src/algebra/basis.cpp:128-185, 304-356
//   HubbardModelMomentum::hamiltonian                    src/algebra/fermion/model/hubbard_model_momentum.h:119-175
//   HubbardModelMomentum::current                        src/algebra/fermion/model/hubbard_model_momentum.h:50-93
//   compute_optical_conductivity                         src/algorithms/optical_conductivity.cpp:8-80
//   FermionExpression / FermionMonomial (+=, add_to)     src/algebra/expression_base.h:85-120
//   cereal::BinaryOutputArchive (vendored third-party/cereal)
//
// The reference key holds at most 7 operators and silently drops longer
// products (src/algebra/fermion/term.h:11-19, src/algebra/expression_base.h:184-186),
// so the caller must keep Nup+Ndn <= 7 for "basis", and Nup+Ndn <= 3 for full
// Hamiltonians (<= 5 when U == 0).  ref_driver refuses larger sectors.
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "algebra/fermion/basis.h"
#include "algebra/fermion/model/hubbard_model.h"
#include "algebra/fermion/model/hubbard_model_momentum.h"
#include "utils/index.h"
#include "algorithms/optical_conductivity.h"
#include "algebra/matrix_elements.h"

#include <cereal/archives/binary.hpp>
#include <cereal/types/complex.hpp>
#include <fstream>

namespace {

// Minimal dense matrix satisfying what compute_matrix_elements needs:
// (rows, cols) construction, zeros() and operator()(i, j).
struct DenseComplex {
  size_t n_rows, n_cols;
  std::vector<std::complex<double>> a;
  DenseComplex(size_t r, size_t c) : n_rows(r), n_cols(c), a(r * c) {}
  void zeros() { std::fill(a.begin(), a.end(), std::complex<double>(0.0, 0.0)); }
  std::complex<double>& operator()(size_t i, size_t j) { return a[i * n_cols + j]; }
};

void print_states(const FermionBasis& basis) {
  for (size_t i = 0; i < basis.set.size(); ++i) {
    uint64_t up = 0, dn = 0;
    for (const auto& op : basis.set[i]) {
      if (op.spin() == FermionOperator::Spin::Up)
        up |= (uint64_t{1} << op.value());
      else
        dn |= (uint64_t{1} << op.value());
    }
    std::printf("S %llu %llu\n", (unsigned long long)up, (unsigned long long)dn);
  }
}

template <typename Expr>
void print_operator(const FermionBasis& basis, const Expr& A) {
  DenseComplex M = compute_matrix_elements<DenseComplex>(basis, A);
  for (size_t i = 0; i < M.n_rows; ++i)
    for (size_t j = 0; j < M.n_cols; ++j) {
      const auto v = M(i, j);
      if (v.real() != 0.0 || v.imag() != 0.0)
        std::printf("E %zu %zu %.17g %.17g\n", i, j, v.real(), v.imag());
    }
}

template <typename Model>
void print_matrix(const FermionBasis& basis, const Model& model) {
  print_operator(basis, model.hamiltonian());
}

int usage() {
  std::fprintf(stderr,
               "usage: ref_driver basis|h1d|h2d|h3d|g1d|g2d|g3d|mbasis|mh|mj|sigma|expr|mexpr|"
               "cereal|rexpr|vexpr ...\n");
  return 2;
}

// Terms from stdin into a FermionExpression, each added as the reference's
// FermionExpression(FermionMonomial) does (merge by key, drop below 1000 eps,
// drop keys longer than the monomial's capacity).  Returns the longest term.
size_t read_expression(FermionExpression& e) {
  double re, im;
  int n;
  size_t longest = 0;
  while (std::scanf("%lf %lf %d", &re, &im, &n) == 3) {
    FermionMonomial m(FermionMonomial::complex_type(re, im));
    for (int k = 0; k < n; ++k) {
      int type, spin, orb;
      if (std::scanf("%d %d %d", &type, &spin, &orb) != 3) return (size_t)-1;
      const auto sp = spin ? FermionOperator::Spin::Down : FermionOperator::Spin::Up;
      m.operators.push_back(type ? FermionOperator::annihilation(sp, (size_t)orb)
                                 : FermionOperator::creation(sp, (size_t)orb));
    }
    longest = std::max(longest, (size_t)n);
    e += FermionExpression(m);
  }
  return longest;
}

// A sparse matrix as the reference's arma_cereal.h save() writes it
// (src/utils/arma_cereal.h:9-16): a versioned non-member save emitting
// (n_rows, n_cols, n_nonzero) then (row, col, value) per nonzero in the
// column-major iteration order of arma::SpMat; arma::uword is 64-bit.
struct TripletMatrix {
  uint64_t n_rows = 0, n_cols = 0;
  std::vector<uint64_t> r, c;
  std::vector<std::complex<double>> v;
};

}  // namespace

namespace cereal {
template <class Archive>
void save(Archive& ar, const TripletMatrix& t, unsigned int /* version */) {
  ar(t.n_rows, t.n_cols, (uint64_t)t.v.size());
  for (size_t i = 0; i < t.v.size(); ++i) ar(t.r[i], t.c[i], t.v[i]);
}
}  // namespace cereal

int expr_mode(const std::string& mode, const std::vector<double>& a) {
  FermionBasis basis = [&] {
    if (mode == "nexpr" || mode == "nbasis") {
      if (a.size() != 2) throw std::invalid_argument("nexpr L N");
      return FermionBasis::with_fixed_particle_number((size_t)a[0], (size_t)a[1]);
    }
    if (mode == "expr") {
      if (a.size() != 3) throw std::invalid_argument("expr L Nup Ndn");
      const size_t n = (size_t)a[1] + (size_t)a[2];
      return FermionBasis::with_fixed_particle_number_and_spin((size_t)a[0], n,
                                                               (int)a[1] - (int)a[2]);
    }
    const size_t nd = a.empty() ? 0 : (size_t)a[0];
    if (nd < 1 || nd > 3 || a.size() != 1 + nd + 2 + nd) throw std::invalid_argument("mexpr");
    std::vector<size_t> dims, K;
    size_t sites = 1;
    for (size_t d = 0; d < nd; ++d) {
      dims.push_back((size_t)a[1 + d]);
      sites *= dims.back();
    }
    const size_t nup = (size_t)a[1 + nd], ndn = (size_t)a[2 + nd];
    for (size_t d = 0; d < nd; ++d) K.push_back((size_t)a[3 + nd + d]);
    return FermionBasis::with_fixed_particle_number_spin_momentum(
        sites, nup + ndn, (int)nup - (int)ndn, Index(dims), K);
  }();
  if (mode == "nbasis") {
    print_states(basis);
    return 0;
  }
  FermionExpression e;
  const size_t longest = read_expression(e);
  if (longest == (size_t)-1) return usage();
  const size_t n = basis.set.empty() ? 0 : basis.set[0].size();
  if (longest + n > 7) {
    std::fprintf(stderr, "terms times states exceed the 7-operator key\n");
    return 3;
  }
  print_states(basis);
  print_operator(basis, e);
  return 0;
}

// A basis from argv tokens: "s L nup ndn" (with_fixed_particle_number_and_spin),
// "n L N" (with_fixed_particle_number) or "m nd d_1 .. d_nd nup ndn K_1 .. K_nd"
// (with_fixed_particle_number_spin_momentum); advances `at`.
FermionBasis parse_basis(char** argv, int argc, int& at) {
  auto next = [&]() -> size_t {
    if (at >= argc) throw std::invalid_argument("basis spec too short");
    return (size_t)std::atol(argv[at++]);
  };
  if (at >= argc) throw std::invalid_argument("basis spec missing");
  const std::string kind = argv[at++];
  if (kind == "s") {
    const size_t L = next(), nu = next(), nd = next();
    return FermionBasis::with_fixed_particle_number_and_spin(L, nu + nd, (int)nu - (int)nd);
  }
  if (kind == "n") {
    const size_t L = next(), N = next();
    return FermionBasis::with_fixed_particle_number(L, N);
  }
  if (kind != "m") throw std::invalid_argument("basis kind");
  const size_t nd = next();
  std::vector<size_t> dims, K;
  size_t sites = 1;
  for (size_t d = 0; d < nd; ++d) {
    dims.push_back(next());
    sites *= dims.back();
  }
  const size_t nu = next(), ndn = next();
  for (size_t d = 0; d < nd; ++d) K.push_back(next());
  return FermionBasis::with_fixed_particle_number_spin_momentum(sites, nu + ndn,
                                                                (int)nu - (int)ndn, Index(dims), K);
}

struct DenseVector {
  std::vector<std::complex<double>> a;
  explicit DenseVector(size_t n) : a(n) {}
  void zeros() { std::fill(a.begin(), a.end(), std::complex<double>(0.0, 0.0)); }
  std::complex<double>& operator()(size_t i) { return a[i]; }
};

// rexpr <row basis> <col basis>: compute_rectangular_matrix_elements
// (matrix_elements.h:110-167); vexpr <basis>: compute_vector_elements
// (matrix_elements.h:31-51); terms on stdin.  Prints the row states (S), the
// column states (C) and the nonzero elements (E i j re im).
int rect_mode(int argc, char** argv) {
  const std::string mode = argv[1];
  int at = 2;
  FermionBasis row = parse_basis(argv, argc, at);
  FermionExpression e;
  const size_t longest = read_expression(e);
  if (longest == (size_t)-1) return usage();
  print_states(row);
  if (mode == "vexpr") {
    if (longest > 7) return 3;
    DenseVector v = compute_vector_elements<DenseVector>(row, e);
    for (size_t i = 0; i < v.a.size(); ++i)
      if (v.a[i].real() != 0.0 || v.a[i].imag() != 0.0)
        std::printf("E %zu 0 %.17g %.17g\n", i, v.a[i].real(), v.a[i].imag());
    return 0;
  }
  FermionBasis col = parse_basis(argv, argc, at);
  const size_t n = col.set.empty() ? 0 : col.set[0].size();
  if (longest + n > 7) {
    std::fprintf(stderr, "terms times states exceed the 7-operator key\n");
    return 3;
  }
  for (size_t i = 0; i < col.set.size(); ++i) {
    uint64_t up = 0, dn = 0;
    for (const auto& op : col.set[i])
      (op.spin() == FermionOperator::Spin::Up ? up : dn) |= (uint64_t{1} << op.value());
    std::printf("C %llu %llu\n", (unsigned long long)up, (unsigned long long)dn);
  }
  DenseComplex M = compute_rectangular_matrix_elements<DenseComplex>(row, col, e);
  for (size_t i = 0; i < M.n_rows; ++i)
    for (size_t j = 0; j < M.n_cols; ++j) {
      const auto v = M(i, j);
      if (v.real() != 0.0 || v.imag() != 0.0)
        std::printf("E %zu %zu %.17g %.17g\n", i, j, v.real(), v.imag());
    }
  return 0;
}

int cereal_mode(int argc, char** argv) {
  if (argc != 3) return usage();
  TripletMatrix t;
  unsigned long long r, c;
  if (std::scanf("%llu %llu", &r, &c) != 2) return usage();
  t.n_rows = r;
  t.n_cols = c;
  double re, im;
  while (std::scanf("%llu %llu %lf %lf", &r, &c, &re, &im) == 4) {
    t.r.push_back(r);
    t.c.push_back(c);
    t.v.emplace_back(re, im);
  }
  std::ofstream ofs(argv[2], std::ios::binary);
  cereal::BinaryOutputArchive archive(ofs);
  archive(t);
  return 0;
}

// mbasis / mh: ndim L.. Nup Ndn K.. [t U]
int momentum_mode(const std::string& mode, const std::vector<double>& a) {
  if (a.empty()) return usage();
  const size_t nd = (size_t)a[0];
  const size_t need = 1 + nd + 2 + nd + (mode == "mh" ? 2 : 0);
  if (nd < 1 || nd > 3 || a.size() != need) return usage();
  std::vector<size_t> dims, K;
  size_t sites = 1;
  for (size_t d = 0; d < nd; ++d) {
    dims.push_back((size_t)a[1 + d]);
    sites *= dims.back();
  }
  const size_t nup = (size_t)a[1 + nd], ndn = (size_t)a[2 + nd];
  for (size_t d = 0; d < nd; ++d) K.push_back((size_t)a[3 + nd + d]);
  const size_t n = nup + ndn;
  if (n > 7) { std::fprintf(stderr, "sector exceeds the 7-operator key\n"); return 3; }
  double t = 0, u = 0;
  if (mode == "mh") {
    t = a[3 + 2 * nd];
    u = a[4 + 2 * nd];
    // the interaction adds 4 operators to a state's key, the kinetic term 2
    if (n > (u == 0.0 ? 5u : 3u)) {
      std::fprintf(stderr, "sector exceeds the exact range of compute_matrix_elements\n");
      return 3;
    }
  }
  const Index index(dims);
  auto basis = FermionBasis::with_fixed_particle_number_spin_momentum(sites, n, (int)nup - (int)ndn,
                                                                     index, K);
  print_states(basis);
  if (mode == "mh") print_operator(basis, HubbardModelMomentum(t, u, dims).hamiltonian());
  return 0;
}

// mj: ndim L.. Nup Ndn Q.. dir t -- the current operator on the full basis
int current_mode(const std::vector<double>& a) {
  if (a.empty()) return usage();
  const size_t nd = (size_t)a[0];
  if (nd < 1 || nd > 3 || a.size() != 1 + nd + 2 + nd + 2) return usage();
  std::vector<size_t> dims, Q;
  size_t sites = 1;
  for (size_t d = 0; d < nd; ++d) {
    dims.push_back((size_t)a[1 + d]);
    sites *= dims.back();
  }
  const size_t nup = (size_t)a[1 + nd], ndn = (size_t)a[2 + nd];
  for (size_t d = 0; d < nd; ++d) Q.push_back((size_t)a[3 + nd + d]);
  const size_t dir = (size_t)a[3 + 2 * nd];
  const double t = a[4 + 2 * nd];
  // two operators per term: exact up to 5 particles
  if (nup + ndn > 5) { std::fprintf(stderr, "sector exceeds the exact range\n"); return 3; }
  auto basis = FermionBasis::with_fixed_particle_number_and_spin(sites, nup + ndn, (int)nup - (int)ndn);
  print_states(basis);
  print_operator(basis, HubbardModelMomentum(t, 0.0, dims).current(Q, dir));
  return 0;
}

// sigma: dt beta volume, correlation values "re im" on stdin
int sigma_mode(const std::vector<double>& a) {
  if (a.size() != 3) return usage();
  std::vector<std::complex<double>> c;
  double re, im;
  while (std::scanf("%lf %lf", &re, &im) == 2) c.emplace_back(re, im);
  const OpticalConductivityResult r = compute_optical_conductivity(c, a[0], a[1], a[2]);
  for (size_t i = 0; i < r.frequencies.size(); ++i)
    std::printf("W %.17g %.17g %.17g\n", r.frequencies[i], r.sigma[i].real(), r.sigma[i].imag());
  std::printf("R %.17g\n", r.sum_rule);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  const std::string mode = argv[1];
  std::vector<double> a;
  for (int i = 2; i < argc; ++i) a.push_back(std::atof(argv[i]));
  if (mode == "mbasis" || mode == "mh") return momentum_mode(mode, a);
  if (mode == "mj") return current_mode(a);
  if (mode == "sigma") return sigma_mode(a);
  if (mode == "expr" || mode == "mexpr" || mode == "nexpr" || mode == "nbasis")
    return expr_mode(mode, a);
  if (mode == "cereal") return cereal_mode(argc, argv);
  if (mode == "rexpr" || mode == "vexpr") return rect_mode(argc, argv);

  size_t sites = 0, nup = 0, ndn = 0;
  double t = 0, u = 0;
  if (mode == "basis" && a.size() == 3) {
    sites = (size_t)a[0]; nup = (size_t)a[1]; ndn = (size_t)a[2];
  } else if (mode == "h1d" && a.size() == 5) {
    sites = (size_t)a[0]; nup = (size_t)a[1]; ndn = (size_t)a[2]; t = a[3]; u = a[4];
  } else if (mode == "h2d" && a.size() == 6) {
    sites = (size_t)a[0] * (size_t)a[1]; nup = (size_t)a[2]; ndn = (size_t)a[3]; t = a[4]; u = a[5];
  } else if (mode == "h3d" && a.size() == 7) {
    sites = (size_t)a[0] * (size_t)a[1] * (size_t)a[2];
    nup = (size_t)a[3]; ndn = (size_t)a[4]; t = a[5]; u = a[6];
  } else if (mode == "g1d" && a.size() == 4) {  // L nup ndn r
    sites = (size_t)a[0]; nup = (size_t)a[1]; ndn = (size_t)a[2];
  } else if (mode == "g2d" && a.size() == 6) {  // Lx Ly nup ndn rx ry
    sites = (size_t)a[0] * (size_t)a[1]; nup = (size_t)a[2]; ndn = (size_t)a[3];
  } else if (mode == "g3d" && a.size() == 8) {  // Lx Ly Lz nup ndn rx ry rz
    sites = (size_t)a[0] * (size_t)a[1] * (size_t)a[2]; nup = (size_t)a[3]; ndn = (size_t)a[4];
  } else {
    return usage();
  }
  const size_t n = nup + ndn;
  if (n > 7) { std::fprintf(stderr, "sector exceeds the 7-operator key\n"); return 3; }
  const bool corr = mode[0] == 'g';  // density-density operator: the interaction's range
  if (mode != "basis" && n > ((u == 0.0 && !corr) ? 5u : 3u)) {
    std::fprintf(stderr, "sector exceeds the exact range of compute_matrix_elements\n");
    return 3;
  }

  auto basis = FermionBasis::with_fixed_particle_number_and_spin(sites, n, (int)nup - (int)ndn);
  print_states(basis);
  if (mode == "h1d") print_matrix(basis, HubbardModel(t, u, sites));
  if (mode == "h2d") print_matrix(basis, HubbardModel2D(t, u, (size_t)a[0], (size_t)a[1]));
  if (mode == "h3d")
    print_matrix(basis, HubbardModel3D(t, u, (size_t)a[0], (size_t)a[1], (size_t)a[2]));
  // G_updn(r) = (1/N) sum_i n_{i,up} n_{i+r,dn} (hubbard_model.h:40-53, 105-120, 184-200)
  if (mode == "g1d") print_operator(basis, HubbardModel(1.0, 0.0, sites).opposite_spin_correlation((size_t)a[3]));
  if (mode == "g2d")
    print_operator(basis, HubbardModel2D(1.0, 0.0, (size_t)a[0], (size_t)a[1])
                              .opposite_spin_correlation((size_t)a[4], (size_t)a[5]));
  if (mode == "g3d")
    print_operator(basis, HubbardModel3D(1.0, 0.0, (size_t)a[0], (size_t)a[1], (size_t)a[2])
                              .opposite_spin_correlation((size_t)a[5], (size_t)a[6], (size_t)a[7]));
  return 0;
}
