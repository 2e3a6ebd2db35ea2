// C++ host API test (include/manybody_amd.hpp), written the way the
// reference's Catch2 tests read (tests/numerics/test_lanczos.cpp,
// tests/algebra/test_basis.cpp).  Run on a GPU by tests/test_gpu_cpp_host.py;
// prints one line per case and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "manybody_amd.hpp"

static int g_fail = 0;
#define CHECK(cond)                                                         \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++g_fail;                                                             \
    }                                                                       \
  } while (0)
#define CASE(name) std::printf("case %s\n", name); std::fflush(stdout);

int main() {
  using namespace mbamd;
  {
    CASE("basis_order_small_example");
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(2, 2, 0);
    CHECK(basis.size() == 4);
    CHECK((basis.state(0) == std::pair<uint64_t, uint64_t>{1, 1}));
    CHECK((basis.state(3) == std::pair<uint64_t, uint64_t>{2, 1}));
    CHECK(basis.index_of(1, 2) == 1);
    CHECK(!basis.contains(3, 1));
  }
  {
    CASE("basis_errors_are_exceptions");
    bool thrown = false;
    try {
      (void)FermionBasis::with_fixed_particle_number_and_spin(4, 3, 0);
    } catch (const std::invalid_argument &) {
      thrown = true;
    }
    CHECK(thrown);
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(4, 3, 1);
    thrown = false;
    try {
      (void)basis.state(basis.size());
    } catch (const std::out_of_range &) {
      thrown = true;
    }
    CHECK(thrown);
  }
  {
    CASE("ground_state_c1_u0_closed_form");
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(8, 8, 0);
    auto H = HubbardModel(1.0, 0.0, 8).op(basis);
    CHECK(H.dimension() == 4900);
    auto b = H.gaussian(11);
    auto ep = find_min_eigenpair(H, 200, b);
    std::printf("  E0 = %.15f\n", ep.value);
    CHECK(std::abs(ep.value + 9.656854249492380) < 1e-9);
    // residual of the rebuilt eigenvector
    auto Hx = H.apply(ep.vector);
    auto x = ep.vector.to_host(), hx = Hx.to_host();
    double r = 0, n = 0;
    for (size_t i = 0; i < x.size(); ++i) {
      r += std::norm(hx[i] - ep.value * x[i]);
      n += std::norm(x[i]);
    }
    CHECK(std::sqrt(r / n) < 1e-6);
    // the same on the device, and the correlation sum rule sum_r G(r) = N_up N_dn / N
    auto er = H.energy_residual(ep.vector);
    CHECK(std::abs(er.first - ep.value) < 1e-10 * std::abs(ep.value));
    CHECK(std::abs(er.second - std::sqrt(r / n)) < 1e-8);
    auto g = H.opposite_spin_correlation(ep.vector);
    double gs = 0;
    for (double v : g) gs += v;
    CHECK(g.size() == 8 && std::abs(gs - 2.0) < 1e-12);
  }
  {
    CASE("momentum_sector_ground_state");
    // 1D ring L = 8, 4 + 4, U = 4: the lowest energy over the momentum sectors is
    // the real-space ground energy
    auto rbasis = FermionBasis::with_fixed_particle_number_and_spin(8, 8, 0);
    auto R = HubbardModel(1.0, 4.0, 8).op(rbasis);
    const double e_real = find_min_eigenpair(R, 200, R.gaussian(5)).value;
    double e_min = 1e300;
    size_t dims = 0;
    for (int K = 0; K < 8; ++K) {
      auto mbasis = FermionBasis::with_fixed_particle_number_spin_momentum(8, 8, 0, {8}, {K});
      auto M = HubbardMomentumFactorizedFixedK(mbasis, {8}, 1.0, 4.0, {K});
      dims += M.dimension();
      CHECK(mbasis.size() == M.dimension());
      auto st = mbasis.state(0);
      CHECK(mbasis.index_of(st.first, st.second) == 0);
      const int k = (int)std::min<size_t>(200, M.dimension());
      e_min = std::min(e_min, find_min_eigenpair(M, k, M.gaussian(7 + K)).value);
    }
    CHECK(dims == 4900);
    CHECK(std::abs(e_min - e_real) < 1e-10 * std::abs(e_real));
    bool thrown = false;
    try {
      auto bad = FermionBasis::with_fixed_particle_number_spin_momentum(8, 8, 0, {8}, {0});
      (void)HubbardMomentumFactorizedFixedK(bad, {8}, 1.0, 4.0, {3});
    } catch (const std::invalid_argument &) {
      thrown = true;
    }
    CHECK(thrown);
  }
  {
    CASE("lanczos_decomposition_shape");
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(6, 5, 1);
    auto H = HubbardModel2D(1.0, 4.0, 3, 2).op(basis);
    auto b = H.gaussian(3);
    auto d = lanczos_pass_one(H, b, 25);
    CHECK(d.steps_taken == 25);
    CHECK(d.alphas.size() == 25 && d.betas.size() == 24);
    for (double beta : d.betas) CHECK(beta > 0.0);
  }
  {
    CASE("time_evolution_conserves_norm");
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(8, 6, 0);
    auto H = HubbardModel(1.0, 4.0, 8).op(basis);
    auto psi = H.gaussian(5);
    auto out = time_evolve_state_steps(H, psi, 0.0, 0.5, 0.1, EvolutionOptions{30});
    auto a = psi.to_host(), c = out.to_host();
    double na = 0, nc = 0;
    for (size_t i = 0; i < a.size(); ++i) { na += std::norm(a[i]); nc += std::norm(c[i]); }
    CHECK(std::abs(std::sqrt(nc) - std::sqrt(na)) / std::sqrt(na) < 1e-10);
  }
  {
    CASE("kpm_moments_and_kernel");
    auto g = kpm::jackson_kernel(50);
    CHECK(std::abs(g[0] - 1.0) < 1e-10);
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(8, 6, 0);
    auto H = HubbardModel(1.0, 4.0, 8).op(basis);
    // Gershgorin: at most 12 hops of size t per state, diagonal in [0, 3U]
    auto r = rescaling::from_bounds(-12.0, 24.0, 0.05);
    auto mu = kpm::dos_moments(H, r.a, r.b, 64, 4, 1, true);
    CHECK(std::abs(mu[0] - 1.0) < 1e-6);
    for (double m : mu) CHECK(std::abs(m) <= 1.0 + 1e-6);
  }
  {  // reorthogonalised Lanczos and FP32 time evolution through the C++ layer
    auto basis = mbamd::FermionBasis::with_fixed_particle_number_and_spin(8, 8, 0);
    auto H = mbamd::HubbardModel(1.0, 0.0, 8).op(basis);
    auto b = H.gaussian(7);
    auto ep = mbamd::find_min_eigenpair(H, 120, b, nullptr, /*reorthogonalize=*/true);
    CHECK(std::abs(ep.value + 9.656854249492380) < 1e-9);
    auto d = mbamd::lanczos_pass_one(H, b, 60, nullptr, true);
    CHECK(d.steps_taken == 60);
    mbamd::EvolutionOptions o;
    o.fp32 = true;
    auto psi = mbamd::time_evolve_state(H, b, 0.5, o);
    auto h64 = b.to_host(), h32 = psi.to_host();
    double n0 = 0.0, n1 = 0.0;
    for (size_t i = 0; i < h64.size(); ++i) {
      n0 += std::norm(h64[i]);
      n1 += std::norm(h32[i]);
    }
    CHECK(std::abs(std::sqrt(n1) - std::sqrt(n0)) / std::sqrt(n0) < 1e-6);
  }
  {  // serving loop: a ground-state request through the C++ layer
    mbamd::Server srv(2, int64_t(1) << 30, int64_t(256) << 20);
    mb_request r{};
    r.kind = MB_REQ_GROUND;
    r.ndim = 1;
    r.sizes[0] = 8;
    r.n_up = 4;
    r.n_dn = 4;
    r.t = 1.0;
    r.u = 0.0;
    r.steps = 150;
    r.seed = 3;
    r.arrival = -1.0;
    const mb_result res = srv.wait(srv.submit(r));
    CHECK(res.status == MB_OK);
    CHECK(std::abs(res.e0 + 9.656854249492380) < 1e-9);
    CHECK(res.dim == 4900);
  }
  {
    CASE("optical_conductivity_pipeline");
    // ring L = 6, 3 + 2, K = 1 -> K + Q = 2: <y|J x> = <J^dagger y|x>, C(0) real
    // and positive, and sigma(omega) on pi/dt / (eta/4) + 1 frequencies
    auto bk = FermionBasis::with_fixed_particle_number_spin_momentum(6, 5, 1, {6}, {1});
    auto bq = FermionBasis::with_fixed_particle_number_spin_momentum(6, 5, 1, {6}, {2});
    auto HK = HubbardMomentumFactorizedFixedK(bk, {6}, 1.0, 4.0, {1});
    auto HQ = HubbardMomentumFactorizedFixedK(bq, {6}, 1.0, 4.0, {2});
    mbamd::CurrentOperatorFixedK J(bk, bq, {6}, 1.0, 1, 0);
    CHECK(J.source_dimension() == HK.dimension() && J.target_dimension() == HQ.dimension());
    auto x = HK.gaussian(1), y = HQ.gaussian(2);
    auto jx = J.apply(x).to_host(), jty = J.adjoint_apply(y).to_host();
    auto xh = x.to_host(), yh = y.to_host();
    std::complex<double> lhs = 0.0, rhs = 0.0;
    for (size_t i = 0; i < yh.size(); ++i) lhs += std::conj(yh[i]) * jx[i];
    for (size_t i = 0; i < xh.size(); ++i) rhs += std::conj(jty[i]) * xh[i];
    CHECK(std::abs(lhs - rhs) < 1e-12 * std::abs(lhs));
    auto C = mbamd::optical_correlator(HK, HQ, J, 2, 1.0, 0.1, 21, 5);
    CHECK(C.size() == 21 && C[0].real() > 0.0 && std::abs(C[0].imag()) < 1e-12 * C[0].real());
    auto r = mbamd::compute_optical_conductivity(C, 0.1, 1.0, 6.0);
    CHECK(r.frequencies.size() == 63 && r.sigma.size() == 63);  // eta = 4 / 2: floor(10 pi / 0.5) + 1
    auto z = mbamd::imaginary_time_evolve_state(HK, x, 0.0);
    auto zh = z.to_host();
    double dz = 0.0;
    for (size_t i = 0; i < zh.size(); ++i) dz = std::max(dz, std::abs(zh[i] - xh[i]));
    CHECK(dz < 1e-12);  // exp(0) = 1
  }

  {
    CASE("expression_operator_and_sparse_export");
    // HubbardModel(1, 4, 6).hamiltonian() written as terms: H.v equals the native
    // operator, the exported matrix is Hermitian and survives a cereal round trip
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(6, 5, 1);
    auto H = HubbardModel(1.0, 4.0, 6).op(basis);
    std::vector<mbamd::ExpressionTerm> terms;
    auto op = [](bool ann, int spin, int orb) {
      return (uint16_t)((ann ? 0x8000 : 0) | (spin ? 0x4000 : 0) | orb);
    };
    for (int i = 0; i < 6; ++i)
      for (int s = 0; s < 2; ++s) {
        const int j = (i + 1) % 6;
        terms.push_back({-1.0, {op(false, s, i), op(true, s, j)}});
        terms.push_back({-1.0, {op(false, s, j), op(true, s, i)}});
      }
    for (int i = 0; i < 6; ++i)
      terms.push_back({4.0, {op(false, 0, i), op(true, 0, i), op(false, 1, i), op(true, 1, i)}});
    auto E = mbamd::expression_operator(basis, terms);
    CHECK(E.dimension() == H.dimension());
    auto x = H.gaussian(4);
    auto y0 = H.apply(x).to_host(), y1 = E.apply(x).to_host();
    double d = 0.0, m = 0.0;
    for (size_t i = 0; i < y0.size(); ++i) {
      d = std::max(d, std::abs(y0[i] - y1[i]));
      m = std::max(m, std::abs(y0[i]));
    }
    CHECK(d <= 1e-13 * m);
    auto sp = mbamd::compute_matrix_elements(E);
    std::vector<uint64_t> r, c;
    std::vector<std::complex<double>> v;
    sp.entries(r, c, v);
    CHECK(sp.n_rows() == H.dimension() && (int64_t)v.size() == sp.n_nonzero());
    bool sorted = true;
    for (size_t i = 1; i < v.size(); ++i)
      sorted &= c[i] > c[i - 1] || (c[i] == c[i - 1] && r[i] > r[i - 1]);
    CHECK(sorted);
    const std::string path = "/tmp/mbamd_host_test_matrix.bin";
    sp.save(path);
    auto back = mbamd::SparseMatrix::load(path);
    std::vector<uint64_t> r2, c2;
    std::vector<std::complex<double>> v2;
    back.entries(r2, c2, v2);
    CHECK(r2 == r && c2 == c && v2 == v);
    std::remove(path.c_str());
  }

  {
    CASE("generic_solve_max_eigenpair_exp_bounds");
    // ring L = 8, 4 + 4, U = 0: spectrum within [-9.657, 9.657] (free fermions)
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(8, 8, 0);
    auto H = HubbardModel(1.0, 0.0, 8).op(basis);
    auto b = H.gaussian(9);
    auto mx = mbamd::find_max_eigenpair(H, 150, b);
    CHECK(std::abs(mx.value - 9.656854249492380) < 1e-8);
    // solve with the lowest-eigenvector callback reproduces find_min_eigenpair
    auto x = mbamd::solve(H, b, 150, [](const std::vector<double> &al,
                                         const std::vector<double> &be) {
      const int m = (int)al.size();
      std::vector<double> w(m), z((size_t)m * m);
      mb_tridiag_eig(m, al.data(), be.data(), w.data(), z.data());
      std::vector<std::complex<double>> y(m);
      for (int i = 0; i < m; ++i) y[i] = z[i];
      return y;
    });
    auto mn = find_min_eigenpair(H, 150, b);
    auto xh = x.to_host(), mh = mn.vector.to_host();
    double d = 0.0, n = 0.0;
    for (size_t i = 0; i < xh.size(); ++i) {
      d = std::max(d, std::abs(xh[i] - mh[i]));
      n = std::max(n, std::abs(mh[i]));
    }
    CHECK(d <= 1e-10 * n);
    auto bounds = mbamd::estimate_bounds(H, b, 60, 0.1);
    CHECK(bounds.alpha <= bounds.beta && bounds.beta <= 9.656854249492380 + 1e-6);
    auto e = mbamd::Exp(H, 30).apply(b);  // exp(H) b: norm grows by at most e^{9.66}
    double ne = 0.0, nb = 0.0;
    auto eh = e.to_host(), bh = b.to_host();
    for (size_t i = 0; i < eh.size(); ++i) {
      ne += std::norm(eh[i]);
      nb += std::norm(bh[i]);
    }
    CHECK(std::sqrt(ne) > std::sqrt(nb) && std::sqrt(ne) < std::exp(9.66) * std::sqrt(nb));
  }

  {
    CASE("exact_diagonalization_example");
    // examples/hubbard_exact_diagonalization.cpp: L = 4, N = 2, t = 1, U = 4 on
    // the full fixed-N basis and on the spin-0 sector; the ground state is a
    // singlet, so both give the same lowest eigenvalue
    HubbardModel hubbard(1.0, 4.0, 4);
    auto full_basis = FermionBasis::with_fixed_particle_number(4, 2);
    auto spin_zero_basis = FermionBasis::with_fixed_particle_number_and_spin(4, 2, 0);
    CHECK(full_basis.size() == 28 && spin_zero_basis.size() == 16);
    auto H = hubbard.op(full_basis);
    auto H0 = hubbard.op(spin_zero_basis);
    auto e = find_min_eigenpair(H, 28, H.gaussian(5));
    auto e0 = find_min_eigenpair(H0, 16, H0.gaussian(5));
    std::printf("  ground state (full basis) %.12f, (spin 0) %.12f\n", e.value, e0.value);
    CHECK(std::abs(e.value - e0.value) < 1e-10);
    auto sp = mbamd::compute_matrix_elements(H);  // H as an arma::sp_cx_mat
    CHECK(sp.n_rows() == 28 && sp.n_cols() == 28 && sp.n_nonzero() == 100);
  }

  {
    CASE("rectangular_and_vector_elements");
    // c+_{up 0} from (1 up, 1 down) to (2 up, 1 down) on 3 orbitals: every column
    // whose up orbital 0 is empty maps to one row with sign +1 (nothing passed);
    // c+_{dn 1} c+_{up 0}|0> = -c+_{up 0} c+_{dn 1}|0> (one transposition)
    auto row = FermionBasis::with_fixed_particle_number_and_spin(3, 3, 1);
    auto col = FermionBasis::with_fixed_particle_number_and_spin(3, 2, 0);
    const uint16_t up0 = 0, dn1 = 0x4000 | 1;
    auto sp = compute_rectangular_matrix_elements(row, col, {{{1.0, 0.0}, {up0}}});
    CHECK(sp.n_rows() == row.size() && sp.n_cols() == col.size());
    std::vector<uint64_t> r, c;
    std::vector<std::complex<double>> v;
    sp.entries(r, c, v);
    CHECK(v.size() == 6);  // 2 empty up orbitals 1, 2 (up 0 empty) x 3 down orbitals
    for (size_t k = 0; k < v.size(); ++k) {
      const auto cs = col.state(c[k]), rs = row.state(r[k]);
      CHECK(rs.first == (cs.first | 1u) && rs.second == cs.second && v[k] == 1.0);
    }
    auto vec = compute_vector_elements(col, {{{2.0, 0.0}, {dn1, up0}}});
    CHECK(vec.size() == col.size());
    for (size_t i = 0; i < vec.size(); ++i)
      CHECK(vec[i] == (i == col.index_of(1, 2) ? std::complex<double>(-2.0, 0.0) : 0.0));
  }

  {
    CASE("operator_algebra");
    // ring L = 8, 4 + 4, U = 0: E0 = -9.656854249492380; 2 H + 1 has 2 E0 + 1
    // and H - H is zero
    auto basis = FermionBasis::with_fixed_particle_number_and_spin(8, 8, 0);
    auto H = HubbardModel(1.0, 0.0, 8).op(basis);
    auto twoH = 2.0 * H;
    auto A = Shifted(twoH, 1.0);
    auto e = find_min_eigenpair(A, 150, H.gaussian(2));
    CHECK(std::abs(e.value - (2.0 * -9.656854249492380 + 1.0)) < 1e-9);
    auto Z = H - H;
    auto x = H.gaussian(5);
    auto z = Z.apply(x).to_host();
    double m = 0.0;
    for (auto v : z) m = std::max(m, std::abs(v));
    CHECK(m == 0.0);
  }

  std::printf(g_fail ? "FAILED (%d)\n" : "ALL PASSED\n", g_fail);
  return g_fail ? 1 : 0;
}
