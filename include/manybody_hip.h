/* manybody_hip.h — C ABI of the MI355X (gfx950) Hubbard H.v path.
 *
 * Drop-in boundary for keyehzy/manybody's fixed-particle-number Hubbard path:
 * the basis (FermionBasis), the model (HubbardModel / 2D / 3D), the operator
 * behind LinearOperator::apply, and the Krylov consumers (Lanczos, Krylov
 * exp(-iHt), KPM).  Each entry point cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns an int status (MB_OK = 0) and never throws; the
 *    message of the last failure on the calling thread is mb_last_error().
 *    Status codes map to the reference's exceptions: MB_ERR_RUNTIME ~
 *    std::runtime_error, MB_ERR_INVALID ~ std::invalid_argument,
 *    MB_ERR_RANGE ~ std::out_of_range.
 *  - Handles are opaque.  A model handle keeps a reference to its basis handle:
 *    the caller keeps the basis alive at least as long as the model (as the
 *    reference's operators hold `const FermionBasis&`).
 *  - Vectors are device pointers.  "ref" vectors are indexed by the reference
 *    basis index (the DFS position, basis.cpp:267-302) and are interleaved
 *    complex (re, im per element), like arma::cx_vec.  "dev" vectors use the
 *    device order: planar, all D real parts then all D imaginary parts, each
 *    plane in product order x[u * D_dn + d] (u, d = combinatorial ranks of the
 *    up and down occupation bitstrings).  nvec complex vectors in device order
 *    are 2 nvec consecutive planes (vector r = planes 2r, 2r+1).
 *  - `stream` is a hipStream_t (NULL = default stream).  Entry points that
 *    return host scalars synchronize that stream before returning.
 *  - Concurrent calls on one const model from different host threads and
 *    streams are safe for mb_hv_* and the permutation calls (read-only tables,
 *    per-call scratch), mirroring the reference's concurrent const apply().
 */
#ifndef MANYBODY_HIP_H
#define MANYBODY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MB_OK 0
#define MB_ERR_INVALID 1 /* std::invalid_argument */
#define MB_ERR_RUNTIME 2 /* std::runtime_error */
#define MB_ERR_RANGE 3   /* std::out_of_range */
#define MB_ERR_DEVICE 4  /* HIP runtime failure */
#define MB_ERR_NOMEM 5   /* device allocation failed or admission refused */

typedef struct mb_basis mb_basis;
typedef struct mb_hubbard mb_hubbard;
typedef void *mb_stream; /* hipStream_t */

/* Message of the last failed call on this thread ("" if none). */
const char *mb_last_error(void);

/* Library version string. */
const char *mb_version(void);

/* ------------------------------------------------------------------ basis */

/* FermionBasis::with_fixed_particle_number_and_spin(orbitals, particles,
 * spin_projection)  — src/algebra/basis.h:25-26, basis.cpp:91-126.
 * up = (particles + spin_projection) / 2, down = particles - up.  Violations of
 * the reference's asserts (basis.cpp:97-117) return MB_ERR_INVALID.
 * orbitals <= 32. */
int mb_basis_create(int orbitals, int particles, int spin_projection, mb_basis **out);
void mb_basis_destroy(mb_basis *b);

/* basis.set.size() — compute_basis_size_spin, basis.cpp:39-50. */
int64_t mb_basis_size(const mb_basis *b);

/* Occupations of basis.set[index] as bitmasks (bit i = orbital i) —
 * IndexedHashSet::operator[] (indexed_hash_set.h).  MB_ERR_RANGE if index is
 * out of range. */
int mb_basis_state(const mb_basis *b, int64_t index, uint64_t *up_bits, uint64_t *dn_bits);

/* basis.set.index_of(key) — indexed_hash_set.h:158-164.  MB_ERR_RANGE if the
 * state is not in the basis (the reference throws std::out_of_range). */
int mb_basis_index_of(const mb_basis *b, uint64_t up_bits, uint64_t dn_bits, int64_t *index);

/* Reference index -> device product index, for n consecutive reference indices
 * starting at `first` (host arrays).  Used by hosts that keep their own vectors
 * in reference order. */
int mb_basis_product_index(const mb_basis *b, int64_t first, int64_t n, int64_t *product_index);

/* ------------------------------------------------------------------ model */

/* HubbardModel(t, u, size) / HubbardModel2D(t, u, sx, sy) /
 * HubbardModel3D(t, u, sx, sy, sz) bound to a basis —
 * src/algebra/fermion/model/hubbard_model.h:10-212.  ndim in {1, 2, 3};
 * sizes[0..ndim) are the linear sizes; their product must equal the basis'
 * orbital count.  Bonds, their loop order, the merge of repeated bonds
 * (linear size 2) and self bonds (linear size 1) follow the reference.
 * Allocates the model's device tables on the current HIP device. */
int mb_hubbard_create(const mb_basis *b, int ndim, const int *sizes, double t, double u,
                      mb_hubbard **out);

/* Same Hamiltonian form over an explicit bond list: H = sum_k -t (hopping(from[k], to[k], up)
 * + hopping(from[k], to[k], down)) + u sum_i n_i,up n_i,dn, hopping as in
 * hopping_generic (expression_base.h:394-403). */
int mb_hubbard_create_bonds(const mb_basis *b, int n_bonds, const int *from, const int *to,
                            double t, double u, mb_hubbard **out);
void mb_hubbard_destroy(mb_hubbard *h);

/* LinearOperator::dimension() — linear_operator.h:42. */
int64_t mb_hubbard_dimension(const mb_hubbard *h);

/* Device-order bookkeeping: D_up, D_dn (dimension = D_up * D_dn). */
int mb_hubbard_shape(const mb_hubbard *h, int64_t *d_up, int64_t *d_dn);

/* --------------------------------------------------------------- H . v */

/* y = H x, complex128, device order.  Replaces compute_matrix_elements
 * (matrix_elements.h:74-105) followed by SparseComplexMatrixOperator::apply
 * (linear_operator.h:81-91).  x and y must not alias. */
int mb_hv_c128(const mb_hubbard *h, const void *x_dev, void *y_dev, mb_stream stream);

/* y = H x, complex64, device order (FP32 arithmetic; the reference has no
 * FP32 path).  Real-space Hubbard operators and expressions of Hubbard form
 * (see mb_expression_create); MB_ERR_INVALID for other handles. */
int mb_hv_c64(const mb_hubbard *h, const void *x_dev, void *y_dev, mb_stream stream);

/* y = H x for nvec complex64 vectors in device order (2 nvec planes). */
int mb_hv_c64_multi(const mb_hubbard *h, int nvec, const void *x_dev, void *y_dev,
                    mb_stream stream);

/* Reference order <-> device order for complex128 / complex64 vectors
 * (a gather and a scatter; the permutation is evaluated by rank arithmetic on
 * the device, no table is stored). */
int mb_to_device_order_c128(const mb_hubbard *h, const void *x_ref, void *x_dev, mb_stream stream);
int mb_to_reference_order_c128(const mb_hubbard *h, const void *x_dev, void *x_ref,
                               mb_stream stream);

/* The whole reference-order apply: y_ref = H x_ref, using `work` (2 device
 * complex128 vectors of the operator's dimension, caller-owned). */
int mb_apply_ref_c128(const mb_hubbard *h, const void *x_ref, void *y_ref, void *work,
                      mb_stream stream);

/* Seeded start vectors, written in device order but drawn per reference index
 * i: re, im ~ N(0,1) i.i.d. from (seed, i) (SplitMix64 -> Box-Muller).  Not
 * normalized (Lanczos normalizes, lanczos.h:96). */
int mb_fill_gaussian_c128(const mb_hubbard *h, uint64_t seed, void *x_dev, mb_stream stream);

/* nvec random-phase complex64 vectors e^{i phi}, phi ~ U[0, 2 pi) from
 * (seed, vector, reference index), device order (2 nvec planes). */
int mb_fill_phase_c64_multi(const mb_hubbard *h, uint64_t seed, int nvec, void *x_dev,
                            mb_stream stream);

/* The same random-phase vectors in complex128 (e^{i phi} evaluated in FP64);
 * any operator handle (the complex64 form needs a real-space Hubbard operator). */
int mb_fill_phase_c128_multi(const mb_hubbard *h, uint64_t seed, int nvec, void *x_dev,
                             mb_stream stream);

/* ----------------------------------------------------------- Krylov layer */

/* lanczos_pass_one(op, b, k) — lanczos.h:130-175 (recurrence :15-40, no
 * reorthogonalization).  b is a device-order complex128 vector.  alphas holds k
 * doubles, betas k doubles (the first steps-1 are meaningful, as
 * LanczosDecomposition::betas); *steps = steps_taken, *b_norm = ||b||.
 * Zero-norm b -> MB_ERR_RUNTIME "Input vector has zero norm". */
int mb_lanczos_pass_one(const mb_hubbard *h, const void *b_dev, int k, double *alphas,
                        double *betas, int *steps, double *b_norm, mb_stream stream);

/* No counterpart in the reference: mb_lanczos_pass_one for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results; every
 * buffer beyond b (three complex128 vectors, alpha/beta, partial sums, flags)
 * comes from a static 48 MiB device array that each GPU holds until the
 * process ends, so it calls no hipMalloc/hipFree.  Limit: 48 D + 16 k bytes
 * plus a few KB of partial sums must fit, i.e. D <= ~1,040,000 (every sector
 * with L <= 12) for k up to a few thousand; larger runs -> MB_ERR_INVALID with
 * the byte counts, before any device work.  Real-space Hubbard operators only
 * (MB_ERR_INVALID otherwise).  Calls on one GPU share the array; they are
 * serialised by a host lock held for the whole call, which returns after its
 * stream has drained. */
int mb_lanczos_pass_one_noalloc(const mb_hubbard *h, const void *b_dev, int k, double *alphas,
                                double *betas, int *steps, double *b_norm, mb_stream stream);

/* find_min_eigenpair(op, k) with an explicit start vector b (the reference
 * draws it with arma::randu, linear_operator.h:26-32) — lanczos.h:246-290.
 * *value = lowest Ritz value; if x_dev != NULL it receives the eigenvector
 * (scaled by ||b|| as lanczos.h:276, rebuilt by pass two :177-218). */
int mb_find_min_eigenpair(const mb_hubbard *h, const void *b_dev, int k, double *value,
                          void *x_dev, mb_stream stream);

/* time_evolve_state(H, psi, t, {krylov_steps}) — evolve_state.h:164-178:
 * out = exp(-i t H) psi, complex128 device order.  krylov_steps 0 -> D. */
/* Lanczos options (no reference counterpart): MB_LANCZOS_REORTHOGONALIZE
 * orthogonalises every new vector twice (classical Gram-Schmidt, FP64 on the
 * device) against all previous ones and keeps the k + 1 vectors in HBM
 * (16 D (k + 1) bytes; MB_ERR_NOMEM if they do not fit).  The recurrence of
 * lanczos.h:15-40 is otherwise unchanged; flags = 0 is mb_lanczos_pass_one. */
#define MB_LANCZOS_REORTHOGONALIZE 1
int mb_lanczos_pass_one_ex(const mb_hubbard *h, const void *b_dev, int k, int flags, double *alphas,
                           double *betas, int *steps, double *b_norm, mb_stream stream);
/* As mb_find_min_eigenpair; with MB_LANCZOS_REORTHOGONALIZE the eigenvector is
 * assembled from the stored vectors (no second pass). */
int mb_find_min_eigenpair_ex(const mb_hubbard *h, const void *b_dev, int k, int flags,
                             double *value, void *x_dev, mb_stream stream);
int mb_time_evolve(const mb_hubbard *h, const void *psi_dev, double t, int krylov_steps,
                   void *out_dev, mb_stream stream);

/* time_evolve_state_steps(H, psi, t0, t1, dt, opts) — evolve_state.h:204-231
 * (the callback is not exposed).  out = psi(t1). */
int mb_time_evolve_steps(const mb_hubbard *h, const void *psi_dev, double t0, double t1,
                         double dt, int krylov_steps, void *out_dev, mb_stream stream);

/* Symmetric tridiagonal eigen-decomposition on the host (replaces arma::eig_sym
 * of T_k, lanczos.h:258-271): w ascending, V (m*m, column j = eigenvector j,
 * V[i + m*j]) optional. */
/* Time evolution options (no reference counterpart): MB_EVOLVE_FP32 runs the
 * Krylov vectors and H.v in FP32 complex; the state psi, the norms, the Lanczos
 * alpha/beta reductions and the exponential of T_k stay FP64.  flags = 0 is
 * mb_time_evolve / mb_time_evolve_steps. */
#define MB_EVOLVE_FP32 1
int mb_time_evolve_ex(const mb_hubbard *h, const void *psi_dev, double t, int krylov_steps, int flags,
                      void *out_dev, mb_stream stream);
int mb_time_evolve_steps_ex(const mb_hubbard *h, const void *psi_dev, double t0, double t1, double dt,
                            int krylov_steps, int flags, void *out_dev, mb_stream stream);
/* time_evolve_state_steps(H, psi, t0, t1, dt, opts, callback) (evolve_state.h:
 * 199-231): callback(t, state_dev, user) at t0 and after every step, with the
 * state in out_dev (device order) and the stream synchronised; a nonzero
 * return aborts with MB_ERR_RUNTIME. */
typedef int (*mb_step_callback)(double t, const void *state_dev, void *user);
int mb_time_evolve_steps_cb(const mb_hubbard *h, const void *psi_dev, double t0, double t1,
                            double dt, int krylov_steps, int flags, mb_step_callback callback,
                            void *user, void *out_dev, mb_stream stream);
/* eigen-decomposition of the symmetric tridiagonal T (alphas on the diagonal,
 * m - 1 betas beside it; arma::eig_sym at lanczos.h:269): eigenvalues ascending
 * in w[m], eigenvectors as the columns of V[m * m] (column-major; V may be NULL) */
int mb_tridiag_eig(int m, const double *alphas, const double *betas, double *w, double *V);

/* KPM density-of-states moments mu_0..mu_{M-1} of Hs = (H - b)/a, averaged
 * over nvec random-phase vectors (seeded as mb_fill_phase_c64_multi):
 * mu_n = (1/(nvec D)) sum_r <r|T_n(Hs)|r>.  The Chebyshev recurrence of
 * SparseKPMProjector::apply (kpm.h:212-229); H.v in FP32 complex (fp32 = 1) or
 * FP64 complex (fp32 = 0); inner products and moments accumulate in FP64.
 * M even, 1 <= nvec <= 64.  Real-space Hubbard operators, and expressions of
 * Hubbard form on a real-space sector (see mb_expression_create), advance all
 * vectors in one launch; any other handle (momentum sector, fixed-N basis,
 * other expressions, composite) runs FP64 one vector at a time with the same
 * phases (MB_ERR_INVALID for fp32 = 1 there). */
int mb_kpm_dos_moments(const mb_hubbard *h, double a, double b, int M, int nvec, uint64_t seed,
                       int fp32, double *mu, mb_stream stream);

/* kpm::jackson_kernel(N) — kpm.h:19-30: g[0..N]. */
int mb_jackson_kernel(int N, double *g);

/* rescaling::from_bounds(E_min, E_max, padding) — rescaling.h:33-38. */
int mb_rescaling_from_bounds(double e_min, double e_max, double padding, double *a, double *b);

/* ------------------------------------------------------------ momentum sectors */

/* Basis::with_fixed_particle_number_spin_momentum (basis.h:27-29,
 * basis.cpp:128-185, 304-356): the states of total momentum `momentum`
 * (ndim components) on the periodic lattice `sizes` (orbital k is the momentum
 * k_x + L_x k_y + L_x L_y k_z), indexed in the reference's order (its real-space
 * DFS order restricted to the sector).  The handle works with mb_basis_size,
 * mb_basis_state and mb_basis_index_of; MB_ERR_INVALID for the real-space model
 * constructors. */
int mb_momentum_basis_create(int orbitals, int particles, int spin_projection, int ndim,
                             const int *sizes, const int *momentum, mb_basis **out);
/* HubbardMomentumFactorizedFixedK (hubbard_momentum_factorized_fixed_k.h:41-117):
 * H = t sum_{k,s} eps(k) n_{k,s} + (U/N) sum_q rho_{q,up} rho_{-q,dn},
 * eps(k) = -2 sum_d cos(2 pi k_d / L_d), on a fixed-momentum basis (sizes and
 * momentum must be the basis's).  The handle works with the FP64 entry points:
 * mb_hubbard_dimension, mb_hv_c128, mb_apply_ref_c128, the order conversions,
 * mb_fill_gaussian_c128, Lanczos (plain and reorthogonalised), the eigenpair,
 * FP64 time evolution and mb_energy_residual; the others return MB_ERR_INVALID.
 * Device order: blocks of (up strings of momentum p) x (down strings of
 * momentum K - p); mb_hubbard_shape reports (D, 1). */
int mb_momentum_hubbard_create(const mb_basis *b, int ndim, const int *sizes, double t, double u,
                               const int *momentum, mb_hubbard **out);

/* ------------------------------------------------------------ optical conductivity */

typedef struct mb_current mb_current;
/* CurrentOperatorFixedK(source_basis, target_basis, lattice_size, t, Q, direction)
 * (current_operator_fixed_k.h:28-69): J_d(Q) = sum_{k,s} v_d(k) c+_{k+Q,s} c_{k,s},
 * v_d(k) = 2 t (2 pi / L_d) sin(2 pi k_d / L_d), from the fixed-momentum basis K
 * (source) to K + Q (target); Q is the flat momentum index.  MB_ERR_INVALID if
 * the bases are not fixed-momentum bases of that lattice or target != K + Q. */
int mb_current_create(const mb_basis *source, const mb_basis *target, int ndim, const int *sizes,
                      double t, int64_t Q, int direction, mb_current **out);
void mb_current_destroy(mb_current *j);
/* y = J x (source -> target) and x = J^dagger y (adjoint_apply, :96-117), device order */
int mb_current_apply(const mb_current *j, const void *x_src, void *y_tgt, mb_stream stream);
int mb_current_adjoint_apply(const mb_current *j, const void *y_tgt, void *x_src, mb_stream stream);
/* imaginary_time_evolve_state (evolve_state.h:184-197): out = exp(-tau H) psi */
int mb_imaginary_time_evolve(const mb_hubbard *h, const void *psi_dev, double tau, int krylov_steps,
                             void *out_dev, mb_stream stream);
/* The stochastic current-current correlator of
 * hubbard_momentum_optical_conductivity_factorized.cpp:226-283: per sample n a
 * seeded N(0,1) vector (seed + n, per reference index) in sector K, normalised,
 * v_b = exp(-beta H_K / 2) v, Z = <v_b|v_b>, then C(i dt) += <phi1|J^dagger phi2>/Z
 * with phi1 = exp(-i H_K t) v_b, phi2 = exp(-i H_{K+Q} t) J v_b; averaged over the
 * samples.  corr: 2 * steps doubles (re, im).  hk and hkq may be any operators on
 * the current operator's source and target bases (the factorized momentum-space
 * model, or a compiled expression on those fixed-momentum bases). */
int mb_optical_correlator(const mb_hubbard *hk, const mb_hubbard *hkq, const mb_current *j,
                          int samples, double beta, double dt, int steps, int krylov_steps,
                          uint64_t seed, double *corr, mb_stream stream);
/* compute_optical_conductivity (optical_conductivity.cpp:8-80) on the host:
 * corr holds n complex values (re, im); writes *n_freqs (omega and sigma NULL:
 * size query), omega[n_freqs], sigma[2 n_freqs] (re, im) and *sum_rule. */
int mb_optical_conductivity(const double *corr, int n, double dt, double beta, double volume,
                            double *omega, double *sigma, int capacity, int *n_freqs,
                            double *sum_rule);

/* ------------------------------------------------------------ KPM projector */

/* kpm::projector_moments(M, epsilon) (kpm.h:46-59): mu[M + 1] of the step
 * theta(epsilon - x); MB_ERR_INVALID for |epsilon| > 1. */
int mb_projector_moments(int M, double epsilon, double *mu);
/* kpm::lorentz_kernel(N, lambda) (kpm.h:32-43): g[N + 1] */
int mb_lorentz_kernel(int N, double lambda, double *g);
/* SparseKPMProjector(H, M, E_min, E_max, E_fermi, padding).apply(v)
 * (kpm.h:184-241) for any operator: out = sum_m g_m mu_m(eps) T_m(H~) v with
 * H~ = (H - b) / a, a = (E_max - E_min)(1 + padding) / 2, b = (E_max + E_min) / 2,
 * eps = (E_fermi - b) / a, Jackson g; M H.v on the device. */
int mb_kpm_projector_apply(const mb_hubbard *h, int M, double e_min, double e_max,
                           double e_fermi, double padding, const void *v_dev, void *out_dev,
                           mb_stream stream);

/* ------------------------------------------------------------ spectral bounds */

/* power_method(op, v, iterations) (linear_operator_utils.h:18-50) from seed_dev
 * (device order, not modified): the Rayleigh-type estimate
 * Re(sum v_i (Hv)_i) / Re(sum v_i v_i) of the last iteration (arma::dot, no
 * conjugation), v renormalised each step; 0 for a zero seed or iterations 0. */
int mb_power_method(const mb_hubbard *h, const void *seed_dev, int iterations, double *value,
                    mb_stream stream);
/* estimate_bounds(op, power_iterations, spectral_padding) (:52-78): dominant
 * |lambda| by the power method, then the power method on H + shift and
 * shift - H with shift = dominant (1 + padding) + eps; *alpha <= *beta.  The
 * reference draws its seed with make_seed_vector; here the caller passes it. */
int mb_estimate_bounds(const mb_hubbard *h, const void *seed_dev, int power_iterations,
                       double spectral_padding, double *alpha, double *beta, mb_stream stream);

/* Exp(op, ExpOptions{krylov_steps}).apply(v) (exp_operator.h:50-89): exp(H) v
 * by the Krylov solve with exp(T_k) e1 (k = krylov_steps, 0 or > D meaning D);
 * v returned unchanged when |v| <= 1000 eps. */
int mb_exp_apply(const mb_hubbard *h, const void *v_dev, int krylov_steps, void *out_dev,
                 mb_stream stream);

/* solve(op, b, k, solver) (lanczos.h:220-237): Lanczos pass one from b, then
 * the caller's solver fills y[2m] (re, im) from the m alphas and m - 1 betas
 * (return 0; nonzero aborts with MB_ERR_RUNTIME), then x = |b| sum_j y_j v_j
 * (pass two, or the stored Krylov basis).  Eigenpairs, exp and linear solves
 * of the reference are instances; x is zero when no step was taken. */
typedef int (*mb_tridiag_solver)(int m, const double *alphas, const double *betas, double *y,
                                 void *user);
int mb_lanczos_solve(const mb_hubbard *h, const void *b_dev, int k, mb_tridiag_solver solver,
                     void *user, void *x_dev, mb_stream stream);
/* find_max_eigenpair(op, k) (lanczos.h:246-281) with an explicit start vector;
 * x_dev may be NULL (value only) */
int mb_find_max_eigenpair(const mb_hubbard *h, const void *b_dev, int k, double *value,
                          void *x_dev, mb_stream stream);

/* ------------------------------------------------------------ fixed-N bases */

/* Basis::with_fixed_particle_number(orbitals, particles) (basis.cpp:70-89,
 * 240-265): every state with that many particles, any N_up, indexed in the
 * reference's depth-first order (orbitals <= 32, C(2 orbitals, particles) <
 * 2^31).  The handle works with mb_basis_size/state/index_of, with
 * mb_expression_create (terms may move particles between spins: spin flips)
 * and with mb_hubbard_create / mb_hubbard_create_bonds (the Hubbard model is
 * compiled as an expression on it; at most 24 orbitals for operators). */
int mb_fixed_n_basis_create(int orbitals, int particles, mb_basis **out);

/* ------------------------------------------------------------ generic expressions */

/* compute_matrix_elements(basis, A) (matrix_elements.h:74-105) for any
 * FermionExpression A, as an operator handle: the expression's terms are
 * n_terms monomials, coefficient t = coeffs[2t] + i coeffs[2t+1], operators
 * ops[offsets[t] .. offsets[t+1]) in the reference's OperatorStorage encoding
 * (operator.h:14-40: bit 15 annihilation, bit 14 spin down, bits 0-13 orbital),
 * the product o_1 o_2 ... o_n with o_n acting first.  Terms are merged by key and
 * dropped below 1000 eps as FermionExpression does (expression_base.h:85-120);
 * outputs outside the basis are dropped.  Works on fixed-N bases (real space)
 * and fixed-momentum bases (at most 24 orbitals).  The handle works with the
 * FP64 H.v, order, fill, apply_ref, Lanczos, eigenpair, time-evolution and
 * energy/residual entry points; MB_ERR_INVALID for an operator on an orbital
 * >= the basis orbitals.  On a fixed (N_up, N_dn) real-space sector an
 * expression of Hubbard form (after merging and dropping: every hop one-body
 * within a species with a real amplitude, real on-site values, and mixed terms
 * only U n_{i,up} n_{i,dn} with one U on every site) runs on the Hubbard
 * operator's tables and kernels built from its own hop lists, which also gives
 * it FP32 H.v and FP32 KPM; MANYBODY_AMD_EXPR_NATIVE=0 keeps the generic
 * expression kernels. */
int mb_expression_create(const mb_basis *basis, int64_t n_terms, const double *coeffs,
                         const int64_t *offsets, const uint16_t *ops, mb_hubbard **out);

/* The reference's LinearOperator algebra (linear_operator.h:93-281) on operator
 * handles: Negated -A, Scaled c A, Shifted A + c I, NegShift -A + c I, Sum
 * A + B, Difference A - B, Composed A B (B applied first) and Identity(basis).
 * c is complex (re, im).  The result is an operator handle that works with the
 * FP64 H.v, order, fill, apply_ref, Lanczos, eigenpair, time-evolution, Exp,
 * power-method, bounds and energy/residual entry points; composites nest.
 * The operands are referenced, not copied: they must outlive the composite
 * (the reference's wrappers hold copies).  Sum, Difference and Composed need
 * operands on bases of the same kind and parameters (MB_ERR_INVALID
 * otherwise).  Each H.v of a composite takes its scratch vectors with
 * hipMallocAsync on the call's stream. */
int mb_operator_negated(const mb_hubbard *a, mb_hubbard **out);
int mb_operator_scaled(const mb_hubbard *a, double re, double im, mb_hubbard **out);
int mb_operator_shifted(const mb_hubbard *a, double re, double im, mb_hubbard **out);
int mb_operator_neg_shifted(const mb_hubbard *a, double re, double im, mb_hubbard **out);
int mb_operator_sum(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out);
int mb_operator_difference(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out);
int mb_operator_composed(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out);
int mb_operator_identity(const mb_basis *basis, mb_hubbard **out);
/* rescaling::rescale_hamiltonian(H, Rescaling{scale, shift}) (rescaling.h:40-48)
 * on an operator: (A - shift I) / scale, applied as A x / scale - (shift /
 * scale) x.  MB_ERR_INVALID for a zero or non-finite scale. */
int mb_operator_rescaled(const mb_hubbard *a, double scale, double shift, mb_hubbard **out);

/* The sparse matrix of compute_matrix_elements (arma::sp_cx_mat): nonzeros in
 * the reference's basis index order, column-major with rows ascending (SpMat
 * iteration order), built on the device from an expression operator. */
typedef struct mb_sparse mb_sparse;
int mb_compute_matrix_elements(const mb_hubbard *h, mb_sparse **out, mb_stream stream);
/* from entries already column-major with rows ascending and unique */
int mb_sparse_create(uint64_t n_rows, uint64_t n_cols, int64_t nnz, const uint64_t *rows,
                     const uint64_t *cols, const double *vals, mb_sparse **out);
void mb_sparse_destroy(mb_sparse *sp);
int mb_sparse_shape(const mb_sparse *sp, uint64_t *n_rows, uint64_t *n_cols, int64_t *nnz);
/* rows[nnz], cols[nnz], vals[2 nnz] (re, im) */
int mb_sparse_entries(const mb_sparse *sp, uint64_t *rows, uint64_t *cols, double *vals);
/* save / load in the reference's cereal binary format (arma_cereal.h:9-30 through
 * cereal::BinaryOutputArchive, as save_sparse_matrix / load_sparse_matrix of
 * examples/hubbard_momentum_optical_conductivity.cpp:185-203 write them):
 * uint32 class version 0, uint64 n_rows, n_cols, n_nonzero, then per nonzero
 * uint64 row, uint64 col, float64 re, float64 im, little endian. */
int mb_sparse_save_cereal(const mb_sparse *sp, const char *path);
int mb_sparse_load_cereal(const char *path, mb_sparse **out);

/* compute_rectangular_matrix_elements(row_basis, col_basis, A)
 * (matrix_elements.h:110-167): <i|A|j> for i in row_basis and j in col_basis
 * (reference indices), as an n_rows x n_cols sparse matrix in SpMat iteration
 * order.  Terms as for mb_expression_create; A may change the particle numbers
 * (creators, annihilators, pair creation) and the two bases may be of any kinds
 * (fixed N_up/N_dn, fixed N, fixed momentum).  Built on the device.
 * MB_ERR_INVALID if the bases have different orbital counts or more than 24. */
int mb_compute_rectangular_matrix_elements(const mb_basis *row_basis, const mb_basis *col_basis,
                                           int64_t n_terms, const double *coeffs,
                                           const int64_t *offsets, const uint16_t *ops,
                                           mb_sparse **out, mb_stream stream);
/* compute_vector_elements(basis, A) (matrix_elements.h:31-51): <i|A|0> for every
 * state i of the basis, |0> the vacuum, into out[2 * size] (host memory,
 * re, im interleaved, reference order). */
int mb_compute_vector_elements(const mb_basis *basis, int64_t n_terms, const double *coeffs,
                               const int64_t *offsets, const uint16_t *ops, double *out,
                               mb_stream stream);

/* ------------------------------------------------------------ observables */

/* <x|G_updn(r)|x> / <x|x> for every lattice displacement r, written to g[r]
 * (r = rx + Lx ry + Lx Ly rz; one value per site), where
 * G_updn(r) = (1/N) sum_i n_{i,up} n_{i+r,dn} is the reference's
 * HubbardModel*::opposite_spin_correlation (hubbard_model.h:40-53, 105-120,
 * 184-200).  x in device order.  MB_ERR_INVALID for a model made from a bond
 * list (no lattice). */
int mb_opposite_spin_correlation(const mb_hubbard *h, const void *x_dev, double *g,
                                 mb_stream stream);
/* Rayleigh quotient E = <x|H|x>/<x|x> and residual |Hx - E x| / |x| (x in device order). */
int mb_energy_residual(const mb_hubbard *h, const void *x_dev, double *energy, double *residual,
                       mb_stream stream);

/* ------------------------------------------------------------ batched solves */

/* Many independent Lanczos solves advanced in lockstep by one launch per phase
 * of a step for the whole batch (no reference counterpart: the reference runs
 * lanczos_pass_one, lanczos.h:130-175, once per operator).  Models may repeat;
 * the caller keeps every model (and its basis) alive while the batch exists. */
typedef struct mb_batch mb_batch;
int mb_batch_create(int n, const mb_hubbard *const *models, mb_batch **out);
void mb_batch_destroy(mb_batch *b);
/* Device bytes held by the batch's vectors. */
int64_t mb_batch_footprint(const mb_batch *b);
/* Start vector of every instance: seeded N(0,1) per reference index as
 * mb_fill_gaussian_c128 with seeds[i]; or copied from a device-order vector. */
int mb_batch_fill_gaussian(mb_batch *b, const uint64_t *seeds, mb_stream stream);
int mb_batch_set_start(mb_batch *b, int i, const void *b_dev, mb_stream stream);
/* Work mapping: elements per tile (one workgroup) and workgroup size. */
int mb_batch_set_mapping(mb_batch *b, int tile, int block);
/* Instances of dimension up to max_dim (default 2048) run all their Lanczos
 * steps inside one persistent launch, one workgroup per instance, beside the
 * per-phase kernels of the larger ones (0: every instance tiled). */
int mb_batch_set_persistent(mb_batch *b, int64_t max_dim);
/* Times k_probe steps for each candidate mapping and keeps the fastest. */
int mb_batch_autotune(mb_batch *b, int k_probe, int *tile, int *block, mb_stream stream);
/* k steps of lanczos_pass_one for every instance (optionally replayed as one
 * hipGraph).  alphas/betas: n*k (instance-major), steps/b_norms: n. */
int mb_batch_lanczos(mb_batch *b, int k, int use_graph, double *alphas, double *betas,
                     int *steps, double *b_norms, mb_stream stream);
/* Lowest Ritz value of every instance after k steps (FP64 tridiagonal solve on
 * the host), as find_min_eigenpair's value (lanczos.h:283-290). */
int mb_batch_ground_energies(mb_batch *b, int k, int use_graph, double *e0, mb_stream stream);

/* ------------------------------------------------------------ serving loop */

/* Continuous admission of mixed requests (no reference counterpart: the
 * reference solves one model at a time).  `slots` requests run concurrently,
 * each on its own stream; a request is admitted when it has arrived, a slot is
 * free and its device footprint fits in the free part of the server's arena
 * (allocated once at creation).  All of a request's state stays on the device
 * until it finishes.  Results do not depend on what runs beside a request. */
#define MB_REQ_GROUND 0 /* FP64 Lanczos, `steps` steps, lowest Ritz value (find_min_eigenpair) */
#define MB_REQ_KPM 1    /* FP64 Lanczos bounds (`steps`), then `moments` FP32 Chebyshev moments */
#define MB_REQ_EVOLVE 2 /* exp(-iHt) psi0 for t = 0 .. t_final in steps dt, Krylov dim `steps` */

typedef struct mb_request {
  int kind;             /* MB_REQ_* */
  int ndim;             /* 1, 2 or 3: periodic chain / square / cubic lattice */
  int sizes[3];         /* lattice sizes (first ndim used) */
  int n_up, n_dn;
  double t, u;
  int steps;            /* see MB_REQ_* */
  int moments, nvec;    /* KPM: even number of moments, random-phase vectors */
  double t_final, dt;   /* EVOLVE */
  uint64_t seed;        /* start vector: N(0,1) per reference index (mb_fill_gaussian_c128) */
  double arrival;       /* seconds after mb_server_create; < 0: on submission */
} mb_request;

typedef struct mb_result {
  int status;           /* MB_OK or the request's error code */
  int slot;
  int64_t dim, footprint, peak; /* D; admitted device bytes; bytes actually used */
  double e0;            /* GROUND */
  double e_min, e_max;  /* KPM: spectral bounds used for the rescaling */
  double overlap_re, overlap_im, norm; /* EVOLVE: <psi0|psi(t)> / |psi0|^2, |psi(t)| / |psi0| */
  double arrival, admitted, finished;  /* seconds after mb_server_create */
  int n_moments;        /* KPM: moments written to the caller's array */
  char error[256];
} mb_result;

typedef struct mb_server mb_server;
/* lane_bytes > 0 reserves that much of the arena for the Lanczos lane: ground-
 * state requests small enough for it (dimension up to an eighth of a lane
 * buffer, at most 2^20, and at most 512 steps) join a persistent batch at the
 * next Lanczos step and share its launches with every other member, instead of
 * taking a slot; they leave after their last step. */
int mb_server_create(int slots, int64_t arena_bytes, int64_t lane_bytes, mb_server **out);
/* Finishes every submitted request, then frees the arena. */
void mb_server_destroy(mb_server *s);
int mb_server_submit(mb_server *s, const mb_request *r, int64_t *id);
/* Blocks until request id has finished; moments (may be NULL) receives up to
 * moments_cap KPM moments. */
int mb_server_wait(mb_server *s, int64_t id, mb_result *out, double *moments, int moments_cap);
int mb_server_drain(mb_server *s);
double mb_server_now(const mb_server *s);
/* Device bytes a request is admitted with (-1 if the request is invalid). */
int64_t mb_request_footprint(const mb_request *r);

#ifdef __cplusplus
}
#endif
#endif
