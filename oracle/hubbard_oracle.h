/* Oracle: CPU restatement of keyehzy/manybody's Hubbard H.v path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library never links or calls it.
 *
 * Every function restates a reference routine with a deliberately different
 * mechanism from the product (explicit DFS instead of closed-form ranks,
 * explicit Fock-space operator action instead of popcount signs, hash lookups
 * instead of combinatorial ranks, Jacobi/bisection instead of implicit QR), so
 * that agreement between the two is evidence rather than a tautology.  How each
 * function is pinned (golden fixtures from the compiled reference, closed forms,
 * numpy/scipy) is listed in tests/test_oracle.py.
 *
 * Complex vectors are interleaved (re, im) doubles, index = reference basis
 * index unless a function says "product order".
 */
#ifndef HUBBARD_ORACLE_H
#define HUBBARD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* basis.cpp:13-26 */
uint64_t orc_choose(uint64_t n, uint64_t m);

/* BasisImpl::with_fixed_particle_number_and_spin + its DFS
 * (basis.cpp:91-126, 267-302).  Writes states in reference index order as
 * occupation bitmasks.  up/dn may be NULL to count only.  Returns the number of
 * states, or -1 if cap is too small. */
int64_t orc_basis_enumerate(int n_sites, int n_up, int n_dn, uint64_t *up, uint64_t *dn,
                            int64_t cap);

/* The kinetic-term bonds of HubbardModel / HubbardModel2D / HubbardModel3D in
 * the reference's loop order (hubbard_model.h:16-20, 72-82, 144-159); site =
 * x + Lx*y + Lx*Ly*z (index.h:82-103).  Each bond (from[k], to[k]) stands for
 * -t * hopping(from, to, spin) for both spins.  Returns the number of bonds or
 * -1 if cap is too small. */
int orc_hubbard_bonds(int ndim, const int *sizes, int *from, int *to, int cap);

/* Column-parallel matrix build with one hash lookup per nonzero, the cost
 * structure of compute_matrix_elements<sp_cx_mat> (matrix_elements.h:74-105).
 * The action of each hopping monomial on a basis key is evaluated operator by
 * operator on the ordered creator string (canonicalize.h:19-78).  Output is CSC
 * with sorted row indices and complex values.  With row_idx == NULL only nnz is
 * returned.  Returns nnz or -1 on overflow / missing state. */
int64_t orc_build_csc(int n_sites, const uint64_t *up, const uint64_t *dn, int64_t D,
                      int n_bonds, const int *from, const int *to, double t, double U,
                      int64_t *col_ptr, int64_t *row_idx, double *val, int64_t cap, int threads);

/* Threads of the OpenMP loops that follow (the vector passes of the Lanczos,
 * Krylov and KPM routines; the CSC product itself is always one thread, as the
 * reference configures Armadillo, src/CMakeLists.txt:20). */
void orc_set_threads(int threads);

/* every > 0: the Lanczos routines print a line to stderr every `every` steps
 * (long host runs show that they progress); 0 (default): silent. */
void orc_set_progress(int every);

/* y = A x with A in CSC, complex values, column scatter order (Armadillo
 * sp_cx_mat * cx_vec, linear_operator.h:87). */
void orc_spmv_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx, const double *val,
                  const double *x, double *y);

/* Single-species hop table in ascending-bitstring order (product order).
 * strings[] must hold C(n_sites, n) entries; rows are CSR (row_ptr[r] ..
 * row_ptr[r+1]) with target row and real amplitude, found by binary search.
 * diag_onsite[r] receives the one-body diagonal (from self bonds).  Returns nnz
 * or -1 on overflow. */
int64_t orc_species_table(int n_sites, int n, int n_bonds, const int *from, const int *to,
                          double t, uint64_t *strings, int64_t *row_ptr, int32_t *cols,
                          double *vals, double *diag_onsite, int64_t cap);

/* Matrix-free y = H x in product order x[u * Ddn + d] (complex interleaved).
 * diagU[k] is the interaction energy of k doubly occupied sites. */
void orc_hv_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up, const uint64_t *s_dn,
                    const int64_t *up_ptr, const int32_t *up_cols, const double *up_vals,
                    const double *up_onsite, const int64_t *dn_ptr, const int32_t *dn_cols,
                    const double *dn_vals, const double *dn_onsite, const double *diagU,
                    const double *x, double *y, int threads);

/* Interaction table: diagU[k] = U added k times, sequentially, as
 * ExpressionBase::add_to accumulates it (expression_base.h:85-97). */
void orc_diag_table(double U, int n_max, double *diagU);

/* Lanczos pass one (lanczos.h:15-40, 104-175) on the CSC operator, with or
 * without negation (Negated<Op>, linear_operator.h:93-105).  b is complex.
 * Returns steps_taken, or -1 for a zero-norm start. */
int orc_lanczos_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                    const double *val, int negate, const double *b, int k, double *alphas,
                    double *betas, double *b_norm);

/* Same recurrence on the product-order matrix-free operator. */
int orc_lanczos_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up, const uint64_t *s_dn,
                        const int64_t *up_ptr, const int32_t *up_cols, const double *up_vals,
                        const double *up_onsite, const int64_t *dn_ptr, const int32_t *dn_cols,
                        const double *dn_vals, const double *dn_onsite, const double *diagU,
                        const double *b, int k, double *alphas, double *betas, double *b_norm,
                        int threads);

/* Eigenvalues of the symmetric tridiagonal T (diag a[0..m), off b[0..m-1)) by
 * Sturm-sequence bisection, ascending. */
void orc_tridiag_eigvals(int m, const double *a, const double *b, double *w);

/* Full eigen-decomposition of a dense symmetric m x m matrix (row-major) by
 * cyclic Jacobi rotations.  w ascending; V[i*m + j] = component i of vector j. */
void orc_jacobi_eig(int m, const double *A, double *w, double *V);

/* Lanczos pass two (lanczos.h:177-218) on the CSC operator: x = sum y_j v_j.
 * y is real (length steps).  Output x is complex. */
void orc_lanczos_pass_two_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                              const double *val, int negate, const double *b, int steps,
                              const double *alphas, const double *betas, double b_norm,
                              const double *y, double *x);

/* psi(t) = exp(-i t H) psi (evolve_state.h:22-61, 120-178) on the CSC
 * operator with k Krylov steps (k = min(k, D); 0 means D). */
void orc_time_evolve_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                         const double *val, const double *psi, double t, int k, double *out);

/* The same on the matrix-free product-order operator of orc_hv_product. */
void orc_time_evolve_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up,
                             const uint64_t *s_dn, const int64_t *up_ptr, const int32_t *up_cols,
                             const double *up_vals, const double *up_onsite,
                             const int64_t *dn_ptr, const int32_t *dn_cols, const double *dn_vals,
                             const double *dn_onsite, const double *diagU, const double *psi,
                             double t, int k, double *out, int threads);

/* Chebyshev moments mu_n = (1/R) sum_r <r|T_n(Hs)|r> / D for Hs = (H - b)/a,
 * evaluated by the plain three-term recurrence (kpm.h:212-229) with one
 * inner product per order, FP64.  vecs holds R complex start vectors. */
void orc_kpm_moments_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                         const double *val, double a, double b, int M, int R,
                         const double *vecs, double *mu);

/* The same on the matrix-free product-order operator of orc_hv_product. */
void orc_kpm_moments_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up,
                             const uint64_t *s_dn, const int64_t *up_ptr, const int32_t *up_cols,
                             const double *up_vals, const double *up_onsite,
                             const int64_t *dn_ptr, const int32_t *dn_cols, const double *dn_vals,
                             const double *dn_onsite, const double *diagU, double a, double b,
                             int M, int R, const double *vecs, double *mu, int threads);

/* Jackson kernel g_0..g_N (kpm.h:19-30). */
void orc_jackson_kernel(int N, double *g);

#ifdef __cplusplus
}
#endif
#endif
