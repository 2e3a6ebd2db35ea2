// Krylov consumers of H.v on the device: Lanczos pass one / pass two, the
// lowest eigenpair, Krylov exp(-iHt), and KPM Chebyshev moments; plus the FP64
// symmetric tridiagonal eigensolver on the host.
//
// Reference: src/numerics/lanczos.h:15-290, src/numerics/evolve_state.h:22-231,
// src/numerics/kpm.h:19-30, 184-241.  The reference keeps its vectors on the
// host and calls LAPACK (arma::eig_sym) for T_k; here the vectors stay in HBM,
// alpha/beta are reduced on the device without host round trips, and T_k is
// solved on the host by implicit QL.
#include "krylov.hpp"

#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <complex>
#include <type_traits>

namespace mb {

// ------------------------------------------------------ tridiagonal eig

// Implicit QL with Wilkinson-style shifts on a symmetric tridiagonal matrix
// (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.3).  d = diagonal,
// e = off-diagonal.  Eigenvectors accumulate into Z (column-major m x m) when
// Z != nullptr.
void tridiag_eig(int m, const double *a, const double *b, double *w, double *V) {
  if (m <= 0) return;
  std::vector<double> d(a, a + m), e(m, 0.0);
  for (int i = 0; i + 1 < m; ++i) e[i] = b[i];
  std::vector<double> Z;
  if (V) {
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[i + (size_t)m * i] = 1.0;
  }
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    while (true) {
      int n = l;
      for (; n + 1 < m; ++n) {
        const double dd = std::fabs(d[n]) + std::fabs(d[n + 1]);
        if (std::fabs(e[n]) <= eps * dd) break;
      }
      if (n == l) break;
      if (++iter > 100) throw std::runtime_error("Eigenvalue decomposition of T_k failed");
      // shift from the trailing 2x2 of the active block [l, n]
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[n] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = n - 1;
      bool deflated = false;
      for (; i >= l; --i) {
        const double f = s * e[i];
        const double h = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[n] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * h;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - h;
        if (V) {
          double *zi = &Z[(size_t)m * i], *zi1 = &Z[(size_t)m * (i + 1)];
          for (int k = 0; k < m; ++k) {
            const double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[n] = 0.0;
    }
  }
  // ascending order
  std::vector<int> idx(m);
  for (int i = 0; i < m; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int j = 0; j < m; ++j) {
    w[j] = d[idx[j]];
    if (V)
      for (int k = 0; k < m; ++k) V[k + (size_t)m * j] = Z[k + (size_t)m * idx[j]];
  }
}

// Lowest eigenvalue of a symmetric tridiagonal matrix by Sturm-sequence
// bisection (Golub & Van Loan sec. 8.4.1): O(m) per bisection step, to the
// last bit; used when only the value is needed.
double tridiag_min_eig(int m, const double *a, const double *b) {
  if (m <= 0) return 0.0;
  if (m == 1) return a[0];
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    double r = 0.0;
    if (i > 0) r += std::fabs(b[i - 1]);
    if (i + 1 < m) r += std::fabs(b[i]);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  const double pad = 1e-12 * (std::fabs(lo) + std::fabs(hi) + 1.0);
  lo -= pad;
  hi += pad;
  auto count_below = [&](double x) {
    int c = 0;
    double q = a[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < m; ++i) {
      if (q == 0.0) q = 1e-300;
      q = a[i] - x - b[i - 1] * b[i - 1] / q;
      if (q < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// ------------------------------------------------------------ Lanczos

LanczosWork::LanczosWork(int64_t D, int k, int64_t nparts, size_t esz, bool vectors)
    : D(D), k(k), esz(esz),
      scal(sizeof(double) *
               (2 * (size_t)std::max(k, 1) + 2 + (size_t)std::max<int64_t>(kRedBlocks, nparts)) +
           16) {
  if (vectors) {
    v.alloc(D * 2 * esz);
    vp.alloc(D * 2 * esz);
    w.alloc(D * 2 * esz);
  }
  double *base = scal.as<double>();
  st.alpha = base;
  st.beta = base + std::max(k, 1);
  st.bnorm = base + 2 * std::max(k, 1);
  st.partial = base + 2 * std::max(k, 1) + 2;
  st.flags = reinterpret_cast<int *>(base + 2 * std::max(k, 1) + 1);
}

bool keep_krylov_basis(size_t bytes) {
  if (thread_dev_alloc()) return true;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return bytes <= free_b / 2;
}

// Pass one of a real-space operator whose H.v runs in the block-staged up
// kernel and the sorted row kernel, with the row pass first: step i's row pass
// stages x_i = w - c x_{i-1} (the previous step's axpy, the same arithmetic,
// done while the row is staged), writes x_i and z = (down hops + diagonal) x_i
// and the row sums of |x_i|^2, which give beta_{i-1}; the up pass then forms
// H x_i = (up hops) x_i + z and applies the Lanczos epilogue.  No separate
// axpy pass: per step the row pass streams w, x_{i-1}, x_i, z and the up pass
// x_i, z, x_{i-1}, w (8 vector passes against 9).  The same recurrence and the
// same per-element arithmetic as the up-then-row step; the alpha and beta sums
// run in another order.  The vectors stay unnormalised (x_i = beta_{i-1} v_i).
void pass_one_fold(const HvTables &t, LanczosWork &W, int k, hipStream_t s) {
  const int64_t D = W.D;
  if (!W.z.p) W.z.alloc((size_t)D * 2 * sizeof(double));
  const size_t np = (size_t)lanczos_fold_partials(t);
  if (W.fold_part.bytes < np * sizeof(double)) W.fold_part.alloc(np * sizeof(double));
  LanczosState st = W.st;
  st.partial = W.fold_part.as<double>();
  double *cur = W.v.as<double>(), *prv = W.vp.as<double>(), *wr = W.w.as<double>(), *z = W.z.as<double>();
  for (int i = 0; i < k; ++i) {
    EpiArgs ra;
    ra.flags = st.flags;
    if (i > 0) {
      ra.fold = 1;
      ra.fold_w = wr;
      ra.fold_v = cur;
      ra.fold_x = prv;  // x_{i-2}'s buffer receives x_i
      ra.fold_partial = st.partial + kRedBlocks;  // summed by launch_lanczos_finish_beta_n
      ra.alpha_dev = st.alpha;
      ra.beta_dev = st.beta;
      ra.fold_step = i - 1;
    }
    launch_row_down(t, cur, z, ra, s);
    if (i > 0) {
      launch_lanczos_finish_beta_n(st, i - 1, 2 * t.d_up, s);
      std::swap(cur, prv);  // cur = x_i, prv = x_{i-1}
    }
    EpiArgs ua;
    ua.prev = prv;
    ua.beta_dev = st.beta;
    ua.flags = st.flags;
    ua.step = i;
    ua.partial = st.partial;
    ua.lazy = 1;
    const int64_t nparts = launch_up_lanczos(t, cur, z, wr, ua, s);
    launch_lanczos_finish_alpha(st, i, nparts, s);
  }
}

// Pass one with Krylov vectors and H.v in T (double, or float for the FP32
// Krylov option); b is FP64, the norm and the alpha/beta reductions are FP64.
// With `basis` (k + 1 vectors of 2 D values of T) the recurrence runs in it:
// step i reads v_i = basis[i] and v_{i-1}, and writes v_{i+1} into basis[i + 1],
// so the Krylov vectors stay in HBM for the reconstruction.
template <typename T>
LanczosResult pass_one_t(const HvOp &op, const void *b, int k, hipStream_t s, LanczosWork *work,
                         T *basis = nullptr) {
  const int64_t D = op.dim();
  std::unique_ptr<LanczosWork> own;
  if (!work || work->D != D || work->k < k || work->esz != sizeof(T) || (!basis && !work->v.p)) {
    own.reset(new LanczosWork(D, k, op.partials(), sizeof(T), basis == nullptr));
    work = own.get();
  }
  LanczosWork &W = *work;
  LanczosResult res;
  launch_norm((const double *)b, 2 * D, W.st.partial, W.st.bnorm, s);
  MB_HIP(hipMemcpyAsync(&res.b_norm, W.st.bnorm, sizeof(double), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (res.b_norm <= kTol) throw std::runtime_error("Input vector has zero norm");
  if (k == 0) return res;
  T *v = basis ? basis : W.v.as<T>(), *vp = basis ? basis : W.vp.as<T>(),
    *w = basis ? basis + 2 * D : W.w.as<T>();
  launch_scale_div<T>((const double *)b, W.st.bnorm, v, 2 * D, s);
  // step 0 does not read v_prev (beta_{-1} = 0); the scratch copy is cleared anyway
  if (!basis) launch_zero(vp, (size_t)D * 2 * sizeof(T), s);
  launch_zero(W.st.flags, 2 * sizeof(int), s);
  // Without a stored basis the vectors are kept unnormalised (w is not divided
  // by beta; the next step's epilogue and axpy apply 1 / beta instead), which
  // saves one pass over w per step.  A stored basis is normalised as it is built.
  const bool lazy = basis == nullptr;
  // the row-first step needs a fourth vector, which request regions and the
  // no-allocation scratch do not budget for: there the step keeps its axpy pass
  const HvTables *ht = op.hubbard_tables();
  const bool fold = std::is_same<T, double>::value && lazy && ht && !thread_dev_alloc() &&
                    lanczos_fold_ok(*ht);
  if (fold) pass_one_fold(*ht, W, k, s);
  for (int i = 0; i < (fold ? 0 : k); ++i) {
    if (basis) {
      v = basis + (int64_t)i * 2 * D;
      vp = i > 0 ? v - 2 * D : v;
      w = v + 2 * D;
    }
    EpiArgs a;
    a.prev = vp;
    a.beta_dev = W.st.beta;
    a.flags = W.st.flags;
    a.step = i;
    a.partial = W.st.partial;
    a.lazy = lazy ? 1 : 0;
    const int64_t slots = op.hv(v, w, Epi::Lanczos, a, s);
    launch_lanczos_finish_alpha(W.st, i, 2 * slots, s);
    launch_lanczos_axpy_norm<T>(v, w, 2 * D, W.st, i, s, lazy);
    launch_lanczos_finish_beta(W.st, i, s);
    if (!lazy) launch_lanczos_normalize<T>(w, 2 * D, W.st, i, s);
    if (!basis) {
      T *tmp = vp;
      vp = v;
      v = w;
      w = tmp;
    }
  }
  std::vector<double> al(k), be(k);
  int flags[2];
  MB_HIP(hipMemcpyAsync(al.data(), W.st.alpha, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(be.data(), W.st.beta, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(flags, W.st.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  res.steps = flags[1];
  res.alphas.assign(al.begin(), al.begin() + res.steps);
  res.betas.assign(be.begin(), be.begin() + std::max(res.steps - 1, 0));
  return res;
}

LanczosResult lanczos_pass_one(const HvOp &op, const void *b, int k, hipStream_t s,
                               LanczosWork *work) {
  return pass_one_t<double>(op, b, k, s, work);
}

// lanczos_pass_one with full reorthogonalisation (an option the reference does
// not have; its recurrence, lanczos.h:15-40, is kept step for step): after
// w -= alpha v, w is orthogonalised twice against every stored Lanczos vector
// (classical Gram-Schmidt twice, FP64 on the device), then beta = |w|.  All
// k + 1 vectors stay in HBM (16 D (k + 1) bytes); they are handed to `basis`
// when it is given.
LanczosResult lanczos_pass_one_reorth(const HvOp &op, const void *b, int k, hipStream_t s,
                                      DevBuf *basis) {
  const int64_t D = op.dim();
  const int64_t nparts = op.partials();
  const int kk = std::max(k, 1);
  DevBuf V((size_t)(kk + 1) * 16 * (size_t)D);
  DevBuf scal(sizeof(double) * (2 * (size_t)kk + 2 + (size_t)std::max<int64_t>(kRedBlocks, nparts)) + 16);
  DevBuf pp((size_t)2 * (kk + 1) * kProjBlocks * sizeof(double));
  DevBuf coef((size_t)2 * (kk + 1) * sizeof(double));
  LanczosState st;
  double *base = scal.as<double>();
  st.alpha = base;
  st.beta = base + kk;
  st.bnorm = base + 2 * kk;
  st.partial = base + 2 * kk + 2;
  st.flags = reinterpret_cast<int *>(base + 2 * kk + 1);
  LanczosResult res;
  launch_norm((const double *)b, 2 * D, st.partial, st.bnorm, s);
  MB_HIP(hipMemcpyAsync(&res.b_norm, st.bnorm, sizeof(double), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (res.b_norm <= kTol) throw std::runtime_error("Input vector has zero norm");
  if (k == 0) return res;
  double *Vb = V.as<double>();
  launch_scale_div((const double *)b, st.bnorm, Vb, 2 * D, s);
  MB_HIP(hipMemsetAsync(st.flags, 0, 2 * sizeof(int), s));
  for (int i = 0; i < k; ++i) {
    double *v = Vb + (int64_t)i * 2 * D, *w = v + 2 * D;
    EpiArgs a;
    a.prev = i > 0 ? v - 2 * D : v;
    a.beta_dev = st.beta;
    a.flags = st.flags;
    a.step = i;
    a.partial = st.partial;
    const int64_t slots = op.hv(v, w, Epi::Lanczos, a, s);
    launch_lanczos_finish_alpha(st, i, 2 * slots, s);
    launch_lanczos_axpy_norm(v, w, 2 * D, st, i, s);
    for (int pass = 0; pass < 2; ++pass)
      launch_reorth(Vb, i + 1, w, D, pp.as<double>(), coef.as<double>(), st.flags, s);
    launch_sumsq_partial(w, 2 * D, st.partial, s);
    launch_lanczos_finish_beta(st, i, s);
    launch_lanczos_normalize(w, 2 * D, st, i, s);
  }
  std::vector<double> al(k), be(k);
  int flags[2];
  MB_HIP(hipMemcpyAsync(al.data(), st.alpha, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(be.data(), st.beta, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(flags, st.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  res.steps = flags[1];
  res.alphas.assign(al.begin(), al.begin() + res.steps);
  res.betas.assign(be.begin(), be.begin() + std::max(res.steps - 1, 0));
  if (basis) *basis = std::move(V);
  return res;
}

// Lowest eigenpair from the reorthogonalised decomposition: x = |b| sum_j y_j V_j
// straight from the stored vectors (no second pass).
double find_min_eigenpair_reorth(const HvOp &op, const void *b, int k, void *x, hipStream_t s) {
  const int64_t D = op.dim();
  DevBuf V;
  const LanczosResult dec = lanczos_pass_one_reorth(op, b, k, s, &V);
  const int m = dec.steps;
  if (m == 0) {
    if (x) MB_HIP(hipMemsetAsync(x, 0, (size_t)D * 16, s));
    MB_HIP(hipStreamSynchronize(s));
    return 0.0;
  }
  if (!x) return tridiag_min_eig(m, dec.alphas.data(), dec.betas.data());
  std::vector<double> wv(m), Q((size_t)m * m);
  tridiag_eig(m, dec.alphas.data(), dec.betas.data(), wv.data(), Q.data());
  std::vector<double> c(2 * (size_t)m, 0.0);
  for (int j = 0; j < m; ++j) c[2 * j] = -Q[j] * dec.b_norm;  // x -= c_j V_j with x = 0
  DevBuf dc = upload(c);
  DevBuf zero_flag = upload(std::vector<int>(2, 0));
  MB_HIP(hipMemsetAsync(x, 0, (size_t)D * 16, s));
  launch_reorth_sub(V.as<double>(), m, dc.as<double>(), (double *)x, D, zero_flag.as<int>(), s);
  MB_HIP(hipStreamSynchronize(s));
  return wv[0];
}

// Pass two: regenerates the Lanczos vectors (in T) and accumulates
// x = sum_j y_j v_j in FP64.
template <typename T>
void pass_two_t(const HvOp &op, const void *b, const LanczosResult &dec,
                const std::vector<std::complex<double>> &y, void *x, hipStream_t s, LanczosWork *work) {
  const int64_t D = op.dim();
  if ((int)y.size() != dec.steps)
    throw std::runtime_error("Dimension mismatch: y_k size must match decomposition steps");
  if (dec.b_norm <= kTol) throw std::runtime_error("Input vector has zero norm");
  if (dec.steps == 0) {
    MB_HIP(hipMemsetAsync(x, 0, (size_t)D * 16, s));
    return;
  }
  std::unique_ptr<LanczosWork> own;
  if (!work || work->D != D || work->esz != sizeof(T) || !work->v.p) {
    own.reset(new LanczosWork(D, 1, op.partials(), sizeof(T)));
    work = own.get();
  }
  LanczosWork &W = *work;
  MB_HIP(hipMemcpyAsync(W.st.bnorm, &dec.b_norm, sizeof(double), hipMemcpyHostToDevice, s));
  T *v = W.v.as<T>(), *vp = W.vp.as<T>(), *w = W.w.as<T>();
  launch_scale_div<T>((const double *)b, W.st.bnorm, v, 2 * D, s);
  MB_HIP(hipMemsetAsync(vp, 0, (size_t)D * 2 * sizeof(T), s));
  launch_scale_complex<T>(v, y[0].real(), y[0].imag(), (double *)x, D, s);
  for (int j = 0; j + 1 < dec.steps; ++j) {
    const double aj = dec.alphas[j], bj = dec.betas[j], bp = j == 0 ? 0.0 : dec.betas[j - 1];
    EpiArgs a;
    a.prev = vp;
    a.alpha = aj;
    a.beta = bp;
    op.hv(v, w, Epi::Recon, a, s);
    launch_recon_update<T>(w, (double *)x, bj, y[j + 1].real(), y[j + 1].imag(), D, s);
    T *tmp = vp;
    vp = v;
    v = w;
    w = tmp;
  }
  MB_HIP(hipStreamSynchronize(s));
}

void lanczos_pass_two(const HvOp &op, const void *b, const LanczosResult &dec,
                      const std::vector<std::complex<double>> &y, void *x, hipStream_t s,
                      LanczosWork *work) {
  pass_two_t<double>(op, b, dec, y, x, s, work);
}

// x = sum_j y_j V_j from the stored Krylov vectors (coefficients uploaded on s)
template <typename T>
void combine_basis(const T *V, const std::vector<std::complex<double>> &y, int64_t D, void *x,
                   hipStream_t s) {
  std::vector<double> yc(2 * y.size());
  for (size_t j = 0; j < y.size(); ++j) {
    yc[2 * j] = y[j].real();
    yc[2 * j + 1] = y[j].imag();
  }
  DevBuf yd(yc.size() * sizeof(double));
  MB_HIP(hipMemcpyAsync(yd.p, yc.data(), yc.size() * sizeof(double), hipMemcpyHostToDevice, s));
  launch_basis_combine<T>(V, (int)y.size(), D, yd.as<double>(), (double *)x, s);
  MB_HIP(hipStreamSynchronize(s));
}

double find_min_eigenpair(const HvOp &op, const void *b, int k, void *x, hipStream_t s) {
  const int64_t D = op.dim();
  // the eigenvector from the k + 1 stored Krylov vectors where they fit in HBM,
  // else by regenerating them (lanczos_pass_two)
  DevBuf V;
  if (x && k > 0) {
    const size_t bytes = (size_t)(k + 1) * 2 * D * sizeof(double);
    if (keep_krylov_basis(bytes)) {
      try {
        V.alloc(bytes);
      } catch (const NoMemError &) {
        V.release();
      }
    }
  }
  LanczosWork work(D, k, op.partials(), sizeof(double), V.p == nullptr);
  LanczosResult dec = V.p ? pass_one_t<double>(op, b, k, s, &work, V.as<double>())
                          : lanczos_pass_one(op, b, k, s, &work);
  const int m = dec.steps;
  if (m == 0) {
    if (x) MB_HIP(hipMemsetAsync(x, 0, (size_t)op.dim() * 16, s));
    MB_HIP(hipStreamSynchronize(s));
    return 0.0;
  }
  if (!x) return tridiag_min_eig(m, dec.alphas.data(), dec.betas.data());
  std::vector<double> wv(m), Z((size_t)m * m);
  tridiag_eig(m, dec.alphas.data(), dec.betas.data(), wv.data(), Z.data());
  {
    std::vector<std::complex<double>> y(m);
    for (int i = 0; i < m; ++i) y[i] = Z[i] * dec.b_norm;  // column 0: lowest
    if (V.p)
      combine_basis<double>(V.as<double>(), y, D, x, s);
    else
      lanczos_pass_two(op, b, dec, y, x, s, &work);
  }
  return wv[0];
}

// y = exp(-i t T) e1 (evolve_state.h:22-61)
std::vector<std::complex<double>> expm_first_column(const LanczosResult &dec, double t) {
  const int m = dec.steps;
  std::vector<double> wv(m), Q((size_t)m * m);
  tridiag_eig(m, dec.alphas.data(), dec.betas.data(), wv.data(), Q.data());
  std::vector<std::complex<double>> y(m, 0.0);
  for (int j = 0; j < m; ++j) {
    const std::complex<double> ph = std::exp(std::complex<double>(0.0, -t) * wv[j]);
    const double q0 = Q[0 + (size_t)m * j];
    for (int i = 0; i < m; ++i) y[i] += Q[i + (size_t)m * j] * (ph * q0);
  }
  return y;
}

// `cache` (optional) keeps the stored Krylov basis between the steps of
// time_evolve_steps; without one a basis is allocated for this call.
// y = exp(-tau T) e1 (expm_tridiagonal_first_column_imaginary, evolve_state.h:64-105)
std::vector<std::complex<double>> expm_first_column_imag(const LanczosResult &dec, double tau) {
  const int m = dec.steps;
  std::vector<double> wv(m), Q((size_t)m * m);
  tridiag_eig(m, dec.alphas.data(), dec.betas.data(), wv.data(), Q.data());
  std::vector<std::complex<double>> y(m, 0.0);
  for (int j = 0; j < m; ++j) {
    const double f = std::exp(-tau * wv[j]) * Q[0 + (size_t)m * j];
    for (int i = 0; i < m; ++i) y[i] += Q[i + (size_t)m * j] * f;
  }
  return y;
}

template <typename T>
void time_evolve_t(const HvOp &op, const void *psi, double t, int krylov_steps, void *out,
                   hipStream_t s, DevBuf *cache = nullptr, bool imag = false) {
  const int64_t D = op.dim();
  int k = krylov_steps;
  if (k == 0 || k > D) k = (int)D;
  // the k + 1 Krylov vectors stay in HBM where they fit: psi(t) is then one
  // pass over them instead of k - 1 more H.v (evolve_state.h:120-151 recomputes)
  DevBuf own_basis;
  DevBuf *basis = cache ? cache : &own_basis;
  const size_t bytes = (size_t)(k + 1) * 2 * D * sizeof(T);
  if (basis->bytes < bytes && keep_krylov_basis(bytes)) {
    try {
      basis->alloc(bytes);
    } catch (const NoMemError &) {
      basis->release();
    }
  }
  const bool stored = basis->p && basis->bytes >= bytes;
  LanczosWork work(D, k, op.partials(), sizeof(T), !stored);
  // norm first: a zero state is returned unchanged (evolve_state.h:131-134)
  double nrm = 0.0;
  launch_norm((const double *)psi, 2 * D, work.st.partial, work.st.bnorm, s);
  MB_HIP(hipMemcpyAsync(&nrm, work.st.bnorm, sizeof(double), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (nrm <= kTol) {
    if (out != psi) MB_HIP(hipMemcpyAsync(out, psi, (size_t)D * 16, hipMemcpyDeviceToDevice, s));
    MB_HIP(hipStreamSynchronize(s));
    return;
  }
  LanczosResult dec = pass_one_t<T>(op, psi, k, s, &work, stored ? basis->as<T>() : nullptr);
  if (dec.steps == 0) {
    MB_HIP(hipMemsetAsync(out, 0, (size_t)D * 16, s));
    MB_HIP(hipStreamSynchronize(s));
    return;
  }
  auto y = imag ? expm_first_column_imag(dec, t) : expm_first_column(dec, t);
  for (auto &v : y) v *= dec.b_norm;
  if (stored) {
    combine_basis<T>(basis->as<T>(), y, D, out, s);  // psi is no longer read: out may be psi
  } else if (out == psi) {
    DevBuf tmp((size_t)D * 16);
    pass_two_t<T>(op, psi, dec, y, tmp.p, s, &work);
    MB_HIP(hipMemcpyAsync(out, tmp.p, (size_t)D * 16, hipMemcpyDeviceToDevice, s));
    MB_HIP(hipStreamSynchronize(s));
  } else {
    pass_two_t<T>(op, psi, dec, y, out, s, &work);
  }
}

void time_evolve_kind(const HvOp &op, const void *psi, double t, int krylov_steps, void *out,
                      hipStream_t s, DevBuf *cache, bool imag) {
  time_evolve_t<double>(op, psi, t, krylov_steps, out, s, cache, imag);
}

void time_evolve(const HvOp &op, const void *psi, double t, int krylov_steps, void *out,
                 hipStream_t s, bool fp32) {
  if (fp32)
    time_evolve_t<float>(op, psi, t, krylov_steps, out, s);
  else
    time_evolve_t<double>(op, psi, t, krylov_steps, out, s);
}

// The device part of one Krylov step with a stored basis, all on the stream:
// norm of b, v_0 = b / |b|, then k Lanczos steps writing basis[1 .. k], with
// alpha, beta, |b| and the step count left in W (the kernels and their order
// are those of pass_one_t).  No host synchronisation, so it can be captured.
template <typename T>
void enqueue_lanczos_basis(const HvOp &op, const void *b, int k, hipStream_t s, LanczosWork &W,
                           T *basis) {
  const int64_t D = op.dim();
  launch_norm((const double *)b, 2 * D, W.st.partial, W.st.bnorm, s);
  launch_scale_div<T>((const double *)b, W.st.bnorm, basis, 2 * D, s);
  launch_zero(W.st.flags, 2 * sizeof(int), s);
  for (int i = 0; i < k; ++i) {
    T *v = basis + (int64_t)i * 2 * D, *vp = i > 0 ? v - 2 * D : v, *w = v + 2 * D;
    EpiArgs a;
    a.prev = vp;
    a.beta_dev = W.st.beta;
    a.flags = W.st.flags;
    a.step = i;
    a.partial = W.st.partial;
    a.lazy = 0;
    const int64_t slots = op.hv(v, w, Epi::Lanczos, a, s);
    launch_lanczos_finish_alpha(W.st, i, 2 * slots, s);
    launch_lanczos_axpy_norm<T>(v, w, 2 * D, W.st, i, s, false);
    launch_lanczos_finish_beta(W.st, i, s);
    launch_lanczos_normalize<T>(w, 2 * D, W.st, i, s);
  }
}

// time_evolve_steps with the device part of every step captured once in a
// hipGraph and replayed: a step of a small sector is a few hundred launches
// (k Lanczos steps of five kernels), which the graph issues without the host.
// The state `out`, the stored basis and the workspace keep their addresses
// across the steps, so one graph serves all of them; the host reads alpha,
// beta, |b| and the step count after each replay, forms exp(-i tau T) e1 and
// enqueues the combine.  Same kernels in the same order as the eager steps,
// so the same results.  Returns false (nothing done) when the stored basis does
// not fit or the capture fails; MANYBODY_AMD_KRYLOV_GRAPH=0 turns it off.
template <typename T>
bool time_evolve_steps_graph(const HvOp &op, double t0, double t1, double dt, int krylov_steps,
                             void *out, hipStream_t s, const std::function<void(double)> &report) {
  const char *e = std::getenv("MANYBODY_AMD_KRYLOV_GRAPH");
  if (e && std::strcmp(e, "0") == 0) return false;
  const int64_t D = op.dim();
  int k = krylov_steps;
  if (k == 0 || k > D) k = (int)D;
  if (k < 1) return false;
  const size_t bytes = (size_t)(k + 1) * 2 * D * sizeof(T);
  if (!keep_krylov_basis(bytes)) return false;
  DevBuf basis;
  try {
    basis.alloc(bytes);
  } catch (const NoMemError &) {
    return false;
  }
  LanczosWork W(D, k, op.partials(), sizeof(T), false);
  DevBuf coef((size_t)2 * k * sizeof(double));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  bool ok = true;
  try {
    enqueue_lanczos_basis<T>(op, out, k, s, W, basis.as<T>());
  } catch (const std::exception &) {
    ok = false;
  }
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (!ok || ec != hipSuccess || !graph ||
      hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    if (graph) (void)hipGraphDestroy(graph);
    return false;
  }
  struct Guard {
    hipGraph_t g;
    hipGraphExec_t x;
    ~Guard() {
      (void)hipGraphExecDestroy(x);
      (void)hipGraphDestroy(g);
    }
  } guard{graph, exec};
  const double direction = (t1 >= t0) ? 1.0 : -1.0;
  dt = std::fabs(dt) * direction;
  double t = t0;
  report(t);
  std::vector<double> scal(2 * (size_t)k + 2), yc(2 * (size_t)k);
  while ((direction > 0 && t < t1) || (direction < 0 && t > t1)) {
    const double step = (direction > 0) ? std::min(dt, t1 - t) : std::max(dt, t1 - t);
    MB_HIP(hipGraphLaunch(exec, s));
    // alpha[0..k), beta[0..k), then |b| and the flags in the next two words
    MB_HIP(hipMemcpyAsync(scal.data(), W.st.alpha, scal.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    MB_HIP(hipStreamSynchronize(s));
    const double bnorm = scal[2 * k];
    int flags[2];
    std::memcpy(flags, &scal[2 * k + 1], sizeof(flags));
    if (bnorm > kTol && flags[1] > 0) {  // a zero state stays as it is (evolve_state.h:131-134)
      LanczosResult dec;
      dec.steps = flags[1];
      dec.b_norm = bnorm;
      dec.alphas.assign(scal.begin(), scal.begin() + dec.steps);
      dec.betas.assign(scal.begin() + k, scal.begin() + k + std::max(dec.steps - 1, 0));
      auto y = expm_first_column(dec, step);
      for (size_t j = 0; j < y.size(); ++j) {
        yc[2 * j] = y[j].real() * bnorm;
        yc[2 * j + 1] = y[j].imag() * bnorm;
      }
      MB_HIP(hipMemcpyAsync(coef.p, yc.data(), y.size() * 2 * sizeof(double), hipMemcpyHostToDevice, s));
      launch_basis_combine<T>(basis.as<T>(), (int)y.size(), D, coef.as<double>(), (double *)out, s);
    } else if (bnorm > kTol) {
      MB_HIP(hipMemsetAsync(out, 0, (size_t)D * 16, s));
    }
    t += step;
    report(t);
  }
  MB_HIP(hipStreamSynchronize(s));
  return true;
}

// FP64 time_evolve_steps for a real-space sector of at most kKrylovSmallMaxDim
// states: each step's k Lanczos steps run in one workgroup (launch_krylov_small,
// the vectors written into the stored basis), then exp(-i tau T) e1 on the host
// and the combine: three device operations per step where the general path
// has five launches per Lanczos step.  Its sums run in another order than the
// general path's, so the states agree to rounding.  Returns false (nothing
// done) for other operators and sizes; MANYBODY_AMD_KRYLOV_SMALL=0 turns it off.
bool time_evolve_steps_small(const HvOp &op, double t0, double t1, double dt, int krylov_steps,
                             void *out, hipStream_t s, const std::function<void(double)> &report) {
  const char *e = std::getenv("MANYBODY_AMD_KRYLOV_SMALL");
  if (e && std::strcmp(e, "0") == 0) return false;
  const HvTables *tb = op.hubbard_tables();
  const int64_t D = op.dim();
  if (!tb || D < 1 || D > kKrylovSmallMaxDim) return false;
  int k = krylov_steps;
  if (k == 0 || k > D) k = (int)D;
  DevBuf basis((size_t)(k + 1) * 2 * D * sizeof(double));
  DevBuf scal((size_t)(2 * k + 2) * sizeof(double));
  DevBuf coef((size_t)2 * k * sizeof(double));
  const double direction = (t1 >= t0) ? 1.0 : -1.0;
  dt = std::fabs(dt) * direction;
  double t = t0;
  report(t);
  std::vector<double> hs(2 * (size_t)k + 2), yc(2 * (size_t)k);
  while ((direction > 0 && t < t1) || (direction < 0 && t > t1)) {
    const double step = (direction > 0) ? std::min(dt, t1 - t) : std::max(dt, t1 - t);
    launch_krylov_small(*tb, D, (const double *)out, basis.as<double>(), scal.as<double>(), k, s);
    MB_HIP(hipMemcpyAsync(hs.data(), scal.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MB_HIP(hipStreamSynchronize(s));
    const double bnorm = hs[2 * k];
    const int m = (int)hs[2 * k + 1];
    if (bnorm > kTol && m > 0) {  // a zero state stays as it is (evolve_state.h:131-134)
      LanczosResult dec;
      dec.steps = m;
      dec.b_norm = bnorm;
      dec.alphas.assign(hs.begin(), hs.begin() + m);
      dec.betas.assign(hs.begin() + k, hs.begin() + k + (m - 1));
      auto y = expm_first_column(dec, step);
      for (size_t j = 0; j < y.size(); ++j) {
        yc[2 * j] = y[j].real() * bnorm;
        yc[2 * j + 1] = y[j].imag() * bnorm;
      }
      MB_HIP(hipMemcpyAsync(coef.p, yc.data(), y.size() * 2 * sizeof(double), hipMemcpyHostToDevice, s));
      launch_basis_combine<double>(basis.as<double>(), m, D, coef.as<double>(), (double *)out, s);
    }
    t += step;
    report(t);
  }
  MB_HIP(hipStreamSynchronize(s));
  return true;
}

void time_evolve_steps(const HvOp &op, const void *psi, double t0, double t1, double dt,
                       int krylov_steps, void *out, hipStream_t s, bool fp32,
                       const StepCallback &callback) {
  const int64_t D = op.dim();
  if (out != psi) MB_HIP(hipMemcpyAsync(out, psi, (size_t)D * 16, hipMemcpyDeviceToDevice, s));
  auto report = [&](double t) {
    if (!callback) return;
    MB_HIP(hipStreamSynchronize(s));
    callback(t, out);
  };
  if (dt == 0.0 || (t1 - t0) == 0.0) {
    MB_HIP(hipStreamSynchronize(s));
    report(t0);
    return;
  }
  const std::function<void(double)> rep = report;
  if (!fp32 && time_evolve_steps_small(op, t0, t1, dt, krylov_steps, out, s, rep)) return;
  if (fp32 ? time_evolve_steps_graph<float>(op, t0, t1, dt, krylov_steps, out, s, rep)
           : time_evolve_steps_graph<double>(op, t0, t1, dt, krylov_steps, out, s, rep))
    return;
  const double direction = (t1 >= t0) ? 1.0 : -1.0;
  dt = std::fabs(dt) * direction;
  double t = t0;
  report(t);
  DevBuf cache;  // the stored Krylov basis, allocated once for all steps
  while ((direction > 0 && t < t1) || (direction < 0 && t > t1)) {
    const double step = (direction > 0) ? std::min(dt, t1 - t) : std::max(dt, t1 - t);
    if (fp32)
      time_evolve_t<float>(op, out, step, krylov_steps, out, s, &cache);
    else
      time_evolve_t<double>(op, out, step, krylov_steps, out, s, &cache);
    t += step;
    report(t);
  }
}

// ------------------------------------------------------------ KPM

void kpm_dos_moments(const Hubbard &h, const Sector &sec, double a, double b, int M, int nvec,
                     uint64_t seed, bool fp32, double *mu, hipStream_t s) {
  if (M < 2 || (M & 1)) throw std::invalid_argument("M must be even and >= 2");
  if (nvec < 1 || nvec > 64) throw std::invalid_argument("nvec must be in [1, 64]");
  if (!(a > 0.0)) throw std::invalid_argument("rescaling factor a must be positive");
  const int64_t D = h.dim();
  const int P = 2 * nvec;
  const size_t esz = fp32 ? 4 : 8;
  const size_t bytes = (size_t)D * P * esz;
  DevBuf T0(bytes), T1(bytes), T2(bytes);
  const HvTables t = tables_of(h);
  const int64_t nslots = hv_partials(t, P);
  const int width = 2 * nvec;
  const int steps = M / 2;
  DevBuf partial((size_t)std::max<int64_t>(2 * nslots, kRedBlocks) * sizeof(double));
  DevBuf raw((size_t)(steps + 1) * width * sizeof(double));
  MB_HIP(hipMemsetAsync(raw.p, 0, raw.bytes, s));
  const PermTables p = perm_of(sec);
  launch_fill_phase(p, seed, nvec, fp32, T0.p, s);
  double *rawp = raw.as<double>();
  // row 0: <T0|T0> per vector in its second half
  for (int r = 0; r < nvec; ++r)
    launch_sumsq((const char *)T0.p + (size_t)2 * r * D * esz, fp32, 2 * D, partial.as<double>(),
                 rawp + nvec + r, false, s);
  void *tn = T0.p, *tprev = T2.p, *tnext = T1.p;
  for (int n = 0; n < steps; ++n) {
    EpiArgs ea;
    ea.prev = tprev;
    ea.first = n == 0;
    ea.scale2 = ea.first ? 1.0 / a : 2.0 / a;
    ea.shift = b;
    ea.partial = partial.as<double>();
    int64_t rows;
    if (fp32)
      rows = launch_hv<float>(t, P, (const float *)tn, (float *)tnext, Epi::Cheb, ea, s);
    else
      rows = launch_hv<double>(t, P, (const double *)tn, (double *)tnext, Epi::Cheb, ea, s);
    launch_reduce_cheb(partial.as<double>(), rows, nvec, rawp + (size_t)(n + 1) * width, s);
    void *old_prev = tprev;
    tprev = tn;
    tn = tnext;
    tnext = old_prev;
  }
  kpm_accumulate(rawp, M, nvec, D, mu, s);
}

// Moments from the raw overlaps (row 0: <T0|T0> per vector in the second half
// of the row; row n + 1: <T_{n+1}|T_n> then <T_{n+1}|T_{n+1}> per vector).
void kpm_accumulate(const double *rawp, int M, int nvec, int64_t D, double *mu, hipStream_t s) {
  const int width = 2 * nvec, steps = M / 2;
  std::vector<double> hr((size_t)(steps + 1) * width);
  MB_HIP(hipMemcpyAsync(hr.data(), rawp, hr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  // per vector r: mu0 = <T0|T0>, mu1 = <T1|T0>; step n (T_{n+1} from T_n):
  // mu_{2n+1} = 2 <T_{n+1}|T_n> - mu1, mu_{2n+2} = 2 <T_{n+1}|T_{n+1}> - mu0
  // (step 0: mu1 itself and mu2 = 2 <T1|T1> - mu0)
  std::vector<double> acc(M + 1, 0.0);
  for (int r = 0; r < nvec; ++r) {
    const double mu0 = hr[nvec + r];
    const double mu1 = hr[(size_t)1 * width + r];
    acc[0] += mu0;
    for (int n = 0; n < steps; ++n) {
      const double cross = hr[(size_t)(n + 1) * width + r];
      const double self = hr[(size_t)(n + 1) * width + nvec + r];
      const int o = 2 * n + 1;
      acc[o] += (n == 0) ? cross : 2.0 * cross - mu1;
      if (o + 1 <= M) acc[o + 1] += 2.0 * self - mu0;
    }
  }
  const double norm = 1.0 / ((double)nvec * (double)D);
  for (int n = 0; n < M; ++n) mu[n] = acc[n] * norm;
}

// The same moments for any operator, FP64, one random-phase vector at a time:
// per step one plain H.v and one combine pass with the Chebyshev epilogue
// (T_{n+1} = scale2 (H T_n - b T_n) - T_{n-1} and its two overlaps), reduced on
// the device; the phases are keyed by reference index, as in the real-space
// path, so both give the same moments for the same operator.
void kpm_dos_moments_op(const HvOp &op, const int64_t *ref, double a, double b, int M, int nvec,
                        uint64_t seed, double *mu, hipStream_t s) {
  if (M < 2 || (M & 1)) throw std::invalid_argument("M must be even and >= 2");
  if (nvec < 1 || nvec > 64) throw std::invalid_argument("nvec must be in [1, 64]");
  if (!(a > 0.0)) throw std::invalid_argument("rescaling factor a must be positive");
  const int64_t D = op.dim();
  const size_t bytes = (size_t)D * 2 * sizeof(double);
  DevBuf T0(bytes), T1(bytes), T2(bytes), z(bytes);
  const int width = 2 * nvec, steps = M / 2;
  const int64_t slots = combine_slots(D);
  DevBuf partial((size_t)std::max<int64_t>(4 * slots, kRedBlocks) * sizeof(double));
  DevBuf raw((size_t)(steps + 1) * width * sizeof(double));
  MB_HIP(hipMemsetAsync(raw.p, 0, raw.bytes, s));
  double *rawp = raw.as<double>();
  for (int r = 0; r < nvec; ++r) {
    launch_fill_phase_ref(ref, D, seed, nvec, r, false, T0.p, s);
    launch_sumsq(T0.p, false, 2 * D, partial.as<double>(), rawp + nvec + r, false, s);
    double *tn = T0.as<double>(), *tprev = T2.as<double>(), *tnext = T1.as<double>();
    for (int n = 0; n < steps; ++n) {
      op.hv(tn, z.as<double>(), Epi::Plain, EpiArgs{}, s);
      EpiArgs ea;
      ea.prev = tprev;
      ea.first = n == 0;
      ea.scale2 = ea.first ? 1.0 / a : 2.0 / a;
      ea.shift = b;
      ea.partial = partial.as<double>();
      CombineCoef c;
      c.ar = 1.0;
      const int64_t per_plane = launch_combine<double>(D, z.as<double>(), nullptr, tn, tnext, c,
                                                       Epi::Cheb, ea, s);
      double *row = rawp + (size_t)(n + 1) * width;
      launch_reduce_pairs(partial.as<double>(), 2 * per_plane, row + r, row + nvec + r, s);
      double *old_prev = tprev;
      tprev = tn;
      tn = tnext;
      tnext = old_prev;
    }
  }
  kpm_accumulate(rawp, M, nvec, D, mu, s);
}

}  // namespace mb

namespace mb {

// ------------------------------------------------------------ power method

double power_method(const HvOp &op, const void *seed, int iterations, double sign, double shift,
                    hipStream_t s) {
  if (iterations < 0) throw std::invalid_argument("iterations must be non-negative");
  if (iterations == 0) return 0.0;
  const int64_t D = op.dim();
  DevBuf v((size_t)D * 16), w((size_t)D * 16);
  DevBuf red(sizeof(double) * (size_t)std::max<int64_t>(3 * kRedBlocks, 2 * kRedBlocks + 4));
  double *partial = red.as<double>();
  std::vector<double> host(3 * kRedBlocks);
  double nrm = 0.0;
  launch_norm((const double *)seed, 2 * D, partial, partial + 2 * kRedBlocks, s);
  MB_HIP(hipMemcpyAsync(&nrm, partial + 2 * kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  if (nrm == 0.0) return 0.0;
  launch_scale_complex<double>((const double *)seed, 1.0 / nrm, 0.0, v.as<double>(), D, s);
  double eigenvalue = 0.0;
  for (int i = 0; i < iterations; ++i) {
    op.hv(v.as<double>(), w.as<double>(), Epi::Plain, EpiArgs{}, s);
    launch_power_step(v.as<double>(), w.as<double>(), D, sign, shift, partial, s);
    MB_HIP(hipMemcpyAsync(host.data(), partial, sizeof(double) * 3 * kRedBlocks,
                          hipMemcpyDeviceToHost, s));
    MB_HIP(hipStreamSynchronize(s));
    double num = 0.0, den = 0.0, ww = 0.0;  // fixed order over the blocks
    for (int b = 0; b < kRedBlocks; ++b) {
      num += host[3 * b];
      den += host[3 * b + 1];
      ww += host[3 * b + 2];
    }
    if (den == 0.0) return 0.0;
    eigenvalue = num / den;
    const double n2 = std::sqrt(ww);
    if (n2 == 0.0) return 0.0;
    launch_scale_complex<double>(w.as<double>(), 1.0 / n2, 0.0, v.as<double>(), D, s);
  }
  MB_HIP(hipStreamSynchronize(s));
  return eigenvalue;
}

std::pair<double, double> estimate_bounds(const HvOp &op, const void *seed, int power_iterations,
                                          double spectral_padding, hipStream_t s) {
  const double dominant = std::fabs(power_method(op, seed, power_iterations, 1.0, 0.0, s));
  if (dominant == 0.0) return {0.0, 0.0};
  const double shift =
      dominant * (1.0 + spectral_padding) + std::numeric_limits<double>::epsilon();
  const double beta = power_method(op, seed, power_iterations, 1.0, shift, s) - shift;
  const double alpha = shift - power_method(op, seed, power_iterations, -1.0, shift, s);
  if (alpha <= beta) return {alpha, beta};
  return {beta, alpha};
}

}  // namespace mb

namespace mb {

// ------------------------------------------------------------ generic solve

// solve(op, b, k, solver) (lanczos.h:220-237): pass one, y = solver(alphas,
// betas), y *= |b|, x = V y (stored basis where it fits, else pass two)
void krylov_solve(const HvOp &op, const void *b, int k, const TridiagSolver &solver, void *x,
                  hipStream_t s) {
  const int64_t D = op.dim();
  DevBuf V;
  if (k > 0) {
    const size_t bytes = (size_t)(k + 1) * 2 * D * sizeof(double);
    if (keep_krylov_basis(bytes)) {
      try {
        V.alloc(bytes);
      } catch (const NoMemError &) {
        V.release();
      }
    }
  }
  LanczosWork work(D, k, op.partials(), sizeof(double), V.p == nullptr);
  LanczosResult dec = V.p ? pass_one_t<double>(op, b, k, s, &work, V.as<double>())
                          : lanczos_pass_one(op, b, k, s, &work);
  if (dec.steps == 0) {
    MB_HIP(hipMemsetAsync(x, 0, (size_t)D * 16, s));
    MB_HIP(hipStreamSynchronize(s));
    return;
  }
  std::vector<std::complex<double>> y = solver(dec.alphas, dec.betas);
  if ((int)y.size() != dec.steps)
    throw std::runtime_error("Dimension mismatch: y_k size must match decomposition steps");
  for (auto &v : y) v *= dec.b_norm;
  if (V.p)
    combine_basis<double>(V.as<double>(), y, D, x, s);
  else
    lanczos_pass_two(op, b, dec, y, x, s, &work);
  MB_HIP(hipStreamSynchronize(s));
}

// find_max_eigenpair (lanczos.h:246-281): the largest Ritz value and its vector
double find_max_eigenpair(const HvOp &op, const void *b, int k, void *x, hipStream_t s) {
  double value = 0.0;
  auto pick = [&](const std::vector<double> &al, const std::vector<double> &be) {
    const int m = (int)al.size();
    std::vector<double> wv(m), Z((size_t)m * m);
    tridiag_eig(m, al.data(), be.data(), wv.data(), Z.data());
    value = wv[m - 1];
    std::vector<std::complex<double>> y(m);
    for (int i = 0; i < m; ++i) y[i] = Z[i + (size_t)m * (m - 1)];
    return y;
  };
  if (x) {
    krylov_solve(op, b, k, pick, x, s);
    return value;
  }
  LanczosResult dec = lanczos_pass_one(op, b, k, s);
  if (dec.steps == 0) return 0.0;
  pick(dec.alphas, dec.betas);
  return value;
}

}  // namespace mb

namespace mb {

// ------------------------------------------------------------ KPM projector

// kpm::jackson_kernel (kpm.h:19-30): N + 1 damping factors
std::vector<double> jackson_kernel(int N) {
  if (N < 0) throw std::invalid_argument("N must be non-negative");
  constexpr double pi = 3.14159265358979323846;
  const double denom = (double)(N + 1);
  const double cot_val = 1.0 / std::tan(pi / denom);
  std::vector<double> g(N + 1);
  for (int n = 0; n <= N; ++n) {
    const double nd = (double)n, arg = pi * nd / denom;
    g[n] = ((denom - nd) * std::cos(arg) + std::sin(arg) * cot_val) / denom;
  }
  return g;
}

std::vector<double> projector_moments(int M, double epsilon) {
  if (M < 0) throw std::invalid_argument("expansion order must be non-negative");
  if (!(epsilon >= -1.0 && epsilon <= 1.0))
    throw std::invalid_argument("Fermi energy outside rescaled spectrum");
  constexpr double pi = 3.14159265358979323846;
  std::vector<double> mu(M + 1);
  const double ac = std::acos(epsilon);
  mu[0] = 1.0 - ac / pi;
  for (int m = 1; m <= M; ++m) mu[m] = -2.0 * std::sin(m * ac) / (m * pi);
  return mu;
}

std::vector<double> lorentz_kernel(int N, double lambda) {
  if (N < 1) throw std::invalid_argument("N must be positive");
  std::vector<double> g(N + 1);
  const double sl = std::sinh(lambda);
  for (int n = 0; n <= N; ++n) g[n] = std::sinh(lambda * (1.0 - (double)n / N)) / sl;
  return g;
}

void kpm_projector_apply(const HvOp &op, int M, double e_min, double e_max, double e_fermi,
                         double padding, const void *v, void *out, hipStream_t s) {
  if (!(e_max > e_min)) throw std::invalid_argument("E_max must exceed E_min");
  const double a = (e_max - e_min) / 2.0 * (1.0 + padding);
  const double b = (e_max + e_min) / 2.0;
  const std::vector<double> mu = projector_moments(M, (e_fermi - b) / a);
  const std::vector<double> g = jackson_kernel(M);
  const int64_t D = op.dim(), n = 2 * D;
  launch_scale_complex<double>((const double *)v, g[0] * mu[0], 0.0, (double *)out, D, s);
  if (M == 0) {
    MB_HIP(hipStreamSynchronize(s));
    return;
  }
  DevBuf A((size_t)D * 16), B((size_t)D * 16), Y((size_t)D * 16);
  double *tp = A.as<double>(), *tc = B.as<double>(), *y = Y.as<double>();
  MB_HIP(hipMemcpyAsync(tp, v, (size_t)D * 16, hipMemcpyDeviceToDevice, s));
  // T_1 = (H - b) v / a into B
  op.hv(tp, y, Epi::Plain, EpiArgs{}, s);
  launch_cheb_accum(y, tp, tc, (double *)out, n, 1.0 / a, b, true, g[1] * mu[1], s);
  for (int m = 2; m <= M; ++m) {  // T_{m} = 2 H~ T_{m-1} - T_{m-2}, written over T_{m-2}
    op.hv(tc, y, Epi::Plain, EpiArgs{}, s);
    launch_cheb_accum(y, tc, tp, (double *)out, n, 2.0 / a, b, false, g[m] * mu[m], s);
    std::swap(tp, tc);
  }
  MB_HIP(hipStreamSynchronize(s));
}

}  // namespace mb
