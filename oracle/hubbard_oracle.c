/* Oracle: CPU restatement of keyehzy/manybody's Hubbard H.v path.
 *
 * TEST INFRASTRUCTURE ONLY (see hubbard_oracle.h).  The product library never
 * links or calls anything in this file.
 *
 * Parity status of each function is documented in tests/test_oracle.py, which
 * pins it against golden vectors written by the compiled reference
 * (tests/golden/), closed forms and numpy/scipy.
 */
#include "hubbard_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* tolerances.h:7-10: 1000 * machine epsilon */
#define ORC_TOL (1000.0 * 2.220446049250313e-16)

/* ------------------------------------------------------------------ basis */

/* basis.cpp:13-26 */
uint64_t orc_choose(uint64_t n, uint64_t m) {
  if (m > n) return 0;
  if (m == 0 || m == n) return 1;
  if (m > n - m) m = n - m;
  uint64_t r = 1;
  for (uint64_t i = 0; i < m; ++i) {
    r *= (n - i);
    r /= (i + 1);
  }
  return r;
}

typedef struct {
  int L, total;
  uint64_t cur_up, cur_dn;
  uint64_t *up, *dn;
  int64_t count, cap;
  int overflow;
} dfs_ctx;

/* generate_restrict_combinations_with_spin (basis.cpp:267-302): children are
 * enumerated spin-major (all up candidates, then all down candidates) with
 * orbital >= the last orbital, accepting a candidate only when it is strictly
 * greater than the last mode in (orbital, spin) order.  A leaf is reached when
 * both budgets are exhausted; the leaf key is the sorted creator string, which
 * this restatement stores as two occupation masks. */
static void dfs(dfs_ctx *c, int depth, int first, int up_left, int dn_left, int last_spin,
                int last_orb) {
  if (up_left == 0 && dn_left == 0) {
    if (c->up) {
      if (c->count >= c->cap) {
        c->overflow = 1;
        return;
      }
      c->up[c->count] = c->cur_up;
      c->dn[c->count] = c->cur_dn;
    }
    c->count++;
    return;
  }
  if (depth >= c->total) return;
  for (int spin = 0; spin < 2; ++spin) {
    for (int i = first; i < c->L; ++i) {
      if ((spin == 0 && up_left == 0) || (spin == 1 && dn_left == 0)) continue;
      int ok = depth == 0 || last_orb < i || (last_orb == i && spin > last_spin);
      if (!ok) continue;
      uint64_t bit = (uint64_t)1 << i;
      if (spin == 0) {
        c->cur_up |= bit;
        dfs(c, depth + 1, i, up_left - 1, dn_left, spin, i);
        c->cur_up &= ~bit;
      } else {
        c->cur_dn |= bit;
        dfs(c, depth + 1, i, up_left, dn_left - 1, spin, i);
        c->cur_dn &= ~bit;
      }
    }
  }
}

int64_t orc_basis_enumerate(int n_sites, int n_up, int n_dn, uint64_t *up, uint64_t *dn,
                            int64_t cap) {
  if (n_sites < 0 || n_sites > 63 || n_up < 0 || n_dn < 0 || n_up > n_sites || n_dn > n_sites)
    return -1;
  dfs_ctx c;
  memset(&c, 0, sizeof c);
  c.L = n_sites;
  c.total = n_up + n_dn;
  c.up = up;
  c.dn = dn;
  c.cap = cap;
  dfs(&c, 0, 0, n_up, n_dn, 0, -1);
  return c.overflow ? -1 : c.count;
}

/* ------------------------------------------------------------------ model */

/* hubbard_model.h:16-20 (1D), 72-82 (2D), 144-159 (3D); index.h:82-103 */
int orc_hubbard_bonds(int ndim, const int *sizes, int *from, int *to, int cap) {
  int n = 0;
  if (ndim == 1) {
    int L = sizes[0];
    for (int i = 0; i < L; ++i) {
      if (n >= cap) return -1;
      from[n] = i;
      to[n] = (i + 1) % L;
      n++;
    }
  } else if (ndim == 2) {
    int Lx = sizes[0], Ly = sizes[1];
    for (int x = 0; x < Lx; ++x)
      for (int y = 0; y < Ly; ++y) {
        int s = x + Lx * y;
        if (n + 2 > cap) return -1;
        from[n] = s; to[n] = ((x + 1) % Lx) + Lx * y; n++;
        from[n] = s; to[n] = x + Lx * ((y + 1) % Ly); n++;
      }
  } else if (ndim == 3) {
    int Lx = sizes[0], Ly = sizes[1], Lz = sizes[2];
    for (int x = 0; x < Lx; ++x)
      for (int y = 0; y < Ly; ++y)
        for (int z = 0; z < Lz; ++z) {
          int s = x + Lx * y + Lx * Ly * z;
          if (n + 3 > cap) return -1;
          from[n] = s; to[n] = ((x + 1) % Lx) + Lx * y + Lx * Ly * z; n++;
          from[n] = s; to[n] = x + Lx * ((y + 1) % Ly) + Lx * Ly * z; n++;
          from[n] = s; to[n] = x + Lx * y + Lx * Ly * ((z + 1) % Lz); n++;
        }
  } else {
    return -1;
  }
  return n;
}

/* A one-body monomial coeff * c+_{a} c_{b} of one spin, after the merge that
 * ExpressionBase performs when identical monomials are added
 * (hopping_generic, expression_base.h:394-403, add_to :85-97). */
typedef struct {
  int a, b;
  double c;
} monomial;

static int merge_monomials(int n_bonds, const int *from, const int *to, double t, monomial *out) {
  int n = 0;
  const double coeff = -t;
  for (int k = 0; k < n_bonds; ++k) {
    int pair[2][2] = {{from[k], to[k]}, {to[k], from[k]}};
    for (int h = 0; h < 2; ++h) {
      int found = -1;
      for (int m = 0; m < n; ++m)
        if (out[m].a == pair[h][0] && out[m].b == pair[h][1]) found = m;
      if (found >= 0)
        out[found].c += coeff;
      else {
        out[n].a = pair[h][0];
        out[n].b = pair[h][1];
        out[n].c = coeff;
        n++;
      }
    }
  }
  /* is_zero (expression_base.h:79-83): |c| < 1000 eps drops the term */
  int w = 0;
  for (int m = 0; m < n; ++m)
    if (!(out[m].c * out[m].c < ORC_TOL * ORC_TOL)) out[w++] = out[m];
  return w;
}

/* The ordered creator string of a basis key: all up creators by ascending
 * orbital, then all down creators by ascending orbital (operator.h:69-79 for
 * creation operators).  Mode code = spin * 64 + orbital. */
static int key_ops(uint64_t up, uint64_t dn, int L, int *ops) {
  int n = 0;
  for (int i = 0; i < L; ++i)
    if (up >> i & 1) ops[n++] = i;
  for (int i = 0; i < L; ++i)
    if (dn >> i & 1) ops[n++] = 64 + i;
  return n;
}

/* Apply c+_{a} c_{b} (same spin) to the ordered creator string, one operator
 * at a time: c_b anticommutes past every creator in front of its partner,
 * c+_a is inserted where the canonical order puts it, each passed operator
 * flipping the sign (canonicalize.h:36-51, 62-78).  Returns 0 if the result
 * vanishes. */
static int apply_hop(int *ops, int n, int a, int b, int *sign) {
  int pos = -1;
  for (int p = 0; p < n; ++p)
    if (ops[p] == b) pos = p;
  if (pos < 0) return 0;
  int s = (pos & 1) ? -1 : 1;
  for (int p = pos; p + 1 < n; ++p) ops[p] = ops[p + 1];
  n--;
  for (int p = 0; p < n; ++p)
    if (ops[p] == a) return 0;
  int q = 0;
  while (q < n && ops[q] < a) q++;
  for (int p = n; p > q; --p) ops[p] = ops[p - 1];
  ops[q] = a;
  if (q & 1) s = -s;
  *sign = s;
  return 1;
}

static void ops_to_masks(const int *ops, int n, uint64_t *up, uint64_t *dn) {
  *up = 0;
  *dn = 0;
  for (int p = 0; p < n; ++p) {
    if (ops[p] >= 64)
      *dn |= (uint64_t)1 << (ops[p] - 64);
    else
      *up |= (uint64_t)1 << ops[p];
  }
}

/* --------------------------------------------------------- hash index */

typedef struct {
  uint64_t *keys;
  int64_t *vals;
  uint64_t mask;
} hindex;

static uint64_t hmix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

static int hindex_build(hindex *h, const uint64_t *up, const uint64_t *dn, int64_t D) {
  uint64_t size = 16;
  while (size < (uint64_t)D * 2) size <<= 1;
  h->keys = (uint64_t *)malloc(size * sizeof(uint64_t));
  h->vals = (int64_t *)malloc(size * sizeof(int64_t));
  if (!h->keys || !h->vals) return -1;
  h->mask = size - 1;
  for (uint64_t i = 0; i < size; ++i) h->keys[i] = ~(uint64_t)0;
  for (int64_t i = 0; i < D; ++i) {
    uint64_t k = up[i] | (dn[i] << 32);
    uint64_t s = hmix(k) & h->mask;
    while (h->keys[s] != ~(uint64_t)0) s = (s + 1) & h->mask;
    h->keys[s] = k;
    h->vals[s] = i;
  }
  return 0;
}

static int64_t hindex_find(const hindex *h, uint64_t up, uint64_t dn) {
  uint64_t k = up | (dn << 32);
  uint64_t s = hmix(k) & h->mask;
  while (h->keys[s] != ~(uint64_t)0) {
    if (h->keys[s] == k) return h->vals[s];
    s = (s + 1) & h->mask;
  }
  return -1;
}

static void hindex_free(hindex *h) {
  free(h->keys);
  free(h->vals);
}

/* ------------------------------------------------------------- CSC build */

typedef struct {
  int64_t row;
  double v;
} entry;

/* All nonzeros of column j (compute_matrix_elements body, matrix_elements.h:84-95):
 * contributions landing on the same key are summed in term order, as add_to
 * does, and vanishing sums are dropped. */
static int column_entries(int L, uint64_t up, uint64_t dn, const monomial *mono, int n_mono,
                          double U, const hindex *h, entry *out, int *missing) {
  int n = 0;
  int ops[128], work[128];
  int nops = key_ops(up, dn, L, ops);
  for (int spin = 0; spin < 2; ++spin)
    for (int m = 0; m < n_mono; ++m) {
      memcpy(work, ops, sizeof(int) * nops);
      int sign;
      if (!apply_hop(work, nops, spin * 64 + mono[m].a, spin * 64 + mono[m].b, &sign)) continue;
      uint64_t u2, d2;
      ops_to_masks(work, nops, &u2, &d2);
      int64_t row = hindex_find(h, u2, d2);
      if (row < 0) {
        *missing = 1;
        continue;
      }
      double v = sign * mono[m].c;
      int f = -1;
      for (int e = 0; e < n; ++e)
        if (out[e].row == row) f = e;
      if (f >= 0)
        out[f].v += v;
      else {
        out[n].row = row;
        out[n].v = v;
        n++;
      }
    }
  /* n_up n_dn on every site: U per doubly occupied orbital */
  if (!(U * U < ORC_TOL * ORC_TOL)) {
    int64_t self = hindex_find(h, up, dn);
    for (int i = 0; i < L; ++i)
      if ((up >> i & 1) && (dn >> i & 1)) {
        int f = -1;
        for (int e = 0; e < n; ++e)
          if (out[e].row == self) f = e;
        if (f >= 0)
          out[f].v += U;
        else {
          out[n].row = self;
          out[n].v = U;
          n++;
        }
      }
  }
  int w = 0;
  for (int e = 0; e < n; ++e)
    if (!(out[e].v * out[e].v < ORC_TOL * ORC_TOL)) out[w++] = out[e];
  /* sort by row (CSC keeps row indices ascending within a column) */
  for (int i = 1; i < w; ++i) {
    entry x = out[i];
    int j = i - 1;
    while (j >= 0 && out[j].row > x.row) {
      out[j + 1] = out[j];
      j--;
    }
    out[j + 1] = x;
  }
  return w;
}

int64_t orc_build_csc(int n_sites, const uint64_t *up, const uint64_t *dn, int64_t D,
                      int n_bonds, const int *from, const int *to, double t, double U,
                      int64_t *col_ptr, int64_t *row_idx, double *val, int64_t cap, int threads) {
  if (n_sites > 31) return -1;
  monomial *mono = (monomial *)malloc(sizeof(monomial) * (size_t)(2 * n_bonds + 1));
  int n_mono = merge_monomials(n_bonds, from, to, t, mono);
  hindex h;
  if (hindex_build(&h, up, dn, D)) return -1;
  int missing = 0;
  const int emax = 2 * n_mono + 2;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  /* pass 1: count per column */
#pragma omp parallel
  {
    entry *buf = (entry *)malloc(sizeof(entry) * (size_t)emax);
#pragma omp for schedule(dynamic, 256)
    for (int64_t j = 0; j < D; ++j) {
      int miss = 0;
      int n = column_entries(n_sites, up[j], dn[j], mono, n_mono, U, &h, buf, &miss);
      col_ptr[j + 1] = n;
      if (miss) missing = 1;
    }
    free(buf);
  }
  col_ptr[0] = 0;
  for (int64_t j = 0; j < D; ++j) col_ptr[j + 1] += col_ptr[j];
  int64_t nnz = col_ptr[D];
  if (row_idx && !missing) {
    if (nnz > cap) {
      hindex_free(&h);
      free(mono);
      return -1;
    }
#pragma omp parallel
    {
      entry *buf = (entry *)malloc(sizeof(entry) * (size_t)emax);
#pragma omp for schedule(dynamic, 256)
      for (int64_t j = 0; j < D; ++j) {
        int miss = 0;
        int n = column_entries(n_sites, up[j], dn[j], mono, n_mono, U, &h, buf, &miss);
        for (int e = 0; e < n; ++e) {
          row_idx[col_ptr[j] + e] = buf[e].row;
          val[2 * (col_ptr[j] + e)] = buf[e].v;
          val[2 * (col_ptr[j] + e) + 1] = 0.0;
        }
      }
      free(buf);
    }
  }
  hindex_free(&h);
  free(mono);
  return missing ? -1 : nnz;
}

static int progress_every = 0;

void orc_set_progress(int every) { progress_every = every; }

void orc_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

void orc_spmv_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx, const double *val,
                  const double *x, double *y) {
  memset(y, 0, sizeof(double) * 2 * (size_t)D);
  for (int64_t j = 0; j < D; ++j) {
    const double xr = x[2 * j], xi = x[2 * j + 1];
    for (int64_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
      const double ar = val[2 * p], ai = val[2 * p + 1];
      const int64_t i = row_idx[p];
      y[2 * i] += ar * xr - ai * xi;
      y[2 * i + 1] += ar * xi + ai * xr;
    }
  }
}

/* ------------------------------------------------- product-order oracle */

static int64_t bsearch_u64(const uint64_t *a, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

int64_t orc_species_table(int n_sites, int n, int n_bonds, const int *from, const int *to,
                          double t, uint64_t *strings, int64_t *row_ptr, int32_t *cols,
                          double *vals, double *diag_onsite, int64_t cap) {
  /* all n-subsets of n_sites in ascending numeric order */
  int64_t cnt = 0;
  for (uint64_t s = 0; s < ((uint64_t)1 << n_sites); ++s)
    if (__builtin_popcountll(s) == n) strings[cnt++] = s;
  monomial *mono = (monomial *)malloc(sizeof(monomial) * (size_t)(2 * n_bonds + 1));
  int n_mono = merge_monomials(n_bonds, from, to, t, mono);
  int64_t nnz = 0;
  int ops[64];
  row_ptr[0] = 0;
  for (int64_t r = 0; r < cnt; ++r) {
    double onsite = 0.0;
    for (int m = 0; m < n_mono; ++m) {
      int nops = 0;
      for (int i = 0; i < n_sites; ++i)
        if (strings[r] >> i & 1) ops[nops++] = i;
      int sign;
      if (!apply_hop(ops, nops, mono[m].a, mono[m].b, &sign)) continue;
      uint64_t tgt = 0;
      for (int p = 0; p < nops; ++p) tgt |= (uint64_t)1 << ops[p];
      if (mono[m].a == mono[m].b) {
        onsite += sign * mono[m].c;
        continue;
      }
      if (nnz >= cap) {
        free(mono);
        return -1;
      }
      cols[nnz] = (int32_t)bsearch_u64(strings, cnt, tgt);
      vals[nnz] = sign * mono[m].c;
      nnz++;
    }
    diag_onsite[r] = onsite;
    row_ptr[r + 1] = nnz;
  }
  free(mono);
  return nnz;
}

void orc_diag_table(double U, int n_max, double *diagU) {
  for (int k = 0; k <= n_max; ++k) {
    double s = 0.0;
    if (!(U * U < ORC_TOL * ORC_TOL))
      for (int i = 0; i < k; ++i) s += U;
    diagU[k] = s;
  }
}

void orc_hv_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up, const uint64_t *s_dn,
                    const int64_t *up_ptr, const int32_t *up_cols, const double *up_vals,
                    const double *up_onsite, const int64_t *dn_ptr, const int32_t *dn_cols,
                    const double *dn_vals, const double *dn_onsite, const double *diagU,
                    const double *x, double *y, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < Dup; ++u) {
    for (int64_t d = 0; d < Ddn; ++d) {
      const int64_t e = u * Ddn + d;
      const double dg =
          diagU[__builtin_popcountll(s_up[u] & s_dn[d])] + up_onsite[u] + dn_onsite[d];
      double yr = dg * x[2 * e], yi = dg * x[2 * e + 1];
      for (int64_t p = up_ptr[u]; p < up_ptr[u + 1]; ++p) {
        const int64_t f = (int64_t)up_cols[p] * Ddn + d;
        yr += up_vals[p] * x[2 * f];
        yi += up_vals[p] * x[2 * f + 1];
      }
      for (int64_t p = dn_ptr[d]; p < dn_ptr[d + 1]; ++p) {
        const int64_t f = u * Ddn + dn_cols[p];
        yr += dn_vals[p] * x[2 * f];
        yi += dn_vals[p] * x[2 * f + 1];
      }
      y[2 * e] = yr;
      y[2 * e + 1] = yi;
    }
  }
}

/* ---------------------------------------------------------------- Lanczos */

typedef void (*apply_fn)(void *ctx, const double *x, double *y);

typedef struct {
  int64_t D;
  const int64_t *cp, *ri;
  const double *val;
  int negate;
} csc_op;

static void csc_apply(void *ctx, const double *x, double *y) {
  csc_op *o = (csc_op *)ctx;
  orc_spmv_csc(o->D, o->cp, o->ri, o->val, x, y);
  if (o->negate)
    for (int64_t i = 0; i < 2 * o->D; ++i) y[i] = -y[i];
}

typedef struct {
  int64_t Dup, Ddn;
  const uint64_t *su, *sd;
  const int64_t *up, *dp;
  const int32_t *uc, *dc;
  const double *uv, *dv, *uo, *dO, *diagU;
  int threads;
} prod_op;

static void prod_apply(void *ctx, const double *x, double *y) {
  prod_op *o = (prod_op *)ctx;
  orc_hv_product(o->Dup, o->Ddn, o->su, o->sd, o->up, o->uc, o->uv, o->uo, o->dp, o->dc, o->dv,
                 o->dO, o->diagU, x, y, o->threads);
}

/* sum of a[i] * b[i]: 256 fixed chunks summed in parallel, then added in chunk
 * order, so the result does not depend on the thread count or schedule */
static double dot_fixed(int64_t n, const double *a, const double *b) {
  enum { NC = 256 };
  double part[NC];
  const int64_t len = (n + NC - 1) / NC;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < NC; ++c) {
    const int64_t lo = c * len, hi = lo + len < n ? lo + len : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (int c = 0; c < NC; ++c) s += part[c];
  return s;
}

static double cnorm(int64_t D, const double *x) { return sqrt(dot_fixed(2 * D, x, x)); }

/* lanczos_pass_one (lanczos.h:130-175) with lanczos_recurrence_step (:15-40)
 * and LanczosIteration::next (:104-127). */
static int lanczos_generic(int64_t D, apply_fn f, void *ctx, const double *b, int k,
                           double *alphas, double *betas, double *b_norm) {
  *b_norm = cnorm(D, b);
  if (*b_norm <= ORC_TOL) return -1;
  if (k == 0) return 0;
  double *vp = (double *)calloc(2 * (size_t)D, sizeof(double));
  double *v = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  /* the vector passes use the threads of the last omp_set_num_threads (the
   * product-order H.v sets them); sums use fixed chunks (dot_fixed) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 2 * D; ++i) v[i] = b[i] / *b_norm;
  double beta_prev = 0.0;
  int steps = 0;
  for (int i = 0; i < k; ++i) {
    f(ctx, v, w);
    if (beta_prev != 0.0) {
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < 2 * D; ++e) w[e] -= beta_prev * vp[e];
    }
    const double alpha = dot_fixed(2 * D, v, w); /* Re(conj(v) w) */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < 2 * D; ++e) w[e] -= alpha * v[e];
    const double beta = cnorm(D, w);
    alphas[steps++] = alpha;
    if (progress_every > 0 && steps % progress_every == 0) {
      fprintf(stderr, "oracle Lanczos: step %d of %d\n", steps, k);
      fflush(stderr);
    }
    if (beta <= ORC_TOL) break;
    if (i + 1 < k) betas[i] = beta;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < 2 * D; ++e) w[e] /= beta;
    double *tmp = vp;
    vp = v;
    v = w;
    w = tmp;
    beta_prev = beta;
  }
  free(vp);
  free(v);
  free(w);
  return steps;
}

int orc_lanczos_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                    const double *val, int negate, const double *b, int k, double *alphas,
                    double *betas, double *b_norm) {
  csc_op o = {D, col_ptr, row_idx, val, negate};
  return lanczos_generic(D, csc_apply, &o, b, k, alphas, betas, b_norm);
}

int orc_lanczos_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up, const uint64_t *s_dn,
                        const int64_t *up_ptr, const int32_t *up_cols, const double *up_vals,
                        const double *up_onsite, const int64_t *dn_ptr, const int32_t *dn_cols,
                        const double *dn_vals, const double *dn_onsite, const double *diagU,
                        const double *b, int k, double *alphas, double *betas, double *b_norm,
                        int threads) {
  prod_op o = {Dup, Ddn, s_up, s_dn, up_ptr, dn_ptr, up_cols, dn_cols, up_vals, dn_vals,
               up_onsite, dn_onsite, diagU, threads};
  return lanczos_generic(Dup * Ddn, prod_apply, &o, b, k, alphas, betas, b_norm);
}

/* ------------------------------------------------------ small eigenproblems */

static int sturm_count(int m, const double *a, const double *b, double x) {
  int c = 0;
  double q = a[0] - x;
  if (q < 0) c++;
  for (int i = 1; i < m; ++i) {
    if (q == 0.0) q = 1e-300;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) c++;
  }
  return c;
}

void orc_tridiag_eigvals(int m, const double *a, const double *b, double *w) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    double r = 0.0;
    if (i > 0) r += fabs(b[i - 1]);
    if (i + 1 < m) r += fabs(b[i]);
    if (a[i] - r < lo) lo = a[i] - r;
    if (a[i] + r > hi) hi = a[i] + r;
  }
  double pad = 1e-12 * (fabs(lo) + fabs(hi) + 1.0);
  lo -= pad;
  hi += pad;
  for (int j = 0; j < m; ++j) {
    /* j-th smallest: largest x with count(x) <= j */
    double l = lo, h = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (l + h);
      if (mid == l || mid == h) break;
      if (sturm_count(m, a, b, mid) <= j)
        l = mid;
      else
        h = mid;
    }
    w[j] = 0.5 * (l + h);
  }
}

void orc_jacobi_eig(int m, const double *A, double *w, double *V) {
  double *S = (double *)malloc(sizeof(double) * (size_t)m * m);
  memcpy(S, A, sizeof(double) * (size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) V[i * m + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += S[p * m + q] * S[p * m + q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        double apq = S[p * m + q];
        if (apq == 0.0) continue;
        double theta = (S[q * m + q] - S[p * m + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          double skp = S[k * m + p], skq = S[k * m + q];
          S[k * m + p] = c * skp - s * skq;
          S[k * m + q] = s * skp + c * skq;
        }
        for (int k = 0; k < m; ++k) {
          double spk = S[p * m + k], sqk = S[q * m + k];
          S[p * m + k] = c * spk - s * sqk;
          S[q * m + k] = s * spk + c * sqk;
        }
        for (int k = 0; k < m; ++k) {
          double vkp = V[k * m + p], vkq = V[k * m + q];
          V[k * m + p] = c * vkp - s * vkq;
          V[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < m; ++i) w[i] = S[i * m + i];
  /* sort ascending with vectors */
  for (int i = 0; i < m; ++i) {
    int mn = i;
    for (int j = i + 1; j < m; ++j)
      if (w[j] < w[mn]) mn = j;
    if (mn != i) {
      double tw = w[i];
      w[i] = w[mn];
      w[mn] = tw;
      for (int k = 0; k < m; ++k) {
        double tv = V[k * m + i];
        V[k * m + i] = V[k * m + mn];
        V[k * m + mn] = tv;
      }
    }
  }
  free(S);
}

/* lanczos_pass_two (lanczos.h:177-218) with complex coefficients y (interleaved). */
static void pass_two_complex(int64_t D, apply_fn f, void *ctx, const double *b, int steps,
                             const double *alphas, const double *betas, double b_norm,
                             const double *y, double *x) {
  memset(x, 0, sizeof(double) * 2 * (size_t)D);
  if (steps == 0) return;
  double *vp = (double *)calloc(2 * (size_t)D, sizeof(double));
  double *v = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  for (int64_t i = 0; i < 2 * D; ++i) v[i] = b[i] / b_norm;
  for (int64_t e = 0; e < D; ++e) {
    x[2 * e] = v[2 * e] * y[0] - v[2 * e + 1] * y[1];
    x[2 * e + 1] = v[2 * e] * y[1] + v[2 * e + 1] * y[0];
  }
  for (int j = 0; j + 1 < steps; ++j) {
    const double aj = alphas[j], bj = betas[j], bp = (j == 0) ? 0.0 : betas[j - 1];
    f(ctx, v, w);
    if (bp != 0.0)
      for (int64_t e = 0; e < 2 * D; ++e) w[e] -= bp * vp[e];
    for (int64_t e = 0; e < 2 * D; ++e) w[e] -= aj * v[e];
    for (int64_t e = 0; e < 2 * D; ++e) w[e] /= bj;
    const double yr = y[2 * (j + 1)], yi = y[2 * (j + 1) + 1];
    for (int64_t e = 0; e < D; ++e) {
      x[2 * e] += w[2 * e] * yr - w[2 * e + 1] * yi;
      x[2 * e + 1] += w[2 * e] * yi + w[2 * e + 1] * yr;
    }
    double *tmp = vp;
    vp = v;
    v = w;
    w = tmp;
  }
  free(vp);
  free(v);
  free(w);
}

void orc_lanczos_pass_two_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                              const double *val, int negate, const double *b, int steps,
                              const double *alphas, const double *betas, double b_norm,
                              const double *y, double *x) {
  csc_op o = {D, col_ptr, row_idx, val, negate};
  double *yc = (double *)calloc(2 * (size_t)(steps > 0 ? steps : 1), sizeof(double));
  for (int j = 0; j < steps; ++j) yc[2 * j] = y[j];
  pass_two_complex(D, csc_apply, &o, b, steps, alphas, betas, b_norm, yc, x);
  free(yc);
}

/* evolve_state (evolve_state.h:120-151) with expm_tridiagonal_first_column_unitary
 * (:22-61) and solve (lanczos.h:220-237), over any operator. */
static void time_evolve_generic(int64_t D, apply_fn f, void *ctx, const double *psi, double t,
                                int k, double *out) {
  memcpy(out, psi, sizeof(double) * 2 * (size_t)D);
  if (D == 0) return;
  const double nrm = cnorm(D, psi);
  if (nrm <= ORC_TOL) return;
  if (k == 0 || k > D) k = (int)D;
  double *al = (double *)malloc(sizeof(double) * (size_t)k);
  double *be = (double *)malloc(sizeof(double) * (size_t)k);
  double bn;
  int m = lanczos_generic(D, f, ctx, psi, k, al, be, &bn);
  if (m <= 0) {
    memset(out, 0, sizeof(double) * 2 * (size_t)D);
    free(al);
    free(be);
    return;
  }
  double *T = (double *)calloc((size_t)m * m, sizeof(double));
  for (int i = 0; i < m; ++i) T[i * m + i] = al[i];
  for (int i = 0; i + 1 < m; ++i) T[i * m + i + 1] = T[(i + 1) * m + i] = be[i];
  double *w = (double *)malloc(sizeof(double) * (size_t)m);
  double *Q = (double *)malloc(sizeof(double) * (size_t)m * m);
  orc_jacobi_eig(m, T, w, Q);
  double *y = (double *)calloc(2 * (size_t)m, sizeof(double));
  for (int i = 0; i < m; ++i) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < m; ++j) {
      const double wt = Q[0 * m + j];
      re += Q[i * m + j] * cos(-w[j] * t) * wt;
      im += Q[i * m + j] * sin(-w[j] * t) * wt;
    }
    y[2 * i] = re * bn;
    y[2 * i + 1] = im * bn;
  }
  pass_two_complex(D, f, ctx, psi, m, al, be, bn, y, out);
  free(T);
  free(w);
  free(Q);
  free(y);
  free(al);
  free(be);
}

void orc_time_evolve_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                         const double *val, const double *psi, double t, int k, double *out) {
  csc_op o = {D, col_ptr, row_idx, val, 0};
  time_evolve_generic(D, csc_apply, &o, psi, t, k, out);
}

void orc_time_evolve_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up,
                             const uint64_t *s_dn, const int64_t *up_ptr, const int32_t *up_cols,
                             const double *up_vals, const double *up_onsite,
                             const int64_t *dn_ptr, const int32_t *dn_cols, const double *dn_vals,
                             const double *dn_onsite, const double *diagU, const double *psi,
                             double t, int k, double *out, int threads) {
  prod_op o = {Dup, Ddn, s_up, s_dn, up_ptr, dn_ptr, up_cols, dn_cols, up_vals, dn_vals,
               up_onsite, dn_onsite, diagU, threads};
  time_evolve_generic(Dup * Ddn, prod_apply, &o, psi, t, k, out);
}

/* ---------------------------------------------------------------- KPM */

/* Density-of-states moments mu_n = (1 / (R D)) sum_r Re <r|T_n(Hs)|r>, Hs =
 * (H - b) / a, by the Chebyshev recurrence of kpm.h:212-229 (T_{n+1} =
 * 2 Hs T_n - T_{n-1}) and a direct overlap per order (no doubling identities). */
static void kpm_generic(int64_t D, apply_fn f, void *ctx, double a, double b, int M, int R,
                        const double *vecs, double *mu) {
  double *t0 = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  double *t1 = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  double *t2 = (double *)malloc(sizeof(double) * 2 * (size_t)D);
  for (int n = 0; n < M; ++n) mu[n] = 0.0;
  for (int r = 0; r < R; ++r) {
    const double *v = vecs + 2 * (size_t)D * r;
    memcpy(t0, v, sizeof(double) * 2 * (size_t)D);
    for (int n = 0; n < M; ++n) {
      if (n == 0) {
        memcpy(t1, t0, sizeof(double) * 2 * (size_t)D);
      } else if (n == 1) {
        f(ctx, t0, t1);
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < 2 * D; ++e) t1[e] = (t1[e] - b * t0[e]) / a;
        /* keep T_0 in t2, T_1 in t1 */
        memcpy(t2, t0, sizeof(double) * 2 * (size_t)D);
      } else {
        /* T_{n} = 2 Hs T_{n-1} - T_{n-2}; t1 = T_{n-1}, t2 = T_{n-2} */
        f(ctx, t1, t0);
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < 2 * D; ++e) t0[e] = 2.0 * (t0[e] - b * t1[e]) / a - t2[e];
        double *tmp = t2;
        t2 = t1;
        t1 = t0;
        t0 = tmp;
      }
      /* <r | T_n r>, real part */
      mu[n] += dot_fixed(2 * D, v, t1);
    }
  }
  for (int n = 0; n < M; ++n) mu[n] /= ((double)R * (double)D);
  free(t0);
  free(t1);
  free(t2);
}

void orc_kpm_moments_csc(int64_t D, const int64_t *col_ptr, const int64_t *row_idx,
                         const double *val, double a, double b, int M, int R,
                         const double *vecs, double *mu) {
  csc_op o = {D, col_ptr, row_idx, val, 0};
  kpm_generic(D, csc_apply, &o, a, b, M, R, vecs, mu);
}

void orc_kpm_moments_product(int64_t Dup, int64_t Ddn, const uint64_t *s_up,
                             const uint64_t *s_dn, const int64_t *up_ptr, const int32_t *up_cols,
                             const double *up_vals, const double *up_onsite,
                             const int64_t *dn_ptr, const int32_t *dn_cols, const double *dn_vals,
                             const double *dn_onsite, const double *diagU, double a, double b,
                             int M, int R, const double *vecs, double *mu, int threads) {
  prod_op o = {Dup, Ddn, s_up, s_dn, up_ptr, dn_ptr, up_cols, dn_cols, up_vals, dn_vals,
               up_onsite, dn_onsite, diagU, threads};
  kpm_generic(Dup * Ddn, prod_apply, &o, a, b, M, R, vecs, mu);
}

/* kpm.h:19-30 */
void orc_jackson_kernel(int N, double *g) {
  const double pi = 3.14159265358979323846;
  const double denom = (double)(N + 1);
  const double cot_val = 1.0 / tan(pi / denom);
  for (int n = 0; n <= N; ++n) {
    const double nd = (double)n;
    const double arg = pi * nd / denom;
    g[n] = ((denom - nd) * cos(arg) + sin(arg) * cot_val) / denom;
  }
}
