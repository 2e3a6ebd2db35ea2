// extern "C" boundary of libmanybody_hip.so (declarations and reference
// citations: include/manybody_hip.h).  Every entry point converts C++
// exceptions into status codes and a thread-local message.
#include <cmath>
#include <cstring>
#include <memory>

#include "../../include/manybody_hip.h"
#include "batch.hpp"
#include "serve.hpp"
#include "observables.hpp"
#include "krylov.hpp"
#include "mb_internal.hpp"
#include "momentum.hpp"
#include "conductivity.hpp"
#include "expression.hpp"
#include "fixed_n.hpp"
#include "rectangular.hpp"
#include "composite.hpp"

#include <cstdio>
#include <map>
#include <fstream>

struct mb_basis {
  mb::Sector sec;
  std::unique_ptr<mb::MomentumSector> ms;  // set for a fixed-momentum basis
  std::unique_ptr<mb::FixedNBasis> fn;     // set for a fixed-N (any N_up) basis
  int orbitals() const { return ms ? ms->L : fn ? fn->L : sec.L; }
  int64_t size() const { return ms ? ms->D : fn ? fn->D : sec.dim(); }
};
struct mb_hubbard {
  mb::Hubbard h;  // real-space model (unused for a momentum-sector operator)
  const mb_basis *basis;
  std::unique_ptr<mb::RealHvOp> real_op;
  std::unique_ptr<mb::MomentumHubbard> mom;
  std::unique_ptr<mb::ExpressionOperator> expr;  // a compiled FermionExpression
  std::unique_ptr<mb::CompositeOp> comp;          // operator algebra over other handles
  const mb::HvOp &op() const {
    if (comp) return *comp;
    if (mom) return *mom;
    if (expr) return *expr;
    return *real_op;
  }
  // built on the basis handle alone (expression or composite operators)
  bool on_basis() const { return expr || comp; }
  // the fixed-momentum sector the operator acts on, or null (real space)
  const mb::MomentumSector *msec() const {
    if (mom) return &mom->sector();
    if (on_basis() && basis->ms) return basis->ms.get();
    return nullptr;
  }
  // the fixed-N basis the operator acts on, or null
  const mb::FixedNBasis *fsec() const {
    return (on_basis() && basis->fn) ? basis->fn.get() : nullptr;
  }
  // the device position -> reference index map of a grouped basis (fixed K or
  // fixed N), uploaded on first use; null for a real-space sector
  const int64_t *ref_map() const {
    if (auto *m = msec()) {
      auto &mm = const_cast<mb::MomentumSector &>(*m);
      mm.ensure_device();
      return mm.d_ref_of_pos.as<int64_t>();
    }
    if (auto *f = fsec()) {
      auto &ff = const_cast<mb::FixedNBasis &>(*f);
      ff.ensure_device();
      return ff.d_ref_of_pos.as<int64_t>();
    }
    return nullptr;
  }
  // the real-space sector the operator acts on, or null (momentum, fixed N)
  const mb::Sector *rsec() const {
    if (mom || (on_basis() && (basis->ms || basis->fn))) return nullptr;
    if (on_basis()) {
      auto &sec = const_cast<mb::Sector &>(basis->sec);
      sec.ensure_device();
      return &sec;
    }
    return h.sec;
  }
};
struct mb_sparse {
  uint64_t n_rows = 0, n_cols = 0;
  std::vector<uint64_t> rows, cols;
  std::vector<std::complex<double>> vals;
};
struct mb_batch {
  std::unique_ptr<mb::Batch> b;
};
struct mb_current {
  std::unique_ptr<mb::CurrentOperator> j;
  const mb_basis *source, *target;
};
struct mb_server {
  std::unique_ptr<mb::Server> s;
};

namespace {

thread_local std::string g_err;

template <typename F>
int guarded(F &&f) {
  try {
    f();
    g_err.clear();
    return MB_OK;
  } catch (const mb::RangeError &e) {
    g_err = e.what();
    return MB_ERR_RANGE;
  } catch (const std::out_of_range &e) {
    g_err = e.what();
    return MB_ERR_RANGE;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return MB_ERR_INVALID;
  } catch (const mb::NoMemError &e) {
    g_err = e.what();
    return MB_ERR_NOMEM;
  } catch (const mb::DeviceError &e) {
    g_err = e.what();
    return MB_ERR_DEVICE;
  } catch (const std::bad_alloc &e) {
    g_err = "host allocation failed";
    return MB_ERR_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return MB_ERR_RUNTIME;
  } catch (...) {
    g_err = "unknown error";
    return MB_ERR_RUNTIME;
  }
}

void need(bool cond, const char *msg) {
  if (!cond) throw std::invalid_argument(msg);
}

hipStream_t S(mb_stream s) { return reinterpret_cast<hipStream_t>(s); }

// the real-space model; entry points without a momentum-sector form reject those
const mb::Hubbard &H(const mb_hubbard *h) {
  need(h != nullptr, "null model handle");
  need(!h->mom, "not available for momentum-sector operators");
  need(!h->expr, "not available for expression operators");
  need(!h->comp, "not available for composite operators");
  return h->h;
}

// device position -> reference index of any operator handle; computed into
// `own` for a real-space sector (free it only after the work using it)
const int64_t *ref_positions(const mb_hubbard *h, mb::DevBuf &own, hipStream_t s) {
  if (const int64_t *m = h->ref_map()) return m;
  const mb::Sector *sec = h->rsec();
  own.alloc((size_t)std::max<int64_t>(sec->dim(), 1) * sizeof(int64_t));
  mb::launch_ref_of_pos(mb::perm_of(*sec), own.as<int64_t>(), s);
  return own.as<int64_t>();
}

// whether two bases have the same device layout (same kind and parameters)
bool same_layout(const mb_basis *a, const mb_basis *b) {
  if (a == b) return true;
  if ((a->ms != nullptr) != (b->ms != nullptr) || (a->fn != nullptr) != (b->fn != nullptr))
    return false;
  if (a->ms) {
    const auto &p = *a->ms, &q = *b->ms;
    return p.L == q.L && p.n_up == q.n_up && p.n_dn == q.n_dn && p.K == q.K &&
           p.ndim == q.ndim && p.dims[0] == q.dims[0] && p.dims[1] == q.dims[1] &&
           p.dims[2] == q.dims[2];
  }
  if (a->fn) return a->fn->L == b->fn->L && a->fn->N == b->fn->N;
  return a->sec.L == b->sec.L && a->sec.n_up == b->sec.n_up && a->sec.n_dn == b->sec.n_dn;
}

int make_composite(mb::CompositeOp::Kind k, const mb_hubbard *a, const mb_hubbard *b,
                   double re, double im, mb_hubbard **out) {
  return guarded([&] {
    need(a && out, "null argument");
    *out = nullptr;
    const bool binary = k == mb::CompositeOp::Kind::Sum ||
                        k == mb::CompositeOp::Kind::Difference ||
                        k == mb::CompositeOp::Kind::Composed;
    if (binary) {
      need(b != nullptr, "null argument");
      need(a->op().dim() == b->op().dim(), "operand dimensions differ");
      need(same_layout(a->basis, b->basis), "operands act on different bases");
    }
    auto h = std::make_unique<mb_hubbard>();
    h->basis = a->basis;
    h->comp = std::make_unique<mb::CompositeOp>(k, &a->op(), binary ? &b->op() : nullptr,
                                                std::complex<double>(re, im), a->op().dim());
    *out = h.release();
  });
}

// the operator the Krylov layer runs on (either kind)
const mb::HvOp &OP(const mb_hubbard *h) {
  need(h != nullptr, "null model handle");
  return h->op();
}

// the real-space sector of a basis; momentum bases are rejected
const mb::Sector &SEC(const mb_basis *b) {
  need(b != nullptr, "null basis handle");
  need(!b->ms, "not available for momentum-sector bases");
  need(!b->fn, "not available for fixed-N bases");
  return b->sec;
}

}  // namespace

extern "C" {

const char *mb_last_error(void) { return g_err.c_str(); }
const char *mb_version(void) { return "manybody_amd 0.1 (gfx950)"; }

int mb_basis_create(int orbitals, int particles, int spin_projection, mb_basis **out) {
  return guarded([&] {
    need(out != nullptr, "null output");
    *out = nullptr;
    need(orbitals >= 1 && orbitals <= mb::kMaxSites, "orbitals must be in [1, 32]");
    need(particles >= 0 && particles <= 2 * orbitals, "particles must be in [0, 2*orbitals]");
    need(std::abs(spin_projection) <= particles, "|spin_projection| must not exceed particles");
    need((particles + spin_projection) % 2 == 0, "particles + spin_projection must be even");
    const int up = (particles + spin_projection) / 2, dn = particles - up;
    need(up <= orbitals && dn <= orbitals, "spin sector exceeds the orbital count");
    auto b = std::make_unique<mb_basis>();
    mb::build_sector(b->sec, orbitals, up, dn);
    *out = b.release();
  });
}

void mb_basis_destroy(mb_basis *b) { delete b; }

int64_t mb_basis_size(const mb_basis *b) { return b ? b->size() : -1; }

int mb_basis_state(const mb_basis *b, int64_t index, uint64_t *up_bits, uint64_t *dn_bits) {
  return guarded([&] {
    need(b && up_bits && dn_bits, "null argument");
    const int64_t D = b->size();
    if (index < 0 || index >= D) throw mb::RangeError("basis index out of range");
    uint32_t u, d;
    if (b->ms)
      b->ms->ref_state(index, &u, &d);
    else if (b->fn)
      b->fn->ref_state(index, &u, &d);
    else
      b->sec.ref_state(index, &u, &d);
    *up_bits = u;
    *dn_bits = d;
  });
}

int mb_basis_index_of(const mb_basis *b, uint64_t up_bits, uint64_t dn_bits, int64_t *index) {
  return guarded([&] {
    need(b && index, "null argument");
    if (b->fn) {
      const uint64_t mask = (1ull << b->fn->L) - 1;
      const int64_t i = ((up_bits & ~mask) || (dn_bits & ~mask))
                            ? -1
                            : b->fn->ref_index((uint32_t)up_bits, (uint32_t)dn_bits);
      if (i < 0) throw mb::RangeError("state not in basis");
      *index = i;
      return;
    }
    if (b->ms) {
      const auto &m = *b->ms;
      const uint64_t mask = (m.L == 64) ? ~0ull : ((1ull << m.L) - 1);
      const int64_t i = ((up_bits & ~mask) || (dn_bits & ~mask))
                            ? -1
                            : m.ref_index((uint32_t)up_bits, (uint32_t)dn_bits);
      if (i < 0) throw mb::RangeError("state not in basis");
      *index = i;
      return;
    }
    const auto &s = b->sec;
    const uint64_t mask = (s.L == 64) ? ~0ull : ((1ull << s.L) - 1);
    if ((up_bits & ~mask) || (dn_bits & ~mask) || __builtin_popcountll(up_bits) != s.n_up ||
        __builtin_popcountll(dn_bits) != s.n_dn)
      throw mb::RangeError("state not in basis");
    *index = s.ref_index((uint32_t)up_bits, (uint32_t)dn_bits);
  });
}

int mb_basis_product_index(const mb_basis *b, int64_t first, int64_t n, int64_t *product_index) {
  return guarded([&] {
    need(b && product_index, "null argument");
    const auto &s = SEC(b);
    if (first < 0 || n < 0 || first + n > s.dim()) throw mb::RangeError("index range out of bounds");
    for (int64_t i = 0; i < n; ++i) {
      uint32_t u, d;
      s.ref_state(first + i, &u, &d);
      product_index[i] = mb::Sector::colex_rank(u) * s.d_dn + mb::Sector::colex_rank(d);
    }
  });
}

// HubbardModel*::hamiltonian() (hubbard_model.h:13-35) on a fixed-N basis: the
// bond list's hoppings (up then down per bond) and the on-site U terms as an
// expression, compiled like any other
static void hubbard_fixed_n(const mb_basis *b, const std::vector<int> &from,
                            const std::vector<int> &to, double t, double u, mb_hubbard *h) {
  std::vector<double> coeffs;
  std::vector<int64_t> offsets{0};
  std::vector<uint16_t> ops;
  auto term = [&](double c, std::initializer_list<uint16_t> o) {
    coeffs.push_back(c);
    coeffs.push_back(0.0);
    ops.insert(ops.end(), o);
    offsets.push_back((int64_t)ops.size());
  };
  for (size_t k = 0; k < from.size(); ++k)
    for (uint16_t spin : {uint16_t{0}, mb::kOpDown}) {
      const uint16_t i = (uint16_t)from[k], j = (uint16_t)to[k];
      term(-t, {(uint16_t)(spin | i), (uint16_t)(mb::kOpAnnihilation | spin | j)});
      term(-t, {(uint16_t)(spin | j), (uint16_t)(mb::kOpAnnihilation | spin | i)});
    }
  for (int i = 0; i < b->fn->L; ++i)
    term(u, {(uint16_t)i, (uint16_t)(mb::kOpAnnihilation | i), (uint16_t)(mb::kOpDown | i),
             (uint16_t)(mb::kOpAnnihilation | mb::kOpDown | i)});
  h->expr = std::make_unique<mb::ExpressionOperator>(nullptr, nullptr, (int64_t)(offsets.size() - 1),
                                                     coeffs.data(), offsets.data(), ops.data(),
                                                     b->fn.get());
}

static int hubbard_common(const mb_basis *b, std::vector<int> from, std::vector<int> to, double t,
                          double u, mb_hubbard **out, int ndim = 0, const int *sizes = nullptr) {
  return guarded([&] {
    need(b && out, "null argument");
    *out = nullptr;
    if (b->fn) {
      for (size_t k = 0; k < from.size(); ++k)
        need(from[k] >= 0 && from[k] < b->fn->L && to[k] >= 0 && to[k] < b->fn->L,
             "bond endpoint outside the lattice");
      auto h = std::make_unique<mb_hubbard>();
      h->basis = b;
      hubbard_fixed_n(b, from, to, t, u, h.get());
      *out = h.release();
      return;
    }
    SEC(b);
    for (size_t k = 0; k < from.size(); ++k)
      need(from[k] >= 0 && from[k] < b->sec.L && to[k] >= 0 && to[k] < b->sec.L,
           "bond endpoint outside the lattice");
    auto h = std::make_unique<mb_hubbard>();
    h->basis = b;
    h->h.sec = &b->sec;
    h->h.t = t;
    h->h.U = u;
    h->h.bond_from = std::move(from);
    h->h.bond_to = std::move(to);
    h->h.ndim = ndim;
    for (int i = 0; i < ndim; ++i) h->h.lsize[i] = sizes[i];
    const_cast<mb::Sector &>(b->sec).ensure_device();
    mb::build_hubbard(h->h);
    h->real_op = std::make_unique<mb::RealHvOp>(h->h);
    *out = h.release();
  });
}

int mb_hubbard_create(const mb_basis *b, int ndim, const int *sizes, double t, double u,
                      mb_hubbard **out) {
  std::vector<int> from, to;
  int st = guarded([&] {
    need(b && sizes, "null argument");
    need(!b->ms, "use the momentum-space model on a fixed-momentum basis");
    need(ndim >= 1 && ndim <= 3, "ndim must be 1, 2 or 3");
    int prod = 1;
    for (int i = 0; i < ndim; ++i) {
      need(sizes[i] >= 1, "Dimension with size zero is invalid.");
      prod *= sizes[i];
    }
    need(prod == b->orbitals(), "lattice size does not match the basis' orbital count");
    mb::lattice_bonds(ndim, sizes, from, to);
  });
  if (st != MB_OK) return st;
  return hubbard_common(b, std::move(from), std::move(to), t, u, out, ndim, sizes);
}

int mb_hubbard_create_bonds(const mb_basis *b, int n_bonds, const int *from, const int *to,
                            double t, double u, mb_hubbard **out) {
  if (n_bonds < 0 || (n_bonds > 0 && (!from || !to))) {
    g_err = "invalid bond list";
    return MB_ERR_INVALID;
  }
  return hubbard_common(b, std::vector<int>(from, from + n_bonds),
                        std::vector<int>(to, to + n_bonds), t, u, out);
}

void mb_hubbard_destroy(mb_hubbard *h) { delete h; }

int64_t mb_hubbard_dimension(const mb_hubbard *h) { return h ? h->op().dim() : -1; }

int mb_hubbard_shape(const mb_hubbard *h, int64_t *d_up, int64_t *d_dn) {
  return guarded([&] {
    need(h && d_up && d_dn, "null argument");
    if (h->msec() || h->fsec()) {  // a grouped basis is not one product block
      *d_up = h->op().dim();
      *d_dn = 1;
      return;
    }
    *d_up = h->rsec()->d_up;
    *d_dn = h->rsec()->d_dn;
  });
}

int mb_hv_c128(const mb_hubbard *h, const void *x_dev, void *y_dev, mb_stream stream) {
  return guarded([&] {
    need(x_dev && y_dev && x_dev != y_dev, "x and y must be distinct device pointers");
    OP(h).hv((const double *)x_dev, (double *)y_dev, mb::Epi::Plain, mb::EpiArgs{}, S(stream));
  });
}

// The tables of an FP32 H.v: a real-space Hubbard model, or a compiled
// expression of Hubbard form (which runs on the same kernels)
mb::HvTables fp32_tables(const mb_hubbard *h) {
  if (h && h->expr) {
    const mb::HvTables *t = h->expr->hubbard_tables();
    need(t != nullptr, "FP32 H.v needs a real-space Hubbard operator or an expression of Hubbard form");
    return *t;
  }
  return mb::tables_of(H(h));
}

int mb_hv_c64(const mb_hubbard *h, const void *x_dev, void *y_dev, mb_stream stream) {
  return guarded([&] {
    need(x_dev && y_dev && x_dev != y_dev, "x and y must be distinct device pointers");
    mb::launch_hv<float>(fp32_tables(h), 2, (const float *)x_dev, (float *)y_dev,
                         mb::Epi::Plain, mb::EpiArgs{}, S(stream));
  });
}

int mb_hv_c64_multi(const mb_hubbard *h, int nvec, const void *x_dev, void *y_dev,
                    mb_stream stream) {
  return guarded([&] {
    need(x_dev && y_dev && x_dev != y_dev, "x and y must be distinct device pointers");
    need(nvec >= 1 && nvec <= 64, "nvec must be in [1, 64]");
    mb::launch_hv<float>(fp32_tables(h), 2 * nvec, (const float *)x_dev, (float *)y_dev,
                         mb::Epi::Plain, mb::EpiArgs{}, S(stream));
  });
}


int mb_to_device_order_c128(const mb_hubbard *h, const void *x_ref, void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(x_ref && x_dev && x_ref != x_dev, "in and out must be distinct device pointers");
    need(h != nullptr, "null model handle");
    if (const int64_t *map = h->ref_map())
      mb::launch_mom_perm(map, h->op().dim(), (const double *)x_ref, (double *)x_dev, true,
                          S(stream));
    else
      mb::launch_gather_c128(mb::perm_of(*h->rsec()), x_ref, x_dev, S(stream));
  });
}

int mb_to_reference_order_c128(const mb_hubbard *h, const void *x_dev, void *x_ref,
                               mb_stream stream) {
  return guarded([&] {
    need(x_ref && x_dev && x_ref != x_dev, "in and out must be distinct device pointers");
    need(h != nullptr, "null model handle");
    if (const int64_t *map = h->ref_map())
      mb::launch_mom_perm(map, h->op().dim(), (const double *)x_dev, (double *)x_ref, false,
                          S(stream));
    else
      mb::launch_scatter_c128(mb::perm_of(*h->rsec()), x_dev, x_ref, S(stream));
  });
}

int mb_apply_ref_c128(const mb_hubbard *h, const void *x_ref, void *y_ref, void *work,
                      mb_stream stream) {
  return guarded([&] {
    need(x_ref && y_ref && work, "null argument");
    need(h != nullptr, "null model handle");
    if (const int64_t *map = h->ref_map()) {
      const int64_t D = h->op().dim();
      double *w0 = (double *)work, *w1 = w0 + 2 * D;
      mb::launch_mom_perm(map, D, (const double *)x_ref, w0, true, S(stream));
      h->op().hv(w0, w1, mb::Epi::Plain, mb::EpiArgs{}, S(stream));
      mb::launch_mom_perm(map, D, w1, (double *)y_ref, false, S(stream));
      return;
    }
    if (h->on_basis()) {
      const int64_t D = h->op().dim();
      double *w0 = (double *)work, *w1 = w0 + 2 * D;
      mb::launch_gather_c128(mb::perm_of(*h->rsec()), x_ref, w0, S(stream));
      h->op().hv(w0, w1, mb::Epi::Plain, mb::EpiArgs{}, S(stream));
      mb::launch_scatter_c128(mb::perm_of(*h->rsec()), w1, y_ref, S(stream));
      return;
    }
    const auto &hh = H(h);
    char *w0 = (char *)work, *w1 = w0 + (size_t)hh.dim() * 16;
    mb::launch_gather_c128(mb::perm_of(*hh.sec), x_ref, w0, S(stream));
    mb::launch_hv<double>(mb::tables_of(hh), 2, (const double *)w0, (double *)w1, mb::Epi::Plain,
                          mb::EpiArgs{}, S(stream));
    mb::launch_scatter_c128(mb::perm_of(*hh.sec), w1, y_ref, S(stream));
  });
}

int mb_fill_gaussian_c128(const mb_hubbard *h, uint64_t seed, void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(x_dev, "null argument");
    need(h != nullptr, "null model handle");
    if (const int64_t *map = h->ref_map())
      mb::launch_mom_fill_gaussian(map, h->op().dim(), seed, (double *)x_dev, S(stream));
    else
      mb::launch_fill_gaussian_c128(mb::perm_of(*h->rsec()), seed, x_dev, S(stream));
  });
}

int mb_fill_phase_c64_multi(const mb_hubbard *h, uint64_t seed, int nvec, void *x_dev,
                            mb_stream stream) {
  return guarded([&] {
    need(x_dev, "null argument");
    need(nvec >= 1 && nvec <= 64, "nvec out of range");
    mb::launch_fill_phase(mb::perm_of(*H(h).sec), seed, nvec, true, x_dev, S(stream));
  });
}

int mb_fill_phase_c128_multi(const mb_hubbard *h, uint64_t seed, int nvec, void *x_dev,
                             mb_stream stream) {
  return guarded([&] {
    need(x_dev, "null argument");
    need(h != nullptr, "null model handle");
    need(nvec >= 1 && nvec <= 64, "nvec out of range");
    if (!h->on_basis() && !h->mom) {
      mb::launch_fill_phase(mb::perm_of(*H(h).sec), seed, nvec, false, x_dev, S(stream));
      return;
    }
    mb::DevBuf own;
    const int64_t *ref = ref_positions(h, own, S(stream));
    const int64_t D = h->op().dim();
    for (int r = 0; r < nvec; ++r)
      mb::launch_fill_phase_ref(ref, D, seed, nvec, r, false, (double *)x_dev + 2 * r * D,
                                S(stream));
    MB_HIP(hipStreamSynchronize(S(stream)));  // own is freed on return
  });
}

int mb_lanczos_pass_one(const mb_hubbard *h, const void *b_dev, int k, double *alphas,
                        double *betas, int *steps, double *b_norm, mb_stream stream) {
  return guarded([&] {
    need(b_dev && alphas && betas && steps && b_norm, "null argument");
    need(k >= 0, "k must be non-negative");
    auto r = mb::lanczos_pass_one(OP(h), b_dev, k, S(stream));
    *steps = r.steps;
    *b_norm = r.b_norm;
    for (int i = 0; i < k; ++i) {
      alphas[i] = i < (int)r.alphas.size() ? r.alphas[i] : 0.0;
      betas[i] = i < (int)r.betas.size() ? r.betas[i] : 0.0;
    }
  });
}

int mb_lanczos_pass_one_noalloc(const mb_hubbard *h, const void *b_dev, int k, double *alphas,
                                double *betas, int *steps, double *b_norm, mb_stream stream) {
  return guarded([&] {
    need(h && b_dev && alphas && betas && steps && b_norm, "null argument");
    need(k >= 0, "k must be non-negative");
    need(h->real_op && !h->comp && !h->mom && !h->expr,
         "the no-allocation Lanczos runs on real-space Hubbard operators only");
    auto r = mb::lanczos_pass_one_noalloc(OP(h), b_dev, k, S(stream));
    *steps = r.steps;
    *b_norm = r.b_norm;
    for (int i = 0; i < k; ++i) {
      alphas[i] = i < (int)r.alphas.size() ? r.alphas[i] : 0.0;
      betas[i] = i < (int)r.betas.size() ? r.betas[i] : 0.0;
    }
  });
}

int mb_find_min_eigenpair(const mb_hubbard *h, const void *b_dev, int k, double *value,
                          void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(b_dev && value, "null argument");
    need(k >= 0, "k must be non-negative");
    *value = mb::find_min_eigenpair(OP(h), b_dev, k, x_dev, S(stream));
  });
}

int mb_lanczos_pass_one_ex(const mb_hubbard *h, const void *b_dev, int k, int flags, double *alphas,
                           double *betas, int *steps, double *b_norm, mb_stream stream) {
  return guarded([&] {
    need(b_dev && alphas && betas && steps && b_norm, "null argument");
    need(k >= 0, "k must be non-negative");
    need((flags & ~MB_LANCZOS_REORTHOGONALIZE) == 0, "unknown Lanczos flags");
    auto r = (flags & MB_LANCZOS_REORTHOGONALIZE)
                 ? mb::lanczos_pass_one_reorth(OP(h), b_dev, k, S(stream))
                 : mb::lanczos_pass_one(OP(h), b_dev, k, S(stream));
    *steps = r.steps;
    *b_norm = r.b_norm;
    for (int i = 0; i < k; ++i) {
      alphas[i] = i < (int)r.alphas.size() ? r.alphas[i] : 0.0;
      betas[i] = i < (int)r.betas.size() ? r.betas[i] : 0.0;
    }
  });
}

int mb_find_min_eigenpair_ex(const mb_hubbard *h, const void *b_dev, int k, int flags,
                             double *value, void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(b_dev && value, "null argument");
    need(k >= 0, "k must be non-negative");
    need((flags & ~MB_LANCZOS_REORTHOGONALIZE) == 0, "unknown Lanczos flags");
    *value = (flags & MB_LANCZOS_REORTHOGONALIZE)
                 ? mb::find_min_eigenpair_reorth(OP(h), b_dev, k, x_dev, S(stream))
                 : mb::find_min_eigenpair(OP(h), b_dev, k, x_dev, S(stream));
  });
}

int mb_time_evolve(const mb_hubbard *h, const void *psi_dev, double t, int krylov_steps,
                   void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    mb::time_evolve(OP(h), psi_dev, t, krylov_steps, out_dev, S(stream));
  });
}

int mb_time_evolve_ex(const mb_hubbard *h, const void *psi_dev, double t, int krylov_steps, int flags,
                      void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    need((flags & ~MB_EVOLVE_FP32) == 0, "unknown time evolution flags");
    mb::time_evolve(OP(h), psi_dev, t, krylov_steps, out_dev, S(stream), (flags & MB_EVOLVE_FP32) != 0);
  });
}

int mb_time_evolve_steps_ex(const mb_hubbard *h, const void *psi_dev, double t0, double t1, double dt,
                            int krylov_steps, int flags, void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    need((flags & ~MB_EVOLVE_FP32) == 0, "unknown time evolution flags");
    mb::time_evolve_steps(OP(h), psi_dev, t0, t1, dt, krylov_steps, out_dev, S(stream),
                          (flags & MB_EVOLVE_FP32) != 0);
  });
}

int mb_time_evolve_steps(const mb_hubbard *h, const void *psi_dev, double t0, double t1,
                         double dt, int krylov_steps, void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    mb::time_evolve_steps(OP(h), psi_dev, t0, t1, dt, krylov_steps, out_dev, S(stream));
  });
}

int mb_tridiag_eig(int m, const double *alphas, const double *betas, double *w, double *V) {
  return guarded([&] {
    need(m >= 0 && (m == 0 || (alphas && w)) && (m <= 1 || betas), "null argument");
    mb::tridiag_eig(m, alphas, betas, w, V);
  });
}

int mb_kpm_dos_moments(const mb_hubbard *h, double a, double b, int M, int nvec, uint64_t seed,
                       int fp32, double *mu, mb_stream stream) {
  return guarded([&] {
    need(mu != nullptr, "null argument");
    need(h != nullptr, "null model handle");
    if (!h->on_basis() && !h->mom) {
      mb::kpm_dos_moments(H(h), *H(h).sec, a, b, M, nvec, seed, fp32 != 0, mu, S(stream));
      return;
    }
    if (h->expr && h->expr->hubbard()) {  // an expression of Hubbard form
      const mb::Hubbard &eh = *h->expr->hubbard();
      mb::kpm_dos_moments(eh, *eh.sec, a, b, M, nvec, seed, fp32 != 0, mu, S(stream));
      return;
    }
    need(fp32 == 0, "FP32 KPM moments need a real-space Hubbard operator or an expression of Hubbard form");
    mb::DevBuf own;
    const int64_t *ref = ref_positions(h, own, S(stream));
    mb::kpm_dos_moments_op(h->op(), ref, a, b, M, nvec, seed, mu, S(stream));
  });
}

int mb_jackson_kernel(int N, double *g) {
  return guarded([&] {
    need(N >= 0 && g, "invalid argument");
    const std::vector<double> v = mb::jackson_kernel(N);
    std::copy(v.begin(), v.end(), g);
  });
}

int mb_rescaling_from_bounds(double e_min, double e_max, double padding, double *a, double *b) {
  return guarded([&] {
    need(a && b, "null argument");
    // rescaling.h:33-38
    *a = (e_max - e_min) / 2.0 * (1.0 + padding);
    *b = (e_max + e_min) / 2.0;
  });
}

int mb_batch_create(int n, const mb_hubbard *const *models, mb_batch **out) {
  return guarded([&] {
    need(out && models && n > 0, "invalid batch");
    *out = nullptr;
    std::vector<const mb::Hubbard *> v;
    for (int i = 0; i < n; ++i) v.push_back(&H(models[i]));
    auto b = std::make_unique<mb_batch>();
    b->b = std::make_unique<mb::Batch>(std::move(v));
    *out = b.release();
  });
}

void mb_batch_destroy(mb_batch *b) { delete b; }

int64_t mb_batch_footprint(const mb_batch *b) { return b ? b->b->footprint_bytes() : -1; }

int mb_batch_fill_gaussian(mb_batch *b, const uint64_t *seeds, mb_stream stream) {
  return guarded([&] {
    need(b && seeds, "null argument");
    b->b->fill_gaussian(seeds, S(stream));
  });
}

int mb_batch_set_start(mb_batch *b, int i, const void *b_dev, mb_stream stream) {
  return guarded([&] {
    need(b && b_dev, "null argument");
    if (i < 0) throw mb::RangeError("instance index out of range");
    b->b->set_start((size_t)i, b_dev, S(stream));
  });
}

int mb_batch_set_mapping(mb_batch *b, int tile, int block) {
  return guarded([&] {
    need(b != nullptr, "null argument");
    b->b->set_mapping(tile, block);
  });
}

int mb_batch_set_persistent(mb_batch *b, int64_t max_dim) {
  return guarded([&] {
    need(b != nullptr, "null argument");
    need(max_dim >= 0, "max_dim must be non-negative");
    b->b->set_persist_max_dim(max_dim);
  });
}

int mb_batch_autotune(mb_batch *b, int k_probe, int *tile, int *block, mb_stream stream) {
  return guarded([&] {
    need(b && k_probe > 0, "invalid argument");
    auto m = b->b->autotune(k_probe, S(stream));
    if (tile) *tile = m.first;
    if (block) *block = m.second;
  });
}

int mb_batch_lanczos(mb_batch *b, int k, int use_graph, double *alphas, double *betas,
                     int *steps, double *b_norms, mb_stream stream) {
  return guarded([&] {
    need(b && alphas && betas && steps && b_norms && k > 0, "invalid argument");
    b->b->run(k, use_graph != 0, S(stream));
    std::vector<double> al, be, bn;
    std::vector<int> st;
    b->b->results(k, al, be, st, bn, S(stream));
    for (size_t i = 0; i < bn.size(); ++i)
      if (bn[i] <= mb::kTol) throw std::runtime_error("Input vector has zero norm");
    std::copy(al.begin(), al.end(), alphas);
    std::copy(be.begin(), be.end(), betas);
    std::copy(st.begin(), st.end(), steps);
    std::copy(bn.begin(), bn.end(), b_norms);
  });
}

int mb_batch_ground_energies(mb_batch *b, int k, int use_graph, double *e0, mb_stream stream) {
  return guarded([&] {
    need(b && e0 && k > 0, "invalid argument");
    auto v = b->b->ground_energies(k, use_graph != 0, S(stream));
    std::copy(v.begin(), v.end(), e0);
  });
}

}  // extern "C"

// ------------------------------------------------------------ serving loop

static mb::ServeRequest to_internal(const mb_request &r) {
  mb::ServeRequest q;
  q.kind = r.kind;
  q.ndim = r.ndim;
  for (int i = 0; i < 3; ++i) q.sizes[i] = r.sizes[i];
  q.n_up = r.n_up;
  q.n_dn = r.n_dn;
  q.t = r.t;
  q.u = r.u;
  q.steps = r.steps;
  q.moments = r.moments;
  q.nvec = r.nvec;
  q.t_final = r.t_final;
  q.dt = r.dt;
  q.seed = r.seed;
  q.arrival = r.arrival;
  return q;
}

int mb_server_create(int slots, int64_t arena_bytes, int64_t lane_bytes, mb_server **out) {
  return guarded([&] {
    need(out != nullptr, "null output");
    *out = nullptr;
    need(lane_bytes >= 0 && lane_bytes < arena_bytes, "lane_bytes must be in [0, arena_bytes)");
    auto s = std::make_unique<mb_server>();
    s->s = std::make_unique<mb::Server>(slots, arena_bytes, lane_bytes);
    *out = s.release();
  });
}

void mb_server_destroy(mb_server *s) { delete s; }

int mb_server_submit(mb_server *s, const mb_request *r, int64_t *id) {
  return guarded([&] {
    need(s && r && id, "null argument");
    need(r->kind >= MB_REQ_GROUND && r->kind <= MB_REQ_EVOLVE, "unknown request kind");
    need(r->steps >= 1, "steps must be positive");
    if (r->kind == MB_REQ_KPM) need(r->nvec >= 1 && r->nvec <= 64, "nvec must be in [1, 64]");
    if (r->kind == MB_REQ_EVOLVE) need(r->t_final >= 0.0 && r->dt > 0.0, "need t_final >= 0, dt > 0");
    *id = s->s->submit(to_internal(*r));
  });
}

int mb_server_wait(mb_server *s, int64_t id, mb_result *out, double *moments, int moments_cap) {
  return guarded([&] {
    need(s && out, "null argument");
    const mb::ServeResult &r = s->s->wait(id);
    std::memset(out, 0, sizeof(*out));
    out->status = r.status;
    out->slot = r.slot;
    out->dim = r.dim;
    out->footprint = r.footprint;
    out->peak = r.peak;
    out->e0 = r.e0;
    out->e_min = r.e_min;
    out->e_max = r.e_max;
    out->overlap_re = r.ov_re;
    out->overlap_im = r.ov_im;
    out->norm = r.norm;
    out->arrival = r.arrival;
    out->admitted = r.admitted;
    out->finished = r.finished;
    const int n = moments ? std::min<int>(moments_cap, (int)r.mu.size()) : 0;
    for (int i = 0; i < n; ++i) moments[i] = r.mu[i];
    out->n_moments = n;
    std::snprintf(out->error, sizeof(out->error), "%s", r.error.c_str());
  });
}

int mb_server_drain(mb_server *s) {
  return guarded([&] {
    need(s != nullptr, "null argument");
    s->s->drain();
  });
}

double mb_server_now(const mb_server *s) { return s ? s->s->now() : -1.0; }

int64_t mb_request_footprint(const mb_request *r) {
  if (!r) return -1;
  try {
    return mb::request_footprint(to_internal(*r));
  } catch (...) {
    return -1;
  }
}

// ------------------------------------------------------------ observables

int mb_opposite_spin_correlation(const mb_hubbard *h, const void *x_dev, double *g,
                                 mb_stream stream) {
  return guarded([&] {
    need(h && x_dev && g, "null argument");
    const auto v = mb::opposite_spin_correlation(H(h), x_dev, S(stream));
    for (size_t i = 0; i < v.size(); ++i) g[i] = v[i];
  });
}

int mb_energy_residual(const mb_hubbard *h, const void *x_dev, double *energy, double *residual,
                       mb_stream stream) {
  return guarded([&] {
    need(h && x_dev && energy && residual, "null argument");
    mb::energy_residual(OP(h), x_dev, energy, residual, S(stream));
  });
}

// ------------------------------------------------------------ momentum sectors

int mb_momentum_basis_create(int orbitals, int particles, int spin_projection, int ndim,
                             const int *sizes, const int *momentum, mb_basis **out) {
  return guarded([&] {
    need(out && sizes && momentum, "null argument");
    *out = nullptr;
    need(ndim >= 1 && ndim <= 3, "ndim must be 1, 2 or 3");
    int prod = 1;
    for (int i = 0; i < ndim; ++i) {
      need(sizes[i] >= 1, "Dimension with size zero is invalid.");
      prod *= sizes[i];
    }
    need(prod == orbitals, "index.size() must equal orbitals");
    need(particles >= 0 && particles <= 2 * orbitals, "particles must be in [0, 2*orbitals]");
    need(std::abs(spin_projection) <= particles, "|spin_projection| must not exceed particles");
    need((particles + spin_projection) % 2 == 0, "particles + spin_projection must be even");
    const int up = (particles + spin_projection) / 2, dn = particles - up;
    auto b = std::make_unique<mb_basis>();
    b->ms = std::make_unique<mb::MomentumSector>();
    mb::build_momentum_sector(*b->ms, ndim, sizes, up, dn, momentum);
    *out = b.release();
  });
}

int mb_momentum_hubbard_create(const mb_basis *b, int ndim, const int *sizes, double t, double u,
                               const int *momentum, mb_hubbard **out) {
  return guarded([&] {
    need(b && sizes && momentum && out, "null argument");
    *out = nullptr;
    need(b->ms != nullptr, "the basis is not a fixed-momentum basis");
    const auto &m = *b->ms;
    need(ndim == m.ndim, "Target momentum dimensionality must match lattice dimensionality.");
    int K = 0, stride = 1;
    for (int i = 0; i < ndim; ++i) {
      need(sizes[i] == m.dims[i], "lattice size does not match the basis");
      need(momentum[i] >= 0 && momentum[i] < sizes[i],
           "Target momentum components must be less than lattice size.");
      K += stride * momentum[i];
      stride *= sizes[i];
    }
    need(K == m.K, "target momentum does not match the basis");
    auto h = std::make_unique<mb_hubbard>();
    h->basis = b;
    h->mom = std::make_unique<mb::MomentumHubbard>(m, t, u);
    const_cast<mb::MomentumSector &>(m).ensure_device();
    *out = h.release();
  });
}

// ------------------------------------------------------------ optical conductivity

int mb_current_create(const mb_basis *source, const mb_basis *target, int ndim, const int *sizes,
                      double t, int64_t Q, int direction, mb_current **out) {
  return guarded([&] {
    need(source && target && sizes && out, "null argument");
    *out = nullptr;
    need(source->ms && target->ms, "the bases must be fixed-momentum bases");
    const auto &S = *source->ms, &T = *target->ms;
    need(ndim == S.ndim, "lattice dimensionality does not match the bases");
    for (int i = 0; i < ndim; ++i) need(sizes[i] == S.dims[i], "lattice size does not match the bases");
    need(Q >= 0 && Q < S.L, "Q must be a flat momentum index");
    need(direction >= 0 && direction < ndim, "Direction must be less than lattice dimensionality.");
    auto j = std::make_unique<mb_current>();
    j->source = source;
    j->target = target;
    j->j = std::make_unique<mb::CurrentOperator>(S, T, t, (int)Q, direction);
    *out = j.release();
  });
}

void mb_current_destroy(mb_current *j) { delete j; }

int mb_current_apply(const mb_current *j, const void *x_src, void *y_tgt, mb_stream stream) {
  return guarded([&] {
    need(j && x_src && y_tgt && x_src != y_tgt, "null or aliased argument");
    j->j->apply((const double *)x_src, (double *)y_tgt, S(stream));
  });
}

int mb_current_adjoint_apply(const mb_current *j, const void *y_tgt, void *x_src, mb_stream stream) {
  return guarded([&] {
    need(j && x_src && y_tgt && x_src != y_tgt, "null or aliased argument");
    j->j->adjoint_apply((const double *)y_tgt, (double *)x_src, S(stream));
  });
}

int mb_imaginary_time_evolve(const mb_hubbard *h, const void *psi_dev, double tau, int krylov_steps,
                             void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    mb::imaginary_time_evolve(OP(h), psi_dev, tau, krylov_steps, out_dev, S(stream));
  });
}

int mb_optical_correlator(const mb_hubbard *hk, const mb_hubbard *hkq, const mb_current *j,
                          int samples, double beta, double dt, int steps, int krylov_steps,
                          uint64_t seed, double *corr, mb_stream stream) {
  return guarded([&] {
    need(hk && hkq && j && corr, "null argument");
    // any operator on the current operator's bases: the factorized momentum-space
    // Hubbard model or a compiled expression on the same fixed-momentum bases
    need(hk->msec() && hkq->msec(), "the Hamiltonians must act on fixed-momentum bases");
    need(hk->msec() == j->source->ms.get() && hkq->msec() == j->target->ms.get(),
         "the Hamiltonians must act on the current operator's source and target bases");
    auto &m = const_cast<mb::MomentumSector &>(*hk->msec());
    m.ensure_device();
    const auto C = mb::optical_correlator(hk->op(), hkq->op(), *j->j, m.d_ref_of_pos.as<int64_t>(),
                                          samples, beta, dt, steps, krylov_steps, seed, S(stream));
    for (int i = 0; i < steps; ++i) {
      corr[2 * i] = C[i].real();
      corr[2 * i + 1] = C[i].imag();
    }
  });
}

int mb_optical_conductivity(const double *corr, int n, double dt, double beta, double volume,
                            double *omega, double *sigma, int capacity, int *n_freqs,
                            double *sum_rule) {
  return guarded([&] {
    need(corr && n_freqs, "null argument");
    std::vector<std::complex<double>> c(n > 0 ? n : 0);
    for (int i = 0; i < n; ++i) c[i] = {corr[2 * i], corr[2 * i + 1]};
    const mb::Conductivity r = mb::optical_conductivity(c, dt, beta, volume);
    *n_freqs = (int)r.omega.size();
    if (!omega && !sigma) return;  // size query
    need(omega && sigma && capacity >= *n_freqs, "output arrays too small");
    for (size_t i = 0; i < r.omega.size(); ++i) {
      omega[i] = r.omega[i];
      sigma[2 * i] = r.sigma[i].real();
      sigma[2 * i + 1] = r.sigma[i].imag();
    }
    if (sum_rule) *sum_rule = r.sum_rule;
  });
}

// ------------------------------------------------------------ generic expressions

int mb_expression_create(const mb_basis *b, int64_t n_terms, const double *coeffs,
                         const int64_t *offsets, const uint16_t *ops, mb_hubbard **out) {
  return guarded([&] {
    need(b && out && n_terms >= 0, "null argument");
    need(n_terms == 0 || (coeffs && offsets), "null term arrays");
    *out = nullptr;
    need(n_terms == 0 || offsets[0] == 0, "offsets[0] must be 0");
    need(n_terms == 0 || offsets[n_terms] == 0 || ops, "null operator array");
    auto h = std::make_unique<mb_hubbard>();
    h->basis = b;
    const int64_t zero = 0;
    if (b->ms) {
      auto &m = const_cast<mb::MomentumSector &>(*b->ms);
      h->expr = std::make_unique<mb::ExpressionOperator>(nullptr, &m, n_terms, coeffs,
                                                         n_terms ? offsets : &zero, ops);
    } else if (b->fn) {
      h->expr = std::make_unique<mb::ExpressionOperator>(nullptr, nullptr, n_terms, coeffs,
                                                         n_terms ? offsets : &zero, ops,
                                                         b->fn.get());
    } else {
      h->expr = std::make_unique<mb::ExpressionOperator>(&b->sec, nullptr, n_terms, coeffs,
                                                         n_terms ? offsets : &zero, ops);
    }
    *out = h.release();
  });
}

int mb_compute_matrix_elements(const mb_hubbard *h, mb_sparse **out, mb_stream stream) {
  return guarded([&] {
    need(h && out, "null argument");
    *out = nullptr;
    need(h->expr != nullptr, "the sparse export needs an expression operator");
    auto sp = std::make_unique<mb_sparse>();
    sp->n_rows = sp->n_cols = (uint64_t)h->expr->dim();
    h->expr->export_coo(sp->rows, sp->cols, sp->vals, S(stream));
    *out = sp.release();
  });
}

namespace {

mb::BasisRef basis_ref(const mb_basis *b) {
  mb::BasisRef r;
  if (b->ms)
    r.mom = b->ms.get();
  else if (b->fn)
    r.fn = b->fn.get();
  else
    r.real = &b->sec;
  return r;
}

void check_terms(int64_t n_terms, const double *coeffs, const int64_t *offsets,
                 const uint16_t *ops) {
  need(n_terms >= 0, "negative term count");
  need(n_terms == 0 || (coeffs && offsets), "null term arrays");
  need(n_terms == 0 || offsets[0] == 0, "offsets[0] must be 0");
  need(n_terms == 0 || offsets[n_terms] == 0 || ops, "null operator array");
}

}  // namespace

int mb_compute_rectangular_matrix_elements(const mb_basis *row_basis, const mb_basis *col_basis,
                                           int64_t n_terms, const double *coeffs,
                                           const int64_t *offsets, const uint16_t *ops,
                                           mb_sparse **out, mb_stream stream) {
  return guarded([&] {
    need(row_basis && col_basis && out, "null argument");
    *out = nullptr;
    check_terms(n_terms, coeffs, offsets, ops);
    auto sp = std::make_unique<mb_sparse>();
    sp->n_rows = (uint64_t)row_basis->size();
    sp->n_cols = (uint64_t)col_basis->size();
    mb::rectangular_matrix_elements(basis_ref(row_basis), basis_ref(col_basis), n_terms, coeffs,
                                    offsets, ops, sp->rows, sp->cols, sp->vals, S(stream));
    *out = sp.release();
  });
}

int mb_compute_vector_elements(const mb_basis *basis, int64_t n_terms, const double *coeffs,
                               const int64_t *offsets, const uint16_t *ops, double *out,
                               mb_stream stream) {
  return guarded([&] {
    need(basis && out, "null argument");
    check_terms(n_terms, coeffs, offsets, ops);
    mb::vector_elements(basis_ref(basis), n_terms, coeffs, offsets, ops,
                        reinterpret_cast<std::complex<double> *>(out), S(stream));
  });
}

int mb_operator_negated(const mb_hubbard *a, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Negated, a, nullptr, 0.0, 0.0, out);
}
int mb_operator_scaled(const mb_hubbard *a, double re, double im, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Scaled, a, nullptr, re, im, out);
}
int mb_operator_shifted(const mb_hubbard *a, double re, double im, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Shifted, a, nullptr, re, im, out);
}
int mb_operator_neg_shifted(const mb_hubbard *a, double re, double im, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::NegShift, a, nullptr, re, im, out);
}
int mb_operator_sum(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Sum, a, b, 0.0, 0.0, out);
}
int mb_operator_difference(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Difference, a, b, 0.0, 0.0, out);
}
int mb_operator_composed(const mb_hubbard *a, const mb_hubbard *b, mb_hubbard **out) {
  return make_composite(mb::CompositeOp::Kind::Composed, a, b, 0.0, 0.0, out);
}
int mb_operator_rescaled(const mb_hubbard *a, double scale, double shift, mb_hubbard **out) {
  const int st = guarded([&] {
    need(scale != 0.0 && std::isfinite(scale) && std::isfinite(shift),
         "the scale must be finite and nonzero");
  });
  if (st != MB_OK) return st;
  return make_composite(mb::CompositeOp::Kind::Rescaled, a, nullptr, scale, shift, out);
}
int mb_operator_identity(const mb_basis *basis, mb_hubbard **out) {
  return guarded([&] {
    need(basis && out, "null argument");
    *out = nullptr;
    auto h = std::make_unique<mb_hubbard>();
    h->basis = basis;
    h->comp = std::make_unique<mb::CompositeOp>(mb::CompositeOp::Kind::Identity, nullptr,
                                                nullptr, std::complex<double>(0.0, 0.0),
                                                basis->size());
    *out = h.release();
  });
}

int mb_sparse_create(uint64_t n_rows, uint64_t n_cols, int64_t nnz, const uint64_t *rows,
                     const uint64_t *cols, const double *vals, mb_sparse **out) {
  return guarded([&] {
    need(out && nnz >= 0 && (nnz == 0 || (rows && cols && vals)), "null argument");
    *out = nullptr;
    auto sp = std::make_unique<mb_sparse>();
    sp->n_rows = n_rows;
    sp->n_cols = n_cols;
    for (int64_t i = 0; i < nnz; ++i) {
      need(rows[i] < n_rows && cols[i] < n_cols, "entry outside the matrix");
      if (i > 0)
        need(cols[i] > cols[i - 1] || (cols[i] == cols[i - 1] && rows[i] > rows[i - 1]),
             "entries must be column-major with rows ascending and unique");
      sp->rows.push_back(rows[i]);
      sp->cols.push_back(cols[i]);
      sp->vals.emplace_back(vals[2 * i], vals[2 * i + 1]);
    }
    *out = sp.release();
  });
}

void mb_sparse_destroy(mb_sparse *sp) { delete sp; }

int mb_sparse_shape(const mb_sparse *sp, uint64_t *n_rows, uint64_t *n_cols, int64_t *nnz) {
  return guarded([&] {
    need(sp && n_rows && n_cols && nnz, "null argument");
    *n_rows = sp->n_rows;
    *n_cols = sp->n_cols;
    *nnz = (int64_t)sp->vals.size();
  });
}

int mb_sparse_entries(const mb_sparse *sp, uint64_t *rows, uint64_t *cols, double *vals) {
  return guarded([&] {
    need(sp && (sp->vals.empty() || (rows && cols && vals)), "null argument");
    for (size_t i = 0; i < sp->vals.size(); ++i) {
      rows[i] = sp->rows[i];
      cols[i] = sp->cols[i];
      vals[2 * i] = sp->vals[i].real();
      vals[2 * i + 1] = sp->vals[i].imag();
    }
  });
}

namespace {
// little-endian fixed-width fields of cereal's BinaryOutputArchive
template <typename T>
void put(std::string &b, T v) {
  b.append(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <typename T>
T get(std::ifstream &f) {
  T v;
  f.read(reinterpret_cast<char *>(&v), sizeof(T));
  if (!f) throw std::runtime_error("truncated sparse-matrix archive");
  return v;
}
}  // namespace

int mb_sparse_save_cereal(const mb_sparse *sp, const char *path) {
  return guarded([&] {
    need(sp && path, "null argument");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("Failed to open file for writing: ") + path);
    std::string b;
    b.reserve(28 + 32 * sp->vals.size());
    put<uint32_t>(b, 0);  // cereal class version of a versioned save
    put<uint64_t>(b, sp->n_rows);
    put<uint64_t>(b, sp->n_cols);
    put<uint64_t>(b, (uint64_t)sp->vals.size());
    for (size_t i = 0; i < sp->vals.size(); ++i) {
      put<uint64_t>(b, sp->rows[i]);
      put<uint64_t>(b, sp->cols[i]);
      put<double>(b, sp->vals[i].real());
      put<double>(b, sp->vals[i].imag());
    }
    f.write(b.data(), (std::streamsize)b.size());
    if (!f) throw std::runtime_error(std::string("write failed: ") + path);
  });
}

int mb_sparse_load_cereal(const char *path, mb_sparse **out) {
  return guarded([&] {
    need(path && out, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("Failed to open file for reading: ") + path);
    auto sp = std::make_unique<mb_sparse>();
    (void)get<uint32_t>(f);  // class version
    sp->n_rows = get<uint64_t>(f);
    sp->n_cols = get<uint64_t>(f);
    const uint64_t nz = get<uint64_t>(f);
    // arma_cereal.h load(): t(row, col) = v per triplet, later ones overwrite
    std::map<std::pair<uint64_t, uint64_t>, std::complex<double>> m;
    for (uint64_t i = 0; i < nz; ++i) {
      const uint64_t r = get<uint64_t>(f), c = get<uint64_t>(f);
      const double re = get<double>(f), im = get<double>(f);
      if (r >= sp->n_rows || c >= sp->n_cols) throw std::out_of_range("entry outside the matrix");
      m[{c, r}] = {re, im};
    }
    for (const auto &[k, v] : m) {
      if (v == 0.0) continue;  // SpMat does not store explicit zeros
      sp->cols.push_back(k.first);
      sp->rows.push_back(k.second);
      sp->vals.push_back(v);
    }
    *out = sp.release();
  });
}

// ------------------------------------------------------------ spectral bounds

int mb_power_method(const mb_hubbard *h, const void *seed_dev, int iterations, double *value,
                    mb_stream stream) {
  return guarded([&] {
    need(seed_dev && value, "null argument");
    *value = mb::power_method(OP(h), seed_dev, iterations, 1.0, 0.0, S(stream));
  });
}

int mb_estimate_bounds(const mb_hubbard *h, const void *seed_dev, int power_iterations,
                       double spectral_padding, double *alpha, double *beta, mb_stream stream) {
  return guarded([&] {
    need(seed_dev && alpha && beta, "null argument");
    need(power_iterations >= 0, "power_iterations must be non-negative");
    const auto b = mb::estimate_bounds(OP(h), seed_dev, power_iterations, spectral_padding,
                                       S(stream));
    *alpha = b.first;
    *beta = b.second;
  });
}

int mb_exp_apply(const mb_hubbard *h, const void *v_dev, int krylov_steps, void *out_dev,
                 mb_stream stream) {
  return guarded([&] {
    need(v_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    // exp(H) v = exp(-tau H) v at tau = -1 (the same Krylov solve)
    mb::imaginary_time_evolve(OP(h), v_dev, -1.0, krylov_steps, out_dev, S(stream));
  });
}

// ------------------------------------------------------------ generic Krylov solve

int mb_lanczos_solve(const mb_hubbard *h, const void *b_dev, int k, mb_tridiag_solver solver,
                     void *user, void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(b_dev && x_dev && solver, "null argument");
    need(k >= 0, "k must be non-negative");
    auto f = [&](const std::vector<double> &al, const std::vector<double> &be) {
      const int m = (int)al.size();
      std::vector<double> buf(2 * (size_t)m, 0.0);
      if (solver(m, al.data(), be.data(), buf.data(), user) != 0)
        throw std::runtime_error("tridiagonal solver callback failed");
      std::vector<std::complex<double>> y(m);
      for (int i = 0; i < m; ++i) y[i] = {buf[2 * i], buf[2 * i + 1]};
      return y;
    };
    mb::krylov_solve(OP(h), b_dev, k, f, x_dev, S(stream));
  });
}

int mb_find_max_eigenpair(const mb_hubbard *h, const void *b_dev, int k, double *value,
                          void *x_dev, mb_stream stream) {
  return guarded([&] {
    need(b_dev && value, "null argument");
    need(k >= 0, "k must be non-negative");
    *value = mb::find_max_eigenpair(OP(h), b_dev, k, x_dev, S(stream));
  });
}

int mb_time_evolve_steps_cb(const mb_hubbard *h, const void *psi_dev, double t0, double t1,
                            double dt, int krylov_steps, int flags, mb_step_callback callback,
                            void *user, void *out_dev, mb_stream stream) {
  return guarded([&] {
    need(psi_dev && out_dev, "null argument");
    need(krylov_steps >= 0, "krylov_steps must be non-negative");
    need((flags & ~MB_EVOLVE_FP32) == 0, "unknown time evolution flags");
    mb::StepCallback cb;
    if (callback)
      cb = [&](double t, const void *state) {
        if (callback(t, state, user) != 0) throw std::runtime_error("step callback failed");
      };
    mb::time_evolve_steps(OP(h), psi_dev, t0, t1, dt, krylov_steps, out_dev, S(stream),
                          (flags & MB_EVOLVE_FP32) != 0, cb);
  });
}

// ------------------------------------------------------------ KPM projector

int mb_projector_moments(int M, double epsilon, double *mu) {
  return guarded([&] {
    need(mu != nullptr, "null argument");
    const std::vector<double> v = mb::projector_moments(M, epsilon);
    std::copy(v.begin(), v.end(), mu);
  });
}

int mb_lorentz_kernel(int N, double lambda, double *g) {
  return guarded([&] {
    need(g != nullptr, "null argument");
    const std::vector<double> v = mb::lorentz_kernel(N, lambda);
    std::copy(v.begin(), v.end(), g);
  });
}

int mb_kpm_projector_apply(const mb_hubbard *h, int M, double e_min, double e_max,
                           double e_fermi, double padding, const void *v_dev, void *out_dev,
                           mb_stream stream) {
  return guarded([&] {
    need(v_dev && out_dev && v_dev != out_dev, "null or aliased argument");
    mb::kpm_projector_apply(OP(h), M, e_min, e_max, e_fermi, padding, v_dev, out_dev, S(stream));
  });
}

// ------------------------------------------------------------ fixed-N bases

int mb_fixed_n_basis_create(int orbitals, int particles, mb_basis **out) {
  return guarded([&] {
    need(out != nullptr, "null argument");
    *out = nullptr;
    auto b = std::make_unique<mb_basis>();
    b->fn = std::make_unique<mb::FixedNBasis>();
    mb::build_fixed_n_basis(*b->fn, orbitals, particles);
    *out = b.release();
  });
}
