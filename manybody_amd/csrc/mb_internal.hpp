// Internal declarations shared by the host code and the HIP kernels of
// libmanybody_hip.so.  Not part of the public C ABI (include/manybody_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mb {

// tolerances::tolerance<double>() — src/utils/tolerances.h:7-10
constexpr double kTol = 1000.0 * 2.220446049250313e-16;
constexpr int kMaxSites = 32;

struct RangeError : std::out_of_range {
  using std::out_of_range::out_of_range;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NoMemError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define MB_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::mb::DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

// Owning device buffer (hipMalloc / hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

template <typename T>
DevBuf upload(const std::vector<T> &v) {
  DevBuf b(v.size() * sizeof(T) + 1);
  if (!v.empty()) MB_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return b;
}

// Binomial table C(n, k) for n, k <= kMaxSites (fits in uint64).
struct Binom {
  uint64_t c[kMaxSites + 1][kMaxSites + 1];
  Binom();
  uint64_t operator()(int n, int k) const {
    if (n < 0 || k < 0 || k > n) return 0;
    return c[n][k];
  }
};
const Binom &binom();

// ---------------------------------------------------------------- basis
// One fixed-(N_up, N_dn) sector.  Host side keeps the species strings in
// combinatorial (colex) rank order, which is ascending bitstring order.
struct Sector {
  int L = 0, n_up = 0, n_dn = 0;
  int64_t d_up = 0, d_dn = 0;
  std::vector<uint32_t> s_up, s_dn;  // strings by rank
  // device copies
  DevBuf d_s_up, d_s_dn, d_binom;
  int device = -1;

  int64_t dim() const { return d_up * d_dn; }
  // colex rank of a bitstring with the species' popcount
  static int64_t colex_rank(uint32_t bits);
  // reference index (DFS position, basis.cpp:267-302) of (up, dn)
  int64_t ref_index(uint32_t up, uint32_t dn) const;
  // inverse of ref_index
  void ref_state(int64_t index, uint32_t *up, uint32_t *dn) const;
  void ensure_device();
};

// ---------------------------------------------------------------- model
// Per-species hop table: CSR rows by string rank; each entry the target rank
// and the signed amplitude (coefficient * fermionic sign).
struct SpeciesHops {
  std::vector<int32_t> row_ptr, tgt;
  std::vector<double> amp;
  std::vector<double> onsite;  // one-body diagonal of self bonds (all zero unless a size is 1)
  int max_row = 0;
  DevBuf d_row_ptr, d_tgt, d_amp, d_ampf, d_onsite, d_onsitef;
};

struct Hubbard {
  const Sector *sec = nullptr;
  double t = 0, U = 0;
  std::vector<int> bond_from, bond_to;
  SpeciesHops up, dn;
  bool has_onsite = false;
  std::vector<double> diagU;  // interaction energy of k double occupations
  DevBuf d_diagU, d_diagUf;
  int device = -1;
  int64_t dim() const { return sec->dim(); }
};

void build_hubbard(Hubbard &h);

// ------------------------------------------------------------ kernels
// Views passed to kernels (plain pointers).
struct HvTables {
  int64_t d_up, d_dn;
  const uint32_t *s_up, *s_dn;
  const int32_t *up_ptr, *up_tgt, *dn_ptr, *dn_tgt;
  const double *up_amp, *dn_amp, *up_on, *dn_on, *diagU;
  const float *up_ampf, *dn_ampf, *up_onf, *dn_onf, *diagUf;
  int has_onsite;
  int max_up_row;
};
HvTables tables_of(const Hubbard &h);

struct PermTables {
  int L, n_up, n_dn;
  int64_t d_up, d_dn;
  const uint32_t *s_up, *s_dn;
  const uint64_t *binom;  // (kMaxSites+1)^2
};
PermTables perm_of(const Sector &s);

// Number of blocks used by the deterministic reductions.
constexpr int kRedBlocks = 512;

void launch_hv_c128(const HvTables &t, const void *x, void *y, hipStream_t s);
void launch_hv_c64(const HvTables &t, const void *x, void *y, hipStream_t s);
void launch_hv_c64_multi(const HvTables &t, int nvec, const void *x, void *y, hipStream_t s);

void launch_gather_c128(const PermTables &p, const void *x_ref, void *x_dev, hipStream_t s);
void launch_scatter_c128(const PermTables &p, const void *x_dev, void *x_ref, hipStream_t s);
void launch_fill_gaussian_c128(const PermTables &p, uint64_t seed, void *x, hipStream_t s);
void launch_fill_phase_c64_multi(const PermTables &p, uint64_t seed, int nvec, void *x,
                                 hipStream_t s);
void launch_fill_phase_c128_multi(const PermTables &p, uint64_t seed, int nvec, void *x,
                                  hipStream_t s);

// Lanczos device-resident scalars for one run of k steps.
struct LanczosState {
  double *alpha;     // [k]
  double *beta;      // [k]
  double *partial;   // [kRedBlocks]
  int *flags;        // [0] = done, [1] = steps taken
  double *bnorm;     // [1]
};
void launch_norm_c128(const double *x, int64_t n, double *partial, double *out, hipStream_t s);
void launch_scale_div_c128(const double *x, const double *den, double *y, int64_t n,
                           hipStream_t s);
// w = H v - beta_prev v_prev; alpha partials; guarded by flags[0]
void launch_lanczos_hv(const HvTables &t, const double *v, const double *vprev, double *w,
                       const LanczosState &st, int step, hipStream_t s);
int lanczos_partial_blocks(int64_t D);
void launch_lanczos_finish_alpha(const LanczosState &st, int step, int nblocks, hipStream_t s);
void launch_lanczos_axpy_norm(const double *v, double *w, int64_t n, const LanczosState &st,
                              int step, hipStream_t s);
void launch_lanczos_finish_beta(const LanczosState &st, int step, int k, hipStream_t s);
void launch_lanczos_normalize(double *w, int64_t n, const LanczosState &st, int step,
                              hipStream_t s);
// pass two: w = H v - beta_prev v_prev - alpha v; w /= beta; x += y_j w
void launch_recon_hv(const HvTables &t, const double *v, const double *vprev, double *w,
                     double alpha, double beta_prev, hipStream_t s);
void launch_recon_update(double *w, double *x, double beta, double yr, double yi, int64_t n,
                         hipStream_t s);
void launch_scale_complex_c128(const double *x, double yr, double yi, double *out, int64_t n,
                               hipStream_t s);

// KPM: element-major nvec complex vectors.
void launch_cheb_c64(const HvTables &t, int nvec, const float *tn, const float *tprev, float *tnext,
                     float scale2, float shift, int first, double *partial, hipStream_t s);
void launch_cheb_c128(const HvTables &t, int nvec, const double *tn, const double *tprev,
                      double *tnext, double scale2, double shift, int first, double *partial,
                      hipStream_t s);
void launch_reduce_partials(const double *partial, int nblocks, int width, double *out,
                            hipStream_t s);
int cheb_partial_blocks();

// host numerics
void tridiag_eig(int m, const double *a, const double *b, double *w, double *V /*nullable*/);

}  // namespace mb
