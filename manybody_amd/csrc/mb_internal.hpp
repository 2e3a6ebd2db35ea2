// Internal declarations shared by the host code and the HIP kernels of
// libmanybody_hip.so.  Not part of the public C ABI (include/manybody_hip.h).
//
// Device vector layout ("device order"): P real planes of D = D_up * D_dn
// values each, plane p at offset p * D, element (u, d) at u * D_dn + d, where u
// and d are the combinatorial (colex) ranks of the up and down bitstrings.  A
// complex vector is P = 2 planes (re, im); nvec complex vectors are P = 2 nvec
// planes, vector r = planes 2r, 2r+1.  H is real, so every plane is transformed
// independently by the same tables.
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

// Source of device memory other than hipMalloc (the server's per-request
// regions).  A thread that sets thread_dev_alloc() gets every DevBuf it
// allocates from that source until it resets it; the buffer remembers its
// owner and returns the memory there.
struct DevAlloc {
  virtual void *alloc(size_t n) = 0;  // nullptr when the source is exhausted
  virtual void free(void *p) = 0;
  virtual ~DevAlloc() = default;
};
DevAlloc *&thread_dev_alloc();
// Stream used by upload() on this thread (null: the null stream).
hipStream_t &thread_upload_stream();

// Owning device buffer (hipMalloc / hipFree, or the thread's DevAlloc).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevAlloc *owner = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), owner(o.owner) {
    o.p = nullptr;
    o.bytes = 0;
    o.owner = nullptr;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      owner = o.owner;
      o.p = nullptr;
      o.bytes = 0;
      o.owner = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Host vector -> new device buffer, copied on the thread's upload stream and
// complete on return.
template <typename T>
DevBuf upload(const std::vector<T> &v) {
  DevBuf b(v.size() * sizeof(T) + 16);
  if (!v.empty()) {
    hipStream_t s = thread_upload_stream();
    MB_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    MB_HIP(hipStreamSynchronize(s));
  }
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
// Fills the species strings of the (L, n_up, n_dn) sector (host only).
void build_sector(Sector &s, int L, int n_up, int n_dn);

// ---------------------------------------------------------------- model
// Per-species hop table: CSR rows by string rank; each entry the target rank
// and the signed amplitude (coefficient * fermionic sign).  The down species
// also gets a column-major packed copy (ELL): entry k of string d at
// ell[k * D + d] = target << 8 | sign << 7 | magnitude index.
struct SpeciesHops {
  std::vector<int32_t> row_ptr, tgt;
  std::vector<double> amp;
  std::vector<double> onsite;  // one-body diagonal of self bonds (all zero unless a size is 1)
  int max_row = 0;
  DevBuf d_row_ptr, d_tgt, d_amp, d_ampf, d_onsite, d_onsitef;
  // the CSR entries packed in 16 bits, target | (amplitude < 0) << 15, when
  // every amplitude has the magnitude mag16 and targets are below 2^15
  bool c16_ok = false;
  double mag16 = 0.0;
  std::vector<uint16_t> c16;
  DevBuf d_c16;
  // packed column-major copy
  bool ell_ok = false;
  int ell_k = 0;
  std::vector<uint32_t> ell;
  std::vector<uint8_t> cnt;
  std::vector<double> mags;
  DevBuf d_ell, d_cnt, d_mags, d_magsf;
  // compact copy (single magnitude): padded target indices (per group of 64
  // strings [chunk of 8 entries][string][8]), sign masks, and the largest hop
  // count per group of 64 strings
  bool comp_ok = false;
  std::vector<uint16_t> comp;
  std::vector<uint32_t> csign;
  std::vector<uint8_t> cgmax;
  DevBuf d_comp, d_csign, d_cgmax;
  // the compact table again in sorted slot order, with one sign pattern per
  // wave: strings sorted by (hop count, positive hop count) share a wave of 64
  // slots; per group P = most positive and N = most negative hops, and every
  // slot lists its positive targets in [0, P) and negative ones in [P, P + N),
  // padded with the zero slot n.  slot -> string, targets (laid out as the
  // compact copy, slots in place of strings), per-group P | N << 8; padded to
  // rounds of kSortedRound slots, pad slots -> string n
  bool srt_ok = false;
  int64_t srt_np = 0;
  std::vector<uint16_t> sperm, scomp, spn;
  DevBuf d_sperm, d_scomp, d_spn;
  // blocks of strings (up species): strings that agree from bit blk_bits
  // upward are a contiguous run of rows; every row's entries list the hops
  // that leave its block first and the ones inside it after (each part in
  // the original order); nin[r] counts the in-block ones.  blk holds
  // (first row, rows) per block, largest first.
  bool blk_ok = false;
  int blk_bits = 0, blk_rows = 0, blk_ent = 0;
  std::vector<int32_t> blk;
  std::vector<uint8_t> nin;
  DevBuf d_blk, d_nin;
};
constexpr int kSortedRound = 1024;  // slots per round (one per thread of the row kernel)
constexpr int kSortedRounds = 12;   // rounds per row the sorted row kernel unrolls

struct Hubbard {
  const Sector *sec = nullptr;
  double t = 0, U = 0;
  std::vector<int> bond_from, bond_to;
  int ndim = 0;                 // periodic lattice the bonds came from (0: a bond list)
  int lsize[3] = {1, 1, 1};
  SpeciesHops up, dn;
  bool has_onsite = false;
  std::vector<double> diagU;  // interaction energy of k double occupations
  DevBuf d_diagU, d_diagUf;
  int device = -1;
  int64_t dim() const { return sec->dim(); }
};

void build_hubbard(Hubbard &h);
// The rest of build_hubbard once h.up and h.dn hold their CSR lists (row_ptr,
// tgt, amp: entry (tgt, amp) of string r adds amp x[tgt] to string r's
// output), on-site values and max_row: packed tables, U per doubly occupied
// site from h.U, upload.  Also used for compiled expressions of that form.
void finish_hubbard(Hubbard &h);
// Bond list of the periodic hypercubic lattice in the reference's loop order
// (HubbardModel / HubbardModel2D / HubbardModel3D).
void lattice_bonds(int ndim, const int *sizes, std::vector<int> &from, std::vector<int> &to);

// ------------------------------------------------------------ kernels
// Views passed to kernels (plain pointers).
struct HvTables {
  int64_t d_up, d_dn;
  const uint32_t *s_up, *s_dn;
  const int32_t *up_ptr, *up_tgt, *dn_ptr, *dn_tgt;
  const double *up_amp, *dn_amp, *up_on, *dn_on, *diagU;
  const float *up_ampf, *dn_ampf, *up_onf, *dn_onf, *diagUf;
  int has_onsite;
  // packed down table
  int ell_ok, ell_k, n_mags;
  const uint32_t *dn_ell;
  const uint8_t *dn_cnt;
  const double *mags;
  const float *magsf;
  // compact down table, one magnitude: 32 uint16 target indices per string
  // (unused slots = d_dn, a zero slot) stored per group of 64 strings as
  // [chunk of 8 entries][string], per-string sign masks, per-64-string
  // largest hop count; padded to a multiple of 64 strings
  int comp_ok;
  const uint16_t *dn_comp;
  const uint32_t *dn_csign;
  const uint8_t *dn_cgmax;
  // the compact table in sorted slot order (see SpeciesHops)
  int srt_ok;
  int64_t srt_np;
  const uint16_t *srt_perm, *srt_comp, *srt_pn;
  int n_diag;  // entries of diagU that can occur: min(n_up, n_dn) + 1
  int max_up_row;
  // up hops packed in 16 bits (see SpeciesHops::c16)
  int up_c16_ok;
  const uint16_t *up_c16;
  double up_mag;
  // blocks of up strings (see SpeciesHops::blk): count, largest block's rows
  // and hop entries, (first row, rows) pairs, in-block hops per row
  int ub_ok, ub_nb, ub_rows, ub_ent;
  const int32_t *ub_blk;
  const uint8_t *ub_nin;
};
// Chunk j (entries 8 j .. 8 j + 7) of string or slot d of a compact down
// table.  The tables are laid out per group of 64 strings as [chunk][string]
// (16 bytes each), so the 64 lanes of a wave read one contiguous KB per chunk.
// The one-pass, compact and multi-plane row kernels load all four chunks:
// skipping the chunks past a group's largest hop count made the FP64 one-pass
// Chebyshev step on the L = 14 ring 4.95 ms against 4.09 ms (the sorted row
// kernel and the batched engine keep the skip).
__device__ __forceinline__ uint4 comp_chunk(const uint16_t *tab, int64_t d, int j) {
  return reinterpret_cast<const uint4 *>(tab)[(d & ~(int64_t)63) * 4 + j * 64 + (d & 63)];
}
// LDS a workgroup of the block-staged up kernel may use (six or more per CU)
// and the bytes of a row's staged strip.
constexpr int kUpBlockLds = 24 * 1024;
constexpr int kUpBlockRowBytes = 256;
HvTables tables_of(const Hubbard &h);

struct PermTables {
  int L, n_up, n_dn;
  int64_t d_up, d_dn;
  const uint32_t *s_up, *s_dn;
  const uint64_t *binom;  // (kMaxSites+1)^2
};
PermTables perm_of(const Sector &s);

// What the second H.v kernel does with (H x) before storing it.
enum class Epi : int {
  Plain = 0,    // y = H x
  Lanczos = 1,  // w = H v - beta[step-1] v_prev ; partial[wg] = sum v . w
  Recon = 2,    // w = H v - beta v_prev - alpha v
  Cheb = 3,     // t' = scale2 (H t - shift t) - t_prev ; partial[2 wg + {0,1}] = <t'|t>, <t'|t'>
  Down = 4,     // z = (down hops + diagonal) x, the row pass of a row-first H.v (see launch_row_down)
};

struct EpiArgs {
  const void *prev = nullptr;     // v_prev / t_prev planes (same layout as x)
  const double *beta_dev = nullptr;  // Lanczos: device beta array, uses beta_dev[step-1]
  const int *flags = nullptr;     // Lanczos: flags[0] != 0 -> kernels are no-ops
  int step = 0;
  double alpha = 0, beta = 0;     // Recon
  double scale2 = 0, shift = 0;   // Cheb
  int first = 0;                  // Cheb: no t_prev term
  double *partial = nullptr;      // per-row-workgroup partial sums
  // Lanczos: x and prev are the unnormalised vectors beta[step-1] v and
  // beta[step-2] v_prev (1 at step 1): the epilogue forms w = H x / beta[step-1]
  // - (beta[step-1] / beta[step-2]) prev and the partial sum of (x / beta[step-1]) . w
  int lazy = 0;
  // Down with a fold (the Lanczos step of the row-first H.v): the row pass
  // stages x = fold_w - c fold_v, c = alpha[fold_step] / beta[fold_step - 1]
  // (alpha[0] at step 0; the axpy of the previous Lanczos step), writes x to
  // fold_x and |x row|^2 to fold_partial[row]
  const void *fold_w = nullptr, *fold_v = nullptr;
  void *fold_x = nullptr;
  double *fold_partial = nullptr;
  const double *alpha_dev = nullptr;
  int fold = 0, fold_step = 0;
};

// y = epilogue(H x) over P planes; T = double or float.  Returns the number of
// partial sums written (one, or two for Cheb, per row workgroup).
template <typename T>
int64_t launch_hv(const HvTables &t, int P, const T *x, T *y, Epi epi, const EpiArgs &a,
                  hipStream_t s);
// Number of row workgroups (= partial slots) for P planes.
int64_t hv_partials(const HvTables &t, int P);

// h = alpha z1 + beta z2 + gamma x (complex coefficients; a null z1 / z2 or
// with_x == 0 drops that term) over planar complex vectors of D elements, then
// the epilogue (Plain, Lanczos or Recon).  Returns the partial slots per plane
// (the launch writes twice that many); combine_slots(D) of them.
struct CombineCoef {
  double ar = 0, ai = 0, br = 0, bi = 0, gr = 0, gi = 0;
  int with_x = 0;
};
int64_t combine_slots(int64_t D);
template <typename T>
int64_t launch_combine(int64_t D, const T *z1, const T *z2, const T *x, T *y, const CombineCoef &c,
                       Epi epi, const EpiArgs &a, hipStream_t s);

// Random-phase vector r of nvec (the phases of launch_fill_phase) for any
// layout, keyed by a device position -> reference index map.
void launch_fill_phase_ref(const int64_t *ref, int64_t D, uint64_t seed, int nvec, int r,
                           bool fp32, void *x, hipStream_t s);
// out0 / out1 = the sums of the even / odd entries of partial[0, 2 n)
void launch_reduce_pairs(const double *partial, int64_t n, double *out0, double *out1,
                         hipStream_t s);

// Number of blocks used by the elementwise deterministic reductions.
constexpr int kRedBlocks = 512;

void launch_gather_c128(const PermTables &p, const void *x_ref, void *x_dev, hipStream_t s);
void launch_scatter_c128(const PermTables &p, const void *x_dev, void *x_ref, hipStream_t s);
void launch_fill_gaussian_c128(const PermTables &p, uint64_t seed, void *x, hipStream_t s);
void launch_fill_phase(const PermTables &p, uint64_t seed, int nvec, bool fp32, void *x,
                       hipStream_t s);

// Lanczos device-resident scalars for one run of k steps.
struct LanczosState {
  double *alpha;     // [k]
  double *beta;      // [k]
  double *partial;   // [max(kRedBlocks, hv_partials)]
  int *flags;        // [0] = done, [1] = steps taken
  double *bnorm;     // [1]
};
void launch_norm(const double *x, int64_t n, double *partial, double *out, hipStream_t s);
// Zeroes `bytes` (a multiple of 4) at p with a fill kernel.
void launch_zero(void *p, size_t bytes, hipStream_t s);
void launch_sumsq_partial(const double *x, int64_t n, double *partial, hipStream_t s);
// w -= V (V^H w) against the m stored planar vectors of V (stride 2 D); partial
// holds 2 m kProjBlocks doubles, c 2 m doubles; no-op once flags[0] is set
constexpr int kProjBlocks = 256;
void launch_reorth(const double *V, int m, double *w, int64_t D, double *partial, double *c,
                   const int *flags, hipStream_t s);
// w -= sum_j c_j V_j only (complex c, 2 m doubles)
void launch_reorth_sub(const double *V, int m, const double *c, double *w, int64_t D,
                       const int *flags, hipStream_t s);
// out[0], out[1] = Re, Im of <a|b> for planar complex vectors of D elements;
// partial holds 2 kRedBlocks doubles
void launch_cdot(const double *a, const double *b, int64_t D, double *partial, double *out,
                 hipStream_t s);
template <typename T>
void launch_scale_div(const double *x, const double *den, T *y, int64_t n, hipStream_t s);
void launch_lanczos_finish_alpha(const LanczosState &st, int step, int64_t nparts, hipStream_t s);
// lazy: v is stored as beta[step - 1] v (see EpiArgs::lazy)
template <typename T>
void launch_lanczos_axpy_norm(const T *v, T *w, int64_t n, const LanczosState &st, int step,
                              hipStream_t s, bool lazy = false);
void launch_lanczos_finish_beta(const LanczosState &st, int step, hipStream_t s);
// the same over nparts partial sums at st.partial + kRedBlocks (summed in two
// levels; st.partial[0, kRedBlocks) is overwritten)
void launch_lanczos_finish_beta_n(const LanczosState &st, int step, int64_t nparts, hipStream_t s);
// Row-first H.v of the Lanczos step on a real-space operator (two-pass path):
// launch_row_down writes z = (down hops + diagonal) x row by row (with a.fold:
// x formed from a.fold_w and a.fold_v and written to a.fold_x); launch_up_lanczos
// then forms H x = (up hops) x + z and applies the Lanczos epilogue (a.prev,
// a.beta_dev, a.step, a.lazy, a.flags), one partial sum per workgroup into
// a.partial; it returns the number of partial sums.  lanczos_fold_ok says
// whether the operator's tables allow it; lanczos_fold_partials bounds the
// partial slots either launch writes.
bool lanczos_fold_ok(const HvTables &t);
int64_t lanczos_fold_partials(const HvTables &t);
void launch_row_down(const HvTables &t, const double *x, double *z, const EpiArgs &a, hipStream_t s);
int64_t launch_up_lanczos(const HvTables &t, const double *x, const double *z, double *w,
                          const EpiArgs &a, hipStream_t s);
template <typename T>
void launch_lanczos_normalize(T *w, int64_t n, const LanczosState &st, int step, hipStream_t s);
// pass two: w /= beta ; x += y w (complex y, planar vectors of D elements)
template <typename T>
void launch_recon_update(T *w, double *x, double beta, double yr, double yi, int64_t D,
                         hipStream_t s);
template <typename T>
void launch_scale_complex(const T *x, double yr, double yi, double *out, int64_t D, hipStream_t s);
// x = sum_{j < m} y_j V_j over Krylov vectors kept in HBM (stride 2 D), y on the device
template <typename T>
void launch_basis_combine(const T *V, int m, int64_t D, const double *y, double *x, hipStream_t s);
// out[r] = sum over row workgroups of planes 2r, 2r+1 of partial[2 wg + j] for j = 0,
// out[nvec + r] likewise for j = 1.
void launch_reduce_cheb(const double *partial, int64_t rows_per_plane, int nvec, double *out,
                        hipStream_t s);

// host numerics
void tridiag_eig(int m, const double *a, const double *b, double *w, double *V /*nullable*/);

}  // namespace mb
