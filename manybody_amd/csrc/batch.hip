// Batched Lanczos: many independent Hubbard instances of uneven dimension
// advanced in lockstep, one launch per phase of a step for the whole batch.
//
// Work mapping: every instance's elements (u, d) are cut into tiles of `tile`
// elements; one workgroup processes one tile of one instance, both planes
// (re, im) at once so the index arithmetic and table reads are shared.  A tile
// table (instance, first element, count) lets one launch cover all instances;
// the tile size and workgroup size are chosen by the autotuner.  Per-instance
// reductions are two-level and atomic-free: per-tile partial sums, then one
// wave per instance adds its tiles in tile order.
//
// The per-step sequence (hv + alpha, axpy + beta^2, beta + breakdown,
// normalise) is the recurrence of lanczos.h:15-40 for every instance; an
// instance whose beta falls to 1000 eps is marked done and skipped from then
// on (lanczos.h:35-37, 165-167).  The whole k-step sequence can be captured in
// one hipGraph: no memset nodes are recorded (initialisation uses kernels).
#include "batch.hpp"
#include "batch_dev.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>

namespace mb {

namespace {

using namespace bdev;

// v = sigma X (sigma = 1 / |b| at step 0, 1 / beta[step-1] after: the previous
// step's normalisation, applied as X is read instead of in a pass of its own);
// w = H v - beta[step-1] v_prev ; partial[tile] = sum v . w.  The normalised
// vectors are not stored: v_prev is read as sigma_prev X_prev (VP = the
// previous step's X), and V, when not null, receives v.
template <int BS>
__global__ __launch_bounds__(BS) void bl_hv_kernel(const BInst *__restrict__ inst,
                                                   const BTile *__restrict__ tiles,
                                                   const double *__restrict__ X,
                                                   double *__restrict__ V,
                                                   const double *__restrict__ VP,
                                                   double *__restrict__ Wv, const double *beta,
                                                   const double *bnorm, const int *done, int k,
                                                   int step, double *partial) {
  __shared__ double lds[BS / kWave];
  extern __shared__ double lds_row[];
  const BTile tl = tiles[blockIdx.x];
  if (done[tl.inst]) {
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const BInst I = inst[tl.inst];
  const double b1 = step > 0 ? beta[(int64_t)tl.inst * k + step - 1] : 0.0;
  const double sigma = 1.0 / (step > 0 ? b1 : bnorm[tl.inst]);
  // beta[step-1] v_prev = beta[step-1] sigma_prev X_prev
  const double bp = step > 0 ? b1 / (step > 1 ? beta[(int64_t)tl.inst * k + step - 2] : bnorm[tl.inst])
                             : 0.0;
  const double *x = X + I.off, *vp = VP + I.off;
  double *v = V ? V + I.off : nullptr, *w = Wv + I.off;
  const int64_t D = I.D;
  double s = 0.0;
  if (I.rows) {
    s = hv_rows_tile<BS>(I, tl, x, sigma, v, vp, w, bp, lds_row);
    s = block_sum<BS>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    return;
  }
  for (int64_t e = tl.e0 + threadIdx.x; e < tl.e0 + tl.n; e += BS) {
    double re, im;
    hv_pair(I.t, x, D, e, re, im);
    re *= sigma;
    im *= sigma;
    const double vr = sigma * x[e], vi = sigma * x[D + e];
    if (v) {
      v[e] = vr;
      v[D + e] = vi;
    }
    if (bp != 0.0) {
      re -= bp * vp[e];
      im -= bp * vp[D + e];
    }
    w[e] = re;
    w[D + e] = im;
    s += vr * re + vi * im;
  }
  s = block_sum<BS>(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// mode 0: w -= alpha v with v = sigma V (sigma as in bl_hv_kernel: V holds the
// step's unnormalised input), partial = |w|^2 ; mode 1: w /= beta ;
// mode 2: partial = |w|^2 ; mode 3: v = w / bnorm (start vector), vp = 0
template <int BS>
__global__ __launch_bounds__(BS) void bl_vec_kernel(const BInst *__restrict__ inst,
                                                    const BTile *__restrict__ tiles,
                                                    double *__restrict__ V, double *__restrict__ VP,
                                                    double *__restrict__ Wv, const double *alpha,
                                                    const double *beta, const double *bnorm,
                                                    const int *done, int k, int step, int mode,
                                                    double *partial) {
  __shared__ double lds[BS / kWave];
  const BTile tl = tiles[blockIdx.x];
  if (done[tl.inst] && mode != 3) {
    if (threadIdx.x == 0 && partial) partial[blockIdx.x] = 0.0;
    return;
  }
  const BInst I = inst[tl.inst];
  const int64_t D = I.D;
  double *v = V + I.off, *vp = VP + I.off, *w = Wv + I.off;
  double s = 0.0;
  const double a =
      mode == 0 ? alpha[(int64_t)tl.inst * k + step] /
                      (step > 0 ? beta[(int64_t)tl.inst * k + step - 1] : bnorm[tl.inst])
                : 0.0;
  const double b = mode == 1 ? beta[(int64_t)tl.inst * k + step] : (mode == 3 ? bnorm[tl.inst] : 1.0);
  for (int64_t e = tl.e0 + threadIdx.x; e < tl.e0 + tl.n; e += BS) {
    for (int p = 0; p < 2; ++p) {
      const int64_t i = p * D + e;
      if (mode == 0) {
        const double q = w[i] - a * v[i];
        w[i] = q;
        s += q * q;
      } else if (mode == 1) {
        w[i] = w[i] / b;
      } else if (mode == 2) {
        s += w[i] * w[i];
      } else {
        v[i] = w[i] / b;
        vp[i] = 0.0;
      }
    }
  }
  if (mode == 0 || mode == 2) {
    s = block_sum<BS>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// One wave per instance: sum of its tiles' partials in tile order.
// what 0: alpha[i][step] = sum, steps[i] = step + 1
// what 1: beta[i][step] = sqrt(sum); breakdown -> done[i] = 1
// what 2: bnorm[i] = sqrt(sum); zero-norm start -> done[i] = 2
// what 3: reset done and steps
__global__ __launch_bounds__(kWave) void bl_reduce_kernel(const BInst *__restrict__ inst, int n,
                                                          const double *partial, double *alpha,
                                                          double *beta, double *bnorm, int *done,
                                                          int *steps, int k, int step, int what) {
  const int i = blockIdx.x;
  if (i >= n) return;
  if (what == 3) {
    if (threadIdx.x == 0) {
      done[i] = 0;
      steps[i] = 0;
    }
    return;
  }
  if (inst[i].small) return;  // the persistent kernel keeps its own scalars
  if (done[i] && what != 2) return;
  const BInst I = inst[i];
  double s = 0.0;
  for (int j = I.tile0 + threadIdx.x; j < I.tile1; j += kWave) s += partial[j];
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kWave);
  if (threadIdx.x == 0) {
    if (what == 0) {
      alpha[(int64_t)i * k + step] = s;
      steps[i] = step + 1;
    } else if (what == 1) {
      const double b = sqrt(s);
      beta[(int64_t)i * k + step] = b;
      if (b <= kTol) done[i] = 1;
    } else {
      const double b = sqrt(s);
      bnorm[i] = b;
      if (b <= kTol) done[i] = 2;
    }
  }
}

// Fixed-order sum over the workgroup, the result in every thread.
template <int BS>
__device__ __forceinline__ double block_allsum(double v, double *lds) {
  const int tid = threadIdx.x;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) lds[tid / kWave] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < BS / kWave; ++w) r += lds[w];
  __syncthreads();
  return r;
}

// One workgroup per small instance runs all k steps of lanczos_pass_one
// (lanczos.h:130-175, recurrence :15-40) without leaving the kernel: |b|, v =
// b / |b|, then per step w = H v - beta v_prev, alpha = Re <v|w>, w -= alpha v,
// beta = |w|, stop at beta <= 1000 eps (the step is counted, its beta kept but
// not used), else v_next = w / beta.  The three vectors rotate through V, VP
// and W; every reduction is a fixed-order workgroup sum, so a replay gives the
// same bits.  Writes alpha / beta / steps / bnorm / done as the per-phase
// kernels do.
template <int BS>
__global__ __launch_bounds__(BS) void bl_persistent_kernel(const BInst *__restrict__ inst,
                                                           const int *__restrict__ small_ids,
                                                           const double *__restrict__ B,
                                                           double *V, double *VP, double *Wv,
                                                           double *alpha, double *beta,
                                                           double *bnorm, int *done, int *steps,
                                                           int k, int kcap) {
  __shared__ double lds[BS / kWave];
  const int i = small_ids[blockIdx.x];
  const BInst I = inst[i];
  const int64_t D = I.D, n2 = 2 * D;
  const int tid = threadIdx.x;
  const double *b = B + I.off;
  double *v = V + I.off, *vp = VP + I.off, *w = Wv + I.off;
  double s = 0.0;
  for (int64_t e = tid; e < n2; e += BS) s += b[e] * b[e];
  const double bn = sqrt(block_allsum<BS>(s, lds));
  if (tid == 0) {
    bnorm[i] = bn;
    steps[i] = 0;
    done[i] = bn <= kTol ? 2 : 0;
  }
  if (bn <= kTol) return;
  for (int64_t e = tid; e < n2; e += BS) {
    v[e] = b[e] / bn;
    vp[e] = 0.0;
  }
  __syncthreads();
  double bprev = 0.0;
  for (int j = 0; j < k; ++j) {
    s = 0.0;
    for (int64_t e = tid; e < D; e += BS) {
      double re, im;
      hv_pair(I.t, v, D, e, re, im);
      if (bprev != 0.0) {
        re -= bprev * vp[e];
        im -= bprev * vp[D + e];
      }
      w[e] = re;
      w[D + e] = im;
      s += v[e] * re + v[D + e] * im;
    }
    const double a = block_allsum<BS>(s, lds);
    s = 0.0;
    for (int64_t e = tid; e < n2; e += BS) {
      const double q = w[e] - a * v[e];
      w[e] = q;
      s += q * q;
    }
    const double bt = sqrt(block_allsum<BS>(s, lds));
    if (tid == 0) {
      alpha[(int64_t)i * kcap + j] = a;
      beta[(int64_t)i * kcap + j] = bt;
      steps[i] = j + 1;
      if (bt <= kTol) done[i] = 1;
    }
    if (bt <= kTol) return;
    for (int64_t e = tid; e < n2; e += BS) w[e] /= bt;
    double *t = vp;
    vp = v;
    v = w;
    w = t;
    bprev = bt;
    __syncthreads();
  }
}

// The Lanczos run of one Krylov time step for one small sector, all k steps in
// one workgroup (the recurrence of bl_persistent_kernel), with the Krylov
// vectors kept: v_j is written to basis + 2 D j (normalised), the scalars to
// scal = alpha[k], beta[k], |b|, steps.  The caller forms exp(-i tau T) e1 and
// combines the basis.  A zero b leaves steps = 0 and |b| in scal.
template <int BS>
__global__ __launch_bounds__(BS) void krylov_small_kernel(HvTables t, int64_t D,
                                                          const double *__restrict__ b,
                                                          double *__restrict__ basis,
                                                          double *__restrict__ scal, int k) {
  __shared__ double lds[BS / kWave];
  const int64_t n2 = 2 * D;
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int64_t e = tid; e < n2; e += BS) s += b[e] * b[e];
  const double bn = sqrt(block_allsum<BS>(s, lds));
  if (tid == 0) {
    scal[2 * k] = bn;
    scal[2 * k + 1] = 0.0;
  }
  if (bn <= kTol) return;
  for (int64_t e = tid; e < n2; e += BS) basis[e] = b[e] / bn;
  __syncthreads();
  double bprev = 0.0;
  for (int j = 0; j < k; ++j) {
    const double *v = basis + (int64_t)j * n2;
    const double *vp = j > 0 ? v - n2 : v;  // read only for j > 0
    double *w = basis + (int64_t)(j + 1) * n2;
    s = 0.0;
    for (int64_t e = tid; e < D; e += BS) {
      double re, im;
      hv_pair(t, v, D, e, re, im);
      if (j > 0) {
        re -= bprev * vp[e];
        im -= bprev * vp[D + e];
      }
      w[e] = re;
      w[D + e] = im;
      s += v[e] * re + v[D + e] * im;
    }
    const double a = block_allsum<BS>(s, lds);
    s = 0.0;
    for (int64_t e = tid; e < n2; e += BS) {
      const double q = w[e] - a * v[e];
      w[e] = q;
      s += q * q;
    }
    const double bt = sqrt(block_allsum<BS>(s, lds));
    if (tid == 0) {
      scal[j] = a;
      scal[k + j] = bt;
      scal[2 * k + 1] = (double)(j + 1);
    }
    if (bt <= kTol) return;
    for (int64_t e = tid; e < n2; e += BS) w[e] /= bt;
    bprev = bt;
    __syncthreads();
  }
}

}  // namespace

void launch_krylov_small(const HvTables &t, int64_t D, const double *b, double *basis, double *scal,
                         int k, hipStream_t s) {
  if (D > kKrylovSmallMaxDim) throw std::invalid_argument("sector too large for the one-workgroup Krylov run");
  hipLaunchKernelGGL((krylov_small_kernel<256>), dim3(1), dim3(256), 0, s, t, D, b, basis, scal, k);
  MB_HIP(hipGetLastError());
}

// ------------------------------------------------------------ host engine

Batch::Batch(std::vector<const Hubbard *> models) : models_(std::move(models)) {
  int64_t off = 0;
  for (const Hubbard *h : models_) {
    BInst I{};
    I.t = tables_of(*h);
    I.D = h->dim();
    I.off = off;
    off += 2 * I.D;
    host_inst_.push_back(I);
  }
  total_ = off;
  const size_t bytes = (size_t)std::max<int64_t>(total_, 1) * sizeof(double);
  V_.alloc(bytes);
  VP_.alloc(bytes);
  W_.alloc(bytes);
  B_.alloc(bytes);
  d_inst_.alloc(sizeof(BInst) * std::max<size_t>(host_inst_.size(), 1));
  done_.alloc(sizeof(int) * std::max<size_t>(host_inst_.size(), 1));
  steps_.alloc(sizeof(int) * std::max<size_t>(host_inst_.size(), 1));
  bnorm_.alloc(sizeof(double) * std::max<size_t>(host_inst_.size(), 1));
  set_mapping(1024, 256);
}

int64_t Batch::footprint_bytes() const {
  return 4 * total_ * (int64_t)sizeof(double);
}

void Batch::set_mapping(int tile, int block) {
  if (block != 128 && block != 256 && block != 512) throw std::invalid_argument("block must be 128, 256 or 512");
  if (tile < block || tile % block) throw std::invalid_argument("tile must be a multiple of block");
  tile_ = tile;
  block_ = block;
  tiles_.clear();
  lds_bytes_ = 0;
  for (size_t i = 0; i < host_inst_.size(); ++i) {
    BInst &I = host_inst_[i];
    I.tile0 = (int)tiles_.size();
    I.small = I.D <= persist_max_dim_ ? 1 : 0;
    if (I.small) {
      I.tile1 = I.tile0;
      I.rows = 0;
      continue;
    }
    const int64_t Ddn = I.t.d_dn;
    // row tiles when the compact down table exists and a row is short
    const int64_t rows_n = std::max<int64_t>(1, tile / Ddn) * Ddn;
    I.rows = (I.t.comp_ok && Ddn <= kMaxRowsTileDn &&
              rows_tile_lds(Ddn, rows_n, I.t.max_up_row) <= kRowsTileLdsMax) ? 1 : 0;
    const int64_t step = I.rows ? rows_n : tile;
    for (int64_t e = 0; e < I.D; e += step) {
      BTile tl;
      tl.inst = (int)i;
      tl.e0 = e;
      tl.n = (int)std::min<int64_t>(step, I.D - e);
      tiles_.push_back(tl);
      if (I.rows) lds_bytes_ = std::max<size_t>(lds_bytes_, rows_tile_lds(Ddn, tl.n, I.t.max_up_row));
    }
    I.tile1 = (int)tiles_.size();
  }
  if (tiles_.empty()) {  // every instance is small: one idle tile keeps the launches valid
    BTile tl;
    tl.inst = 0;
    tl.e0 = 0;
    tl.n = 0;
    tiles_.push_back(tl);
  }
  std::vector<int> small;
  for (size_t i = 0; i < host_inst_.size(); ++i)
    if (host_inst_[i].small) small.push_back((int)i);
  n_small_ = (int)small.size();
  d_small_ = upload(small.empty() ? std::vector<int>{0} : small);
  d_tiles_ = upload(tiles_);
  MB_HIP(hipMemcpy(d_inst_.p, host_inst_.data(), sizeof(BInst) * host_inst_.size(),
                   hipMemcpyHostToDevice));
  partial_.alloc(sizeof(double) * std::max<size_t>(tiles_.size(), 1));
  graph_k_ = -1;  // invalidate a captured graph
  if (graph_exec_) {
    (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
}

Batch::~Batch() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (side_) {
    (void)hipStreamSynchronize(side_);
    (void)hipEventDestroy(fork_);
    (void)hipEventDestroy(join_);
    (void)hipStreamDestroy(side_);
  }
}

template <int BS>
static void launch_hv_bs(const Batch &b, const double *X, double *V, const double *VP, double *W,
                         int k, int step, hipStream_t s) {
  static std::atomic<bool> attr{false};  // idempotent; host threads may race
  if (!attr) {
    MB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&bl_hv_kernel<BS>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL((bl_hv_kernel<BS>), dim3((unsigned)b.tiles_.size()), dim3(BS), b.lds_bytes_, s,
                     b.d_inst_.as<BInst>(), b.d_tiles_.as<BTile>(), X, V, VP, W,
                     b.beta_.as<double>(), b.bnorm_.as<double>(), b.done_.as<int>(), k, step,
                     b.partial_.as<double>());
}

template <int BS>
static void launch_vec_bs(const Batch &b, double *V, double *VP, double *W, int k, int step,
                          int mode, hipStream_t s) {
  hipLaunchKernelGGL((bl_vec_kernel<BS>), dim3((unsigned)b.tiles_.size()), dim3(BS), 0, s,
                     b.d_inst_.as<BInst>(), b.d_tiles_.as<BTile>(), V, VP, W,
                     b.alpha_.as<double>(), b.beta_.as<double>(), b.bnorm_.as<double>(),
                     b.done_.as<int>(), k, step, mode, b.partial_.as<double>());
}

void Batch::hv_step(const double *X, double *V, const double *VP, double *W, int k, int step,
                    hipStream_t s) const {
  switch (block_) {
    case 128: launch_hv_bs<128>(*this, X, V, VP, W, k, step, s); break;
    case 256: launch_hv_bs<256>(*this, X, V, VP, W, k, step, s); break;
    default: launch_hv_bs<512>(*this, X, V, VP, W, k, step, s); break;
  }
}

void Batch::vec_step(double *V, double *VP, double *W, int k, int step, int mode,
                     hipStream_t s) const {
  switch (block_) {
    case 128: launch_vec_bs<128>(*this, V, VP, W, k, step, mode, s); break;
    case 256: launch_vec_bs<256>(*this, V, VP, W, k, step, mode, s); break;
    default: launch_vec_bs<512>(*this, V, VP, W, k, step, mode, s); break;
  }
}

void Batch::reduce(int k, int step, int what, hipStream_t s) const {
  const int n = (int)host_inst_.size();
  hipLaunchKernelGGL(bl_reduce_kernel, dim3((unsigned)std::max(n, 1)), dim3(kWave), 0, s,
                     d_inst_.as<BInst>(), n, partial_.as<double>(), alpha_.as<double>(),
                     beta_.as<double>(), bnorm_.as<double>(), done_.as<int>(), steps_.as<int>(), k,
                     step, what);
}

void Batch::ensure_k(int k) {
  if (k > kcap_) {
    const size_t n = std::max<size_t>(host_inst_.size(), 1);
    alpha_.alloc(sizeof(double) * n * k);
    beta_.alloc(sizeof(double) * n * k);
    kcap_ = k;
    graph_k_ = -1;
    if (graph_exec_) {
      (void)hipGraphExecDestroy(graph_exec_);
      graph_exec_ = nullptr;
    }
  }
}

void Batch::persistent_run(int k, hipStream_t s) const {
  if (n_small_ == 0) return;
  constexpr int BS = 256;
  hipLaunchKernelGGL((bl_persistent_kernel<BS>), dim3((unsigned)n_small_), dim3(BS), 0, s,
                     d_inst_.as<BInst>(), d_small_.as<int>(), B_.as<double>(), V_.as<double>(),
                     VP_.as<double>(), W_.as<double>(), alpha_.as<double>(), beta_.as<double>(),
                     bnorm_.as<double>(), done_.as<int>(), steps_.as<int>(), k, kcap_);
  MB_HIP(hipGetLastError());
}

void Batch::set_persist_max_dim(int64_t max_dim) {
  persist_max_dim_ = max_dim;
  set_mapping(tile_, block_);
}

// Start vectors (in B_) -> normalised V, zero VP, reset flags; then k steps.
// The small instances run in the persistent kernel on a second stream forked
// after the reset (captured as a parallel branch of the graph), beside the
// per-phase kernels of the tiled instances.
void Batch::record_run(int k, hipStream_t s) {
  const int n = (int)host_inst_.size();
  (void)n;
  reduce(kcap_, 0, 3, s);  // reset done / steps
  if (n_small_ > 0) {
    if (!side_) {
      MB_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
      MB_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
      MB_HIP(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
    }
    MB_HIP(hipEventRecord(fork_, s));
    MB_HIP(hipStreamWaitEvent(side_, fork_, 0));
    persistent_run(k, side_);
    MB_HIP(hipEventRecord(join_, side_));
  }
  vec_step(nullptr, nullptr, B_.as<double>(), kcap_, 0, 2, s);  // |b|^2 partials
  reduce(kcap_, 0, 2, s);                                      // bnorm, zero-norm flag
  // x: the unnormalised input of the step (b, then the previous w), read as
  // v = x / (|b| or beta); the previous step's x stands for v_prev.  The
  // normalised vectors are never written (two vector passes per step fewer).
  // B_ is never overwritten, so a replay starts from the same vectors; the
  // steps' outputs rotate through three buffers.
  double *rot[3] = {W_.as<double>(), V_.as<double>(), VP_.as<double>()};
  double *x = B_.as<double>(), *xp = B_.as<double>();  // xp unread at step 0
  for (int i = 0; i < k; ++i) {
    double *w = rot[i % 3];
    hv_step(x, nullptr, xp, w, kcap_, i, s);
    reduce(kcap_, i, 0, s);                   // alpha
    vec_step(x, nullptr, w, kcap_, i, 0, s);  // w -= alpha v, |w|^2
    reduce(kcap_, i, 1, s);                   // beta, breakdown
    xp = x;
    x = w;
  }
  if (n_small_ > 0) MB_HIP(hipStreamWaitEvent(s, join_, 0));
  MB_HIP(hipGetLastError());
}

void Batch::run(int k, bool use_graph, hipStream_t s) {
  if (k <= 0) throw std::invalid_argument("k must be positive");
  ensure_k(k);
  if (!use_graph) {
    record_run(k, s);
    return;
  }
  if (graph_k_ != k || !graph_exec_) {
    if (graph_exec_) {
      (void)hipGraphExecDestroy(graph_exec_);
      graph_exec_ = nullptr;
    }
    hipStream_t cs;
    MB_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    MB_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    record_run(k, cs);
    MB_HIP(hipStreamEndCapture(cs, &g));
    MB_HIP(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    graph_k_ = k;
  }
  MB_HIP(hipGraphLaunch(graph_exec_, s));
}

void Batch::results(int k, std::vector<double> &alphas, std::vector<double> &betas,
                    std::vector<int> &steps, std::vector<double> &bnorms, hipStream_t s) const {
  const size_t n = host_inst_.size();
  std::vector<double> a(n * kcap_), b(n * kcap_);
  steps.assign(n, 0);
  bnorms.assign(n, 0.0);
  std::vector<int> done(n);
  MB_HIP(hipMemcpyAsync(a.data(), alpha_.p, sizeof(double) * n * kcap_, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(b.data(), beta_.p, sizeof(double) * n * kcap_, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(steps.data(), steps_.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(bnorms.data(), bnorm_.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MB_HIP(hipStreamSynchronize(s));
  alphas.assign(n * (size_t)k, 0.0);
  betas.assign(n * (size_t)k, 0.0);
  for (size_t i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) {
      alphas[i * k + j] = a[i * kcap_ + j];
      betas[i * k + j] = (j + 1 < steps[i]) ? b[i * kcap_ + j] : 0.0;
    }
}

std::vector<double> Batch::ground_energies(int k, bool use_graph, hipStream_t s) {
  run(k, use_graph, s);
  std::vector<double> al, be, bn;
  std::vector<int> st;
  results(k, al, be, st, bn, s);
  const size_t n = host_inst_.size();
  for (size_t i = 0; i < n; ++i)
    if (bn[i] <= kTol) throw std::runtime_error("Input vector has zero norm");
  // the tridiagonal solves are independent: split them over host threads
  std::vector<double> e0(n, 0.0);
  const size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  auto work = [&](size_t t0) {
    for (size_t i = t0; i < n; i += nt)
      if (st[i] > 0) e0[i] = tridiag_min_eig(st[i], &al[i * k], &be[i * k]);
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt && t < n; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &th : pool) th.join();
  return e0;
}

void Batch::set_start(size_t i, const void *b_dev, hipStream_t s) {
  if (i >= host_inst_.size()) throw RangeError("instance index out of range");
  const BInst &I = host_inst_[i];
  MB_HIP(hipMemcpyAsync(B_.as<double>() + I.off, b_dev, sizeof(double) * 2 * I.D,
                        hipMemcpyDeviceToDevice, s));
}

void Batch::fill_gaussian(const uint64_t *seeds, hipStream_t s) {
  for (size_t i = 0; i < host_inst_.size(); ++i)
    launch_fill_gaussian_c128(perm_of(*models_[i]->sec), seeds[i], B_.as<double>() + host_inst_[i].off,
                              s);
}

double Batch::time_steps(int k, hipStream_t s) {
  hipEvent_t a, b;
  MB_HIP(hipEventCreate(&a));
  MB_HIP(hipEventCreate(&b));
  ensure_k(k);
  record_run(1, s);  // warm
  MB_HIP(hipEventRecord(a, s));
  record_run(k, s);
  MB_HIP(hipEventRecord(b, s));
  MB_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  MB_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

std::pair<int, int> Batch::autotune(int k_probe, hipStream_t s) {
  const int tiles[] = {256, 512, 1024, 2048, 4096};
  const int blocks[] = {128, 256, 512};
  double best = 1e300;
  std::pair<int, int> arg{tile_, block_};
  for (int tl : tiles)
    for (int bs : blocks) {
      if (tl < bs || tl % bs) continue;
      set_mapping(tl, bs);
      const double ms = time_steps(k_probe, s);
      if (ms < best) {
        best = ms;
        arg = {tl, bs};
      }
    }
  set_mapping(arg.first, arg.second);
  return arg;
}

}  // namespace mb
