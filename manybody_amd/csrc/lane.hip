// Continuous-batching Lanczos lane: independent FP64 Lanczos recurrences
// (lanczos.h:15-40, the pass-one loop of lanczos.h:130-175) of many small and
// medium Hubbard models share every launch, and models join or leave between
// steps without stopping the others.
//
// Per tick, for all members at once: (1) w = H v - beta[s-1] v_prev with
// v = x / (|b| at step 0, beta[s-1] after) and v_prev = x_prev / (|b| at step 1,
// beta[s-2] after), per-tile partial v.w, (2) alpha[s] per member, (3) w -=
// alpha v with |w|^2 partials, (4) beta[s] per member, breakdown (1000 eps) or
// step count reached -> done, next step index.  Only the unnormalised inputs
// x are kept: the normalisations are applied as the vectors are read, so there
// is no pass for them and v is never written.  The vector buffers (x, x_prev,
// w) rotate roles after each tick for every member together; each member keeps its
// own step index (double-buffered: a tick reads one copy and writes the other),
// so a member that joined late simply counts from 0.  Tiles are per member and the
// per-member sums run over its own tiles in order, so a member's numbers do not
// depend on who else is in the lane.
#include "lane.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>

#include "batch_dev.hpp"

namespace mb {

namespace {

using namespace bdev;
constexpr int kLaneBlock = 256;

__global__ __launch_bounds__(kLaneBlock) void ln_hv_kernel(
    const LaneInst *__restrict__ inst, const BTile *__restrict__ tiles, const double *__restrict__ X,
    double *__restrict__ V, const double *__restrict__ VP, double *__restrict__ Wv,
    const double *beta, const double *bnorm, const int *step, const int *done, int kmax,
    double *partial) {
  __shared__ double lds[kLaneBlock / kWave];
  extern __shared__ double lds_row[];
  const BTile tl = tiles[blockIdx.x];
  const LaneInst I = inst[tl.inst];
  if (done[I.slot]) {
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const int s = step[I.slot];
  const double b1 = s > 0 ? beta[(int64_t)I.slot * kmax + s - 1] : 0.0;
  const double sigma = 1.0 / (s > 0 ? b1 : bnorm[I.slot]);
  // beta[s-1] v_prev = beta[s-1] / (beta[s-2] or |b|) x_prev
  const double bp = s > 0 ? b1 / (s > 1 ? beta[(int64_t)I.slot * kmax + s - 2] : bnorm[I.slot]) : 0.0;
  const double *x = X + I.off, *vp = VP + I.off;
  double *v = V ? V + I.off : nullptr, *w = Wv + I.off;
  const int64_t D = I.D;
  double acc = 0.0;
  if (I.rows) {
    BInst B;
    B.t = I.t;
    B.off = I.off;
    B.D = I.D;
    B.tile0 = I.tile0;
    B.tile1 = I.tile1;
    B.rows = 1;
    acc = hv_rows_tile<kLaneBlock>(B, tl, x, sigma, v, vp, w, bp, lds_row);
  } else {
    for (int64_t e = tl.e0 + threadIdx.x; e < tl.e0 + tl.n; e += kLaneBlock) {
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
      acc += vr * re + vi * im;
    }
  }
  acc = block_sum<kLaneBlock>(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// w -= alpha v with v = x / (|b| or beta[s-1]) read from the step's input X,
// partial = |w|^2
__global__ __launch_bounds__(kLaneBlock) void ln_axpy_kernel(
    const LaneInst *__restrict__ inst, const BTile *__restrict__ tiles, const double *__restrict__ V,
    double *__restrict__ Wv, const double *alpha, const double *beta, const double *bnorm,
    const int *step, const int *done, int kmax, double *partial) {
  __shared__ double lds[kLaneBlock / kWave];
  const BTile tl = tiles[blockIdx.x];
  const LaneInst I = inst[tl.inst];
  if (done[I.slot]) {
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const int64_t D = I.D;
  const double *v = V + I.off;
  double *w = Wv + I.off;
  const int s = step[I.slot];
  const double a = alpha[(int64_t)I.slot * kmax + s] /
                   (s > 0 ? beta[(int64_t)I.slot * kmax + s - 1] : bnorm[I.slot]);
  double acc = 0.0;
  for (int64_t e = tl.e0 + threadIdx.x; e < tl.e0 + tl.n; e += kLaneBlock) {
    for (int p = 0; p < 2; ++p) {
      const int64_t i = p * D + e;
      const double q = w[i] - a * v[i];
      w[i] = q;
      acc += q * q;
    }
  }
  acc = block_sum<kLaneBlock>(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// One wave per member: its tiles' partials in tile order.
// what 0: alpha[s] ; what 1: beta[s], done on breakdown or after k steps, next step index.
__global__ __launch_bounds__(kWave) void ln_reduce_kernel(const LaneInst *__restrict__ inst,
                                                          const double *partial, double *alpha,
                                                          double *beta, int *done,
                                                          const int *step_in, int *step_out,
                                                          int kmax, int what) {
  const LaneInst I = inst[blockIdx.x];
  const int slot = I.slot;
  const int s = step_in[slot];
  if (done[slot]) {
    if (what == 1 && threadIdx.x == 0) step_out[slot] = s;
    return;
  }
  double acc = 0.0;
  for (int j = I.tile0 + threadIdx.x; j < I.tile1; j += kWave) acc += partial[j];
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
  if (threadIdx.x != 0) return;
  if (what == 0) {
    alpha[(int64_t)slot * kmax + s] = acc;
  } else {
    const double b = sqrt(acc);
    beta[(int64_t)slot * kmax + s] = b;
    if (b <= kTol || s + 1 >= I.k) done[slot] = b <= kTol ? 1 : 3;
    step_out[slot] = s + 1;
  }
}

// New member: step 0; done = 2 for a zero start vector (its first step divides by |b|).
__global__ void ln_init_kernel(const double *bnorm, int slot, int *step, int *done) {
  if (threadIdx.x == 0) {
    step[slot] = 0;
    done[slot] = bnorm[slot] > kTol ? 0 : 2;
  }
}

int64_t align_up(int64_t n, int64_t a) { return (n + a - 1) / a * a; }

}  // namespace

int64_t LaneEngine::bytes_needed(int64_t buf_doubles, int nslots, int kmax, int tile) {
  const int64_t max_tiles = buf_doubles / tile + nslots + 1;
  return 3 * align_up(buf_doubles * 8, 256) + 3 * align_up((int64_t)nslots * kmax * 8, 256) +
         align_up((int64_t)nslots * 8, 256) + align_up((max_tiles + 2 * kRedBlocks) * 8, 256) +
         3 * align_up((int64_t)nslots * 4, 256) + align_up((int64_t)nslots * sizeof(LaneInst), 256) +
         align_up(max_tiles * (int64_t)sizeof(BTile), 256) + 4096;
}

LaneEngine::LaneEngine(int64_t buf_doubles, int nslots, int kmax, int tile)
    : buf_(buf_doubles), nslots_(nslots), kmax_(kmax), tile_(tile) {
  if (buf_doubles < 1024 || nslots < 1 || kmax < 1 || tile < kLaneBlock || tile % kLaneBlock)
    throw std::invalid_argument("invalid lane configuration");
  max_tiles_ = buf_ / tile_ + nslots_ + 1;
  mem_.alloc((size_t)bytes_needed(buf_, nslots_, kmax_, tile_));
  char *p = mem_.as<char>();
  auto carve = [&](int64_t bytes) {
    char *q = p;
    p += align_up(bytes, 256);
    return q;
  };
  for (auto &r : role_) r = reinterpret_cast<double *>(carve(buf_ * 8));
  alpha_ = reinterpret_cast<double *>(carve((int64_t)nslots_ * kmax_ * 8));
  beta_ = reinterpret_cast<double *>(carve((int64_t)nslots_ * kmax_ * 8));
  (void)carve((int64_t)nslots_ * kmax_ * 8);  // spare row block (keeps the layout of bytes_needed)
  bnorm_ = reinterpret_cast<double *>(carve((int64_t)nslots_ * 8));
  partial_ = reinterpret_cast<double *>(carve((max_tiles_ + 2 * kRedBlocks) * 8));
  done_ = reinterpret_cast<int *>(carve((int64_t)nslots_ * 4));
  step_[0] = reinterpret_cast<int *>(carve((int64_t)nslots_ * 4));
  step_[1] = reinterpret_cast<int *>(carve((int64_t)nslots_ * 4));
  d_inst_ = reinterpret_cast<LaneInst *>(carve((int64_t)nslots_ * sizeof(LaneInst)));
  d_tiles_ = reinterpret_cast<BTile *>(carve(max_tiles_ * (int64_t)sizeof(BTile)));
  const size_t stage = (size_t)nslots_ * sizeof(LaneInst) + (size_t)max_tiles_ * sizeof(BTile);
  for (int i = 0; i < 4; ++i) {
    void *h = nullptr;
    MB_HIP(hipHostMalloc(&h, stage, hipHostMallocDefault));
    staging_.push_back(static_cast<char *>(h));
    hipEvent_t e;
    MB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    staging_ev_.push_back(e);
  }
  for (int i = nslots_ - 1; i >= 0; --i) free_slots_.push_back(i);
  free_vec_.push_back({0, buf_});
  static std::atomic<bool> attr{false};  // idempotent; host threads may race
  if (!attr) {
    MB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&ln_hv_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
    attr = true;
  }
}

LaneEngine::~LaneEngine() {
  for (size_t i = 0; i < staging_.size(); ++i) {
    (void)hipEventSynchronize(staging_ev_[i]);
    (void)hipEventDestroy(staging_ev_[i]);
    (void)hipHostFree(staging_[i]);
  }
}

int64_t LaneEngine::vec_take(int64_t n) {
  n = align_up(std::max<int64_t>(n, 1), 32);  // 256-byte ranges
  for (size_t i = 0; i < free_vec_.size(); ++i) {
    auto &f = free_vec_[i];
    if (f.second < n) continue;
    const int64_t off = f.first;
    f.first += n;
    f.second -= n;
    if (f.second == 0) free_vec_.erase(free_vec_.begin() + i);
    return off;
  }
  return -1;
}

void LaneEngine::vec_give(int64_t off, int64_t n) {
  n = align_up(std::max<int64_t>(n, 1), 32);
  auto it = std::lower_bound(free_vec_.begin(), free_vec_.end(), std::make_pair(off, (int64_t)0));
  it = free_vec_.insert(it, {off, n});
  size_t i = it - free_vec_.begin();
  if (i + 1 < free_vec_.size() && free_vec_[i].first + free_vec_[i].second == free_vec_[i + 1].first) {
    free_vec_[i].second += free_vec_[i + 1].second;
    free_vec_.erase(free_vec_.begin() + i + 1);
  }
  if (i > 0 && free_vec_[i - 1].first + free_vec_[i - 1].second == free_vec_[i].first) {
    free_vec_[i - 1].second += free_vec_[i].second;
    free_vec_.erase(free_vec_.begin() + i);
  }
}

int LaneEngine::slot_take() {
  if (free_slots_.empty()) return -1;
  const int s = free_slots_.back();
  free_slots_.pop_back();
  return s;
}

void LaneEngine::slot_give(int slot) { free_slots_.push_back(slot); }

void LaneEngine::init_member(const LaneInst &I, const PermTables &p, uint64_t seed, hipStream_t s) {
  if (I.off < 0 || I.off + 2 * I.D > buf_) throw std::out_of_range("lane vector range");
  double *x = role_[0] + I.off;  // the input role: the first step normalises it
  launch_fill_gaussian_c128(p, seed, x, s);
  launch_norm(x, 2 * I.D, partial_ + max_tiles_, bnorm_ + I.slot, s);
  hipLaunchKernelGGL(ln_init_kernel, dim3(1), dim3(kWave), 0, s, (const double *)bnorm_, I.slot,
                     step_[cur_], done_);
  MB_HIP(hipGetLastError());
}

void LaneEngine::set_members(const std::vector<LaneInst> &members, hipStream_t s) {
  if ((int)members.size() > nslots_) throw std::invalid_argument("too many lane members");
  std::vector<LaneInst> insts(members);
  std::vector<BTile> tiles;
  lds_bytes_ = 0;
  for (size_t i = 0; i < insts.size(); ++i) {
    LaneInst &I = insts[i];
    I.tile0 = (int)tiles.size();
    const int64_t Ddn = I.t.d_dn;
    const int64_t rows_n = std::max<int64_t>(1, tile_ / Ddn) * Ddn;
    I.rows = (I.t.comp_ok && Ddn <= kMaxRowsTileDn &&
              rows_tile_lds(Ddn, rows_n, I.t.max_up_row) <= kRowsTileLdsMax) ? 1 : 0;
    const int64_t stp = I.rows ? rows_n : tile_;
    for (int64_t e = 0; e < I.D; e += stp) {
      BTile tl;
      tl.inst = (int)i;
      tl.e0 = e;
      tl.n = (int)std::min<int64_t>(stp, I.D - e);
      tiles.push_back(tl);
      if (I.rows) lds_bytes_ = std::max<size_t>(lds_bytes_, rows_tile_lds(Ddn, tl.n, I.t.max_up_row));
    }
    I.tile1 = (int)tiles.size();
  }
  if ((int64_t)tiles.size() > max_tiles_) throw std::runtime_error("lane tile table overflow");
  n_members_ = (int)insts.size();
  n_tiles_ = (int)tiles.size();
  if (n_members_ == 0) return;
  // staged through a pinned ring buffer; a buffer is reused only after its copy ran
  const int b = staging_next_;
  staging_next_ = (staging_next_ + 1) % (int)staging_.size();
  MB_HIP(hipEventSynchronize(staging_ev_[b]));
  char *h = staging_[b];
  const size_t ib = insts.size() * sizeof(LaneInst), tb = tiles.size() * sizeof(BTile);
  std::copy(reinterpret_cast<const char *>(insts.data()), reinterpret_cast<const char *>(insts.data()) + ib, h);
  std::copy(reinterpret_cast<const char *>(tiles.data()), reinterpret_cast<const char *>(tiles.data()) + tb,
            h + ib);
  MB_HIP(hipMemcpyAsync(d_inst_, h, ib, hipMemcpyHostToDevice, s));
  MB_HIP(hipMemcpyAsync(d_tiles_, h + ib, tb, hipMemcpyHostToDevice, s));
  MB_HIP(hipEventRecord(staging_ev_[b], s));
}

void LaneEngine::tick(hipStream_t s) {
  if (n_members_ == 0) return;
  double *x = role_[0], *xp = role_[1], *w = role_[2];
  const int *sin = step_[cur_];
  int *sout = step_[cur_ ^ 1];
  const dim3 gt((unsigned)n_tiles_), gm((unsigned)n_members_), bl(kLaneBlock), bw(kWave);
  hipLaunchKernelGGL(ln_hv_kernel, gt, bl, lds_bytes_, s, d_inst_, d_tiles_, (const double *)x, nullptr,
                     (const double *)xp, w, (const double *)beta_, (const double *)bnorm_, sin,
                     (const int *)done_, kmax_, partial_);
  hipLaunchKernelGGL(ln_reduce_kernel, gm, bw, 0, s, d_inst_, (const double *)partial_, alpha_, beta_,
                     done_, sin, sout, kmax_, 0);
  hipLaunchKernelGGL(ln_axpy_kernel, gt, bl, 0, s, d_inst_, d_tiles_, (const double *)x, w,
                     (const double *)alpha_, (const double *)beta_, (const double *)bnorm_, sin,
                     (const int *)done_, kmax_, partial_);
  hipLaunchKernelGGL(ln_reduce_kernel, gm, bw, 0, s, d_inst_, (const double *)partial_, alpha_, beta_,
                     done_, sin, sout, kmax_, 1);
  MB_HIP(hipGetLastError());
  // rotate: the next input is w, x becomes x_prev; the old x_prev is free
  role_[0] = w;
  role_[1] = x;
  role_[2] = xp;
  cur_ ^= 1;
}

void LaneEngine::fetch(int slot, double *alpha, double *beta, int *steps, hipStream_t s) {
  MB_HIP(hipMemcpyAsync(alpha, alpha_ + (int64_t)slot * kmax_, sizeof(double) * kmax_,
                        hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(beta, beta_ + (int64_t)slot * kmax_, sizeof(double) * kmax_,
                        hipMemcpyDeviceToHost, s));
  MB_HIP(hipMemcpyAsync(steps, step_[cur_] + slot, sizeof(int), hipMemcpyDeviceToHost, s));
}

}  // namespace mb
