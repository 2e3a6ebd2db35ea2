// Serving loop.
//
// A request is one ground-state solve (FP64 Lanczos, lowest Ritz value), one
// KPM density of states (FP64 Lanczos bounds, then FP32 Chebyshev moments over
// random-phase vectors, FP64 accumulation) or one Krylov time evolution
// (exp(-iHt) in steps of dt, two-pass Lanczos per step) on a periodic 1D, 2D or
// 3D Hubbard model of any (N_up, N_dn, t, U).  Its basis, hop tables and every
// Krylov vector live on the device for the whole request.
//
// Continuous admission: `slots` worker threads, each with its own non-blocking
// HIP stream, take, among the arrived requests whose device footprint fits in
// the free part of one pre-allocated arena, the one with the longest expected
// service time (a request that has waited longer than kBypassSeconds goes
// first and is not overtaken).  Longest first keeps a long chain (a Krylov
// evolution is thousands of dependent small steps) from starting last and
// setting the end of a burst.  The admitted range becomes the
// request's private stack allocator, so the loop never calls hipMalloc or
// hipFree (which would synchronise the whole device) and a request can never
// take memory it was not admitted with.  Requests that arrive while others run
// are admitted as soon as a slot and enough memory are free; the kernels of all
// running requests share the device.  Every reduction in the library is
// atomic-free with a fixed order, so a request's result does not depend on what
// else runs beside it.
#include "serve.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/manybody_hip.h"
#include "krylov.hpp"
#include "lane.hpp"

namespace mb {

namespace {

constexpr int64_t kAlign = 256;
constexpr double kBypassSeconds = 0.5;

int64_t round_up(int64_t n, int64_t a) { return (n + a - 1) / a * a; }

int status_of(const std::exception &e) {
  if (dynamic_cast<const NoMemError *>(&e)) return MB_ERR_NOMEM;
  if (dynamic_cast<const DeviceError *>(&e)) return MB_ERR_DEVICE;
  if (dynamic_cast<const std::invalid_argument *>(&e)) return MB_ERR_INVALID;
  if (dynamic_cast<const std::out_of_range *>(&e)) return MB_ERR_RANGE;
  return MB_ERR_RUNTIME;
}

int lattice_sites(const ServeRequest &r) {
  if (r.ndim < 1 || r.ndim > 3) throw std::invalid_argument("ndim must be 1, 2 or 3");
  int L = 1;
  for (int i = 0; i < r.ndim; ++i) {
    if (r.sizes[i] < 1) throw std::invalid_argument("lattice sizes must be positive");
    L *= r.sizes[i];
  }
  if (L > kMaxSites) throw std::invalid_argument("at most 32 sites");
  return L;
}

}  // namespace

// ------------------------------------------------------------ memory

Arena::Arena(int64_t bytes) : cap_(bytes / kAlign * kAlign) {
  if (cap_ <= 0) throw std::invalid_argument("arena size must be positive");
  if (hipMalloc(reinterpret_cast<void **>(&base_), (size_t)cap_) != hipSuccess) {
    (void)hipGetLastError();
    throw NoMemError("cannot allocate a serving arena of " + std::to_string(cap_) + " bytes");
  }
  free_[0] = cap_;
}

Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

int64_t Arena::take(int64_t n) {
  n = round_up(std::max<int64_t>(n, 1), kAlign);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < n) continue;
    const int64_t off = it->first, rest = it->second - n;
    free_.erase(it);
    if (rest > 0) free_[off + n] = rest;
    used_ += n;
    return off;
  }
  return -1;
}

void Arena::give(int64_t off, int64_t n) {
  n = round_up(std::max<int64_t>(n, 1), kAlign);
  used_ -= n;
  auto it = free_.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}

void *Region::alloc(size_t n) {
  const int64_t len = round_up((int64_t)std::max<size_t>(n, 1), kAlign);
  if (top_ + len > size_) return nullptr;
  void *p = base_ + top_;
  stack_.push_back({top_, len, false});
  top_ += len;
  peak_ = std::max(peak_, top_);
  return p;
}

void Region::free(void *p) {
  const int64_t off = static_cast<char *>(p) - base_;
  for (auto it = stack_.rbegin(); it != stack_.rend(); ++it)
    if (it->off == off && !it->freed) {
      it->freed = true;
      break;
    }
  while (!stack_.empty() && stack_.back().freed) {
    top_ = stack_.back().off;
    stack_.pop_back();
  }
}

// ------------------------------------------------------------ footprint

Footprint footprint_parts(const ServeRequest &r) {
  const int L = lattice_sites(r);
  if (r.n_up < 0 || r.n_dn < 0 || r.n_up > L || r.n_dn > L)
    throw std::invalid_argument("spin sector exceeds the orbital count");
  const Binom &C = binom();
  const int64_t du = (int64_t)C(L, r.n_up), dd = (int64_t)C(L, r.n_dn), D = du * dd;
  const int64_t nb = (int64_t)L * r.ndim;  // bonds of the periodic lattice
  const int64_t maxrow = 2 * nb;           // hops per string, both directions
  auto A = [](int64_t n) { return round_up(n + 16, kAlign); };
  auto species = [&](int64_t ds) {
    const int64_t nnz = ds * maxrow, dp = round_up(ds, 64);
    const int64_t ds_sorted = (ds >= kSortedRound && round_up(ds, kSortedRound) <=
                               (int64_t)kSortedRound * kSortedRounds) ? round_up(ds, kSortedRound) : 0;
    return A(4 * ds) /*strings*/ + A(4 * (ds + 1)) + A(4 * nnz) + A(8 * nnz) + A(4 * nnz) +
           A(8 * ds) + A(4 * ds) /*CSR, onsite*/ + A(4 * maxrow * ds) + A(ds) + A(8 * 128) +
           A(4 * 128) /*ELL*/ + A(64 * dp) + A(4 * dp) + A(dp / 64) /*compact*/ +
           A(2 * nnz) /*16-bit*/ +
           (ds_sorted ? A(2 * ds_sorted) + A(64 * ds_sorted) + A(2 * ds_sorted / 64) : 0) /*sorted*/;
  };
  const int64_t tables = species(du) + species(dd) + A(8 * 33 * 33) + 2 * A(8 * (L + 1));
  const int64_t vec = A(16 * D);
  const int64_t rows2 = std::max<int64_t>(2 * du, 2 * 2048);  // partial slots of a 2-plane H.v
  auto lanczos_work = [&](int k) {
    return 3 * vec + A(8 * (2 * (int64_t)std::max(k, 1) + 2 + std::max<int64_t>(512, rows2)));
  };
  int64_t work = 0;
  if (r.kind == kReqGround) {
    work = vec + lanczos_work(r.steps);
  } else if (r.kind == kReqKpm) {
    const int64_t P = 2 * (int64_t)r.nvec;
    const int64_t planes = A(4 * D * P);
    const int64_t slots = std::max<int64_t>(P * du, P * 2048);
    const int64_t cheb = 3 * planes + A(8 * std::max<int64_t>(2 * slots, 512)) +
                         A(8 * (int64_t)(r.moments / 2 + 1) * P);
    work = vec + std::max(lanczos_work(r.steps), cheb);
  } else if (r.kind == kReqEvolve) {
    // start + output vectors, then either the k + 1 stored Krylov vectors (kept
    // between the time steps) with the Lanczos scalars, or the regenerating
    // path's three Lanczos vectors and a temporary
    const int64_t k = (r.steps <= 0 || r.steps > D) ? D : r.steps;
    const int64_t scalars = A(8 * (2 * k + 2 + std::max<int64_t>(512, rows2))) + A(16 * k);
    work = 2 * vec + std::max(A(16 * D * (k + 1)) + scalars,
                              std::max(lanczos_work(r.steps) + vec, A(8 * (2 * 512 + 8))));
  } else {
    throw std::invalid_argument("unknown request kind");
  }
  return {D, tables, work};
}

// Expected service time of a request running alone (seconds): Lanczos-type
// steps times a per-step cost of a launch-bound floor plus the HBM passes over
// the vector.  Only the order matters (admission picks the largest).
double expected_seconds(const ServeRequest &r, int64_t D) {
  const double floor_s = 20e-6, per_elem = 6.0 * 16.0 / 5e12;  // ~6 complex passes at 5 TB/s
  const double step = floor_s + per_elem * (double)D;
  if (r.kind == kReqGround) return r.steps * step;
  if (r.kind == kReqKpm)
    return r.steps * step + 0.5 * r.moments * (floor_s + 0.5 * per_elem * (double)D * r.nvec);
  if (r.kind == kReqEvolve) {
    const double n_t = r.dt > 0.0 ? std::ceil(std::fabs(r.t_final) / r.dt) : 1.0;
    return n_t * std::max(r.steps, 1) * step;
  }
  return 0.0;
}

int64_t request_footprint(const ServeRequest &r) {
  const Footprint f = footprint_parts(r);
  return f.tables + f.work + 64 * 1024;
}

// ------------------------------------------------------------ one request

void run_request(const ServeRequest &r, ServeResult &res, hipStream_t s) {
  const int L = lattice_sites(r);
  Sector sec;
  build_sector(sec, L, r.n_up, r.n_dn);
  sec.ensure_device();
  Hubbard h;
  h.sec = &sec;
  h.t = r.t;
  h.U = r.u;
  lattice_bonds(r.ndim, r.sizes, h.bond_from, h.bond_to);
  h.ndim = r.ndim;
  for (int i = 0; i < r.ndim; ++i) h.lsize[i] = r.sizes[i];
  build_hubbard(h);
  const int64_t D = sec.dim();
  res.dim = D;
  DevBuf b((size_t)D * 16);
  launch_fill_gaussian_c128(perm_of(sec), r.seed, b.p, s);
  if (r.kind == kReqGround) {
    res.e0 = find_min_eigenpair(h, b.p, r.steps, nullptr, s);
  } else if (r.kind == kReqKpm) {
    if (r.moments < 2 || (r.moments & 1)) throw std::invalid_argument("moments must be even and >= 2");
    // spectral bounds: extreme Ritz values of an FP64 Lanczos run, widened by 1 %
    // of the width, then mapped to [-0.99, 0.99] (rescaling.h:33-38)
    const LanczosResult dec = lanczos_pass_one(h, b.p, r.steps, s);
    const int m = dec.steps;
    if (m == 0) throw std::runtime_error("Lanczos produced no steps for the spectral bounds");
    std::vector<double> w(m);
    tridiag_eig(m, dec.alphas.data(), dec.betas.data(), w.data(), nullptr);
    const double lo = w.front(), hi = w.back(), margin = 0.01 * (hi - lo) + 1e-12;
    res.e_min = lo - margin;
    res.e_max = hi + margin;
    const double pad = 1.0 / 0.99 - 1.0;
    const double a = (res.e_max - res.e_min) / 2.0 * (1.0 + pad), c = (res.e_max + res.e_min) / 2.0;
    res.mu.assign(r.moments, 0.0);
    kpm_dos_moments(h, sec, a, c, r.moments, r.nvec, r.seed, true, res.mu.data(), s);
    for (double v : res.mu)
      if (!(std::fabs(v) <= 1.0 + 1e-3))
        throw std::runtime_error("Chebyshev moments left [-1, 1]: spectral bounds too tight");
  } else if (r.kind == kReqEvolve) {
    DevBuf out((size_t)D * 16);
    time_evolve_steps(h, b.p, 0.0, r.t_final, r.dt, r.steps, out.p, s);
    DevBuf red((size_t)(2 * kRedBlocks + 8) * sizeof(double));
    double *rp = red.as<double>(), *res_dev = rp + 2 * kRedBlocks;
    launch_cdot(b.as<double>(), out.as<double>(), D, rp, res_dev, s);  // <b|out>
    launch_norm(out.as<double>(), 2 * D, rp, res_dev + 2, s);          // |out|
    launch_norm(b.as<double>(), 2 * D, rp, res_dev + 3, s);            // |b|
    double hv[4];
    MB_HIP(hipMemcpyAsync(hv, res_dev, sizeof(hv), hipMemcpyDeviceToHost, s));
    MB_HIP(hipStreamSynchronize(s));
    const double nb2 = hv[3] * hv[3];
    res.ov_re = hv[0] / nb2;
    res.ov_im = hv[1] / nb2;
    res.norm = hv[2] / hv[3];
  } else {
    throw std::invalid_argument("unknown request kind");
  }
  MB_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ server

namespace {
constexpr int kLaneSlots = 256;  // members of the Lanczos lane at once
constexpr int kLaneKmax = 512;   // longest Lanczos run the lane takes
}  // namespace

struct Server::LaneMember {
  Job *job = nullptr;
  int64_t arena_off = -1, arena_len = 0;
  std::unique_ptr<Region> reg;  // tables; outlives sec and h (declared first)
  Sector sec;
  Hubbard h;
  int slot = -1, k = 0;
  int64_t voff = -1, vlen = 0;
  int64_t last_tick = 0;  // the member's last step runs in this tick
};

Server::Server(int slots, int64_t arena_bytes, int64_t lane_bytes, int64_t lane_max_dim)
    : t0_(std::chrono::steady_clock::now()), arena_(arena_bytes) {
  if (slots < 1 || slots > 64) throw std::invalid_argument("slots must be in [1, 64]");
  MB_HIP(hipGetDevice(&device_));
  n_slots_ = slots;
  if (lane_bytes > 0) {
    int64_t buf = std::max<int64_t>((lane_bytes - (32 << 20)) / 32, 0);
    while (buf >= 1024 && LaneEngine::bytes_needed(buf, kLaneSlots, kLaneKmax) > lane_bytes)
      buf -= std::max<int64_t>(buf / 64, 1024);
    if (buf < 1024) throw std::invalid_argument("lane_bytes too small for the Lanczos lane");
    lane_bytes_ = LaneEngine::bytes_needed(buf, kLaneSlots, kLaneKmax) + kAlign;
    lane_off_ = arena_.take(lane_bytes_);
    if (lane_off_ < 0) throw std::invalid_argument("lane_bytes exceeds the arena");
    lane_region_ = std::make_unique<Region>(arena_.base() + lane_off_, lane_bytes_);
    thread_dev_alloc() = lane_region_.get();
    try {
      lane_ = std::make_unique<LaneEngine>(buf, kLaneSlots, kLaneKmax);
    } catch (...) {
      thread_dev_alloc() = nullptr;
      throw;
    }
    thread_dev_alloc() = nullptr;
    // a member holds 2 D doubles in each role buffer; keep room for several
    lane_max_dim_ = std::min<int64_t>(lane_max_dim, buf / 8);
  }
  for (int i = 0; i < slots; ++i) threads_.emplace_back(&Server::worker, this, i);
  if (lane_) threads_.emplace_back(&Server::lane_worker, this);
}

Server::~Server() {
  drain();
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto &t : threads_) t.join();
  lane_.reset();
  lane_region_.reset();
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int64_t Server::submit(const ServeRequest &r) {
  auto j = std::make_unique<Job>();
  j->req = r;
  j->res.arrival = r.arrival >= 0.0 ? r.arrival : now();
  j->fp = footprint_parts(r);  // throws on an invalid request
  j->cost = expected_seconds(r, j->fp.dim);
  j->lane = lane_ && r.kind == kReqGround && j->fp.dim <= lane_max_dim_ && r.steps <= kLaneKmax;
  // lane members hold their three vectors in the lane's buffers
  j->res.footprint = j->lane ? j->fp.tables + 64 * 1024 + 3 * 16 * j->fp.dim : request_footprint(r);
  std::lock_guard<std::mutex> lk(m_);
  j->id = next_id_++;
  Job *p = j.get();
  jobs_.emplace(p->id, std::move(j));
  auto pos = std::upper_bound(waiting_.begin(), waiting_.end(), p, [](const Job *a, const Job *b) {
    return a->res.arrival < b->res.arrival;
  });
  waiting_.insert(pos, p);
  ++pending_;
  cv_work_.notify_all();
  return p->id;
}

const ServeResult &Server::wait(int64_t id) {
  std::unique_lock<std::mutex> lk(m_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw std::out_of_range("unknown request id");
  Job *j = it->second.get();
  cv_done_.wait(lk, [&] { return j->done; });
  return j->res;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(m_);
  cv_done_.wait(lk, [&] { return pending_ == 0; });
}

void Server::finish_locked(Job *j) {
  j->done = true;
  --pending_;
}

// The next arrived request of the given kind (lane or slot) that fits; sets
// *off to its arena offset.  Lane requests go in arrival order.  Slot requests:
// the oldest if it has waited longer than kBypassSeconds (and nothing overtakes
// it while it does not fit), else the one with the longest expected service
// time that fits.  For the lane, the arena holds only the tables (the vectors
// go to the lane's buffers, whose range and slot the caller takes).
Server::Job *Server::pick_locked(double t, double *next_arrival, int64_t *off, bool lane) {
  *next_arrival = -1.0;
  std::vector<std::deque<Job *>::iterator> cands;  // arrived slot requests, by arrival
  for (auto it = waiting_.begin(); it != waiting_.end(); ++it) {
    Job *c = *it;
    if (c->res.arrival > t) {
      *next_arrival = c->res.arrival;  // the queue is ordered by arrival
      break;
    }
    if (c->lane != lane) continue;
    if (lane) {
      const int64_t need = c->fp.tables + 64 * 1024;
      const int64_t o = arena_.take(need);
      if (o >= 0) {
        *off = o;
        waiting_.erase(it);
        return c;
      }
      if (t - c->res.arrival > kBypassSeconds) break;
      continue;
    }
    if (c->res.footprint > arena_.capacity()) {  // can never fit: fail it now
      c->res.status = MB_ERR_NOMEM;
      c->res.error = "request needs " + std::to_string(c->res.footprint) +
                     " bytes of device memory; the serving arena holds " +
                     std::to_string(arena_.capacity());
      c->res.admitted = c->res.finished = t;
      c->done = true;
      --pending_;
      waiting_.erase(it);
      cv_done_.notify_all();
      return pick_locked(t, next_arrival, off, lane);
    }
    cands.push_back(it);
  }
  if (cands.empty()) return nullptr;
  if (t - (*cands.front())->res.arrival > kBypassSeconds) {  // a long-waiting request goes first
    const int64_t o = arena_.take((*cands.front())->res.footprint);
    if (o < 0) return nullptr;  // and is not overtaken
    Job *c = *cands.front();
    *off = o;
    waiting_.erase(cands.front());
    return c;
  }
  std::stable_sort(cands.begin(), cands.end(),
                   [](const auto &a, const auto &b) { return (*a)->cost > (*b)->cost; });
  for (auto &it : cands) {
    const int64_t o = arena_.take((*it)->res.footprint);
    if (o >= 0) {
      Job *c = *it;
      *off = o;
      waiting_.erase(it);
      return c;
    }
  }
  return nullptr;
}

void Server::worker(int slot) {
  (void)hipSetDevice(device_);
  hipStream_t s = nullptr;
  const bool have_stream = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  for (;;) {
    Job *j = nullptr;
    int64_t off = -1;
    {
      std::unique_lock<std::mutex> lk(m_);
      for (;;) {
        if (stop_ && waiting_.empty()) {
          lk.unlock();
          if (have_stream) (void)hipStreamDestroy(s);
          return;
        }
        double next = -1.0;
        j = pick_locked(now(), &next, &off, false);
        if (j) break;
        if (next >= 0.0)
          cv_work_.wait_until(lk, t0_ + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                            std::chrono::duration<double>(next)));
        else
          cv_work_.wait(lk);
      }
    }
    ServeResult &res = j->res;
    res.admitted = now();
    res.slot = slot;
    {
      Region region(arena_.base() + off, res.footprint);
      thread_dev_alloc() = &region;
      thread_upload_stream() = s;
      try {
        if (!have_stream) throw DeviceError("cannot create a HIP stream for the serving slot");
        run_request(j->req, res, s);
        res.status = MB_OK;
      } catch (const std::exception &e) {
        res.status = status_of(e);
        res.error = e.what();
        if (have_stream) (void)hipStreamSynchronize(s);
      }
      thread_dev_alloc() = nullptr;
      thread_upload_stream() = nullptr;
      res.peak = region.peak();
    }
    res.finished = now();
    {
      std::lock_guard<std::mutex> lk(m_);
      arena_.give(off, res.footprint);
      finish_locked(j);
    }
    cv_done_.notify_all();
    cv_work_.notify_all();
  }
}

// The Lanczos lane: admits arrived lane requests at step boundaries (slot,
// vector range and table memory permitting), ticks every member one step per
// round, and completes members after their k-th step.  At most kInFlight ticks
// are queued ahead of the device, so a new request waits at most that many
// steps before it joins.
void Server::lane_worker() {
  constexpr int kInFlight = 3;
  (void)hipSetDevice(device_);
  hipStream_t ls = nullptr, us = nullptr;
  std::vector<hipEvent_t> tick_ev(kInFlight + 1, nullptr), slot_ev(kLaneSlots, nullptr);
  double *pin = nullptr;  // per slot: alpha[kmax], beta[kmax], steps
  bool ok = hipStreamCreateWithFlags(&ls, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&us, hipStreamNonBlocking) == hipSuccess &&
            hipHostMalloc(reinterpret_cast<void **>(&pin),
                          sizeof(double) * kLaneSlots * (2 * kLaneKmax + 1), hipHostMallocDefault) ==
                hipSuccess;
  for (auto &e : tick_ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  for (auto &e : slot_ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  auto pin_alpha = [&](int slot) { return pin + (size_t)slot * (2 * kLaneKmax + 1); };
  auto pin_beta = [&](int slot) { return pin_alpha(slot) + kLaneKmax; };
  auto pin_steps = [&](int slot) { return reinterpret_cast<int *>(pin_alpha(slot) + 2 * kLaneKmax); };

  std::vector<std::unique_ptr<LaneMember>> active, finishing;
  int64_t ticks = 0;
  bool changed = false;
  auto release = [&](LaneMember &m) {  // lane resources; the caller holds no lock
    if (m.voff >= 0) lane_->vec_give(m.voff, m.vlen);
    if (m.slot >= 0) lane_->slot_give(m.slot);
    m.voff = -1;
    m.slot = -1;
  };
  for (;;) {
    std::vector<std::unique_ptr<LaneMember>> joining;
    {
      std::unique_lock<std::mutex> lk(m_);
      for (;;) {
        double next = -1.0;
        for (;;) {  // admit every arrived lane request that fits
          if (!ok) break;
          const int slot = lane_->slot_take();
          if (slot < 0) break;
          int64_t off = -1;
          Job *j = pick_locked(now(), &next, &off, true);
          if (!j) {
            lane_->slot_give(slot);
            break;
          }
          const int64_t vlen = 2 * j->fp.dim;
          const int64_t voff = lane_->vec_take(vlen);
          if (voff < 0) {  // no vector room now: put it back at its place in the queue
            lane_->slot_give(slot);
            arena_.give(off, j->fp.tables + 64 * 1024);
            auto pos = std::upper_bound(waiting_.begin(), waiting_.end(), j, [](const Job *a, const Job *b) {
              return a->res.arrival < b->res.arrival;
            });
            waiting_.insert(pos, j);
            break;
          }
          auto m = std::make_unique<LaneMember>();
          m->job = j;
          m->arena_off = off;
          m->arena_len = j->fp.tables + 64 * 1024;
          m->slot = slot;
          m->voff = voff;
          m->vlen = vlen;
          m->k = j->req.steps;
          j->res.admitted = now();
          j->res.slot = n_slots_ + slot;
          joining.push_back(std::move(m));
        }
        if (!ok) {  // no stream: fail every lane request
          for (auto it = waiting_.begin(); it != waiting_.end();) {
            if ((*it)->lane && (*it)->res.arrival <= now()) {
              (*it)->res.status = MB_ERR_DEVICE;
              (*it)->res.error = "cannot create the Lanczos lane's streams";
              finish_locked(*it);
              it = waiting_.erase(it);
            } else {
              ++it;
            }
          }
          cv_done_.notify_all();
        }
        if (!joining.empty() || !active.empty() || !finishing.empty()) break;
        bool lane_waiting = false;
        for (Job *c : waiting_) lane_waiting |= c->lane;
        if (stop_ && !lane_waiting) {
          lk.unlock();
          for (auto &e : tick_ev) if (e) (void)hipEventDestroy(e);
          for (auto &e : slot_ev) if (e) (void)hipEventDestroy(e);
          if (pin) (void)hipHostFree(pin);
          if (ls) (void)hipStreamDestroy(ls);
          if (us) (void)hipStreamDestroy(us);
          return;
        }
        if (next >= 0.0)
          cv_work_.wait_until(lk, t0_ + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                            std::chrono::duration<double>(next)));
        else
          cv_work_.wait(lk);
      }
    }
    // build the joining members' tables (on the upload stream) and start vectors
    for (auto &m : joining) {
      ServeResult &res = m->job->res;
      try {
        m->reg = std::make_unique<Region>(arena_.base() + m->arena_off, m->arena_len);
        thread_dev_alloc() = m->reg.get();
        thread_upload_stream() = us;
        const ServeRequest &r = m->job->req;
        build_sector(m->sec, lattice_sites(r), r.n_up, r.n_dn);
        m->sec.ensure_device();
        m->h.sec = &m->sec;
        m->h.t = r.t;
        m->h.U = r.u;
        lattice_bonds(r.ndim, r.sizes, m->h.bond_from, m->h.bond_to);
        build_hubbard(m->h);
        thread_dev_alloc() = nullptr;
        thread_upload_stream() = nullptr;
        LaneInst I{};
        I.t = tables_of(m->h);
        I.off = m->voff;
        I.D = m->sec.dim();
        I.slot = m->slot;
        I.k = m->k;
        lane_->init_member(I, perm_of(m->sec), r.seed, ls);
        res.dim = I.D;
        m->last_tick = ticks + m->k - 1;
        active.push_back(std::move(m));
        changed = true;
      } catch (const std::exception &e) {
        thread_dev_alloc() = nullptr;
        thread_upload_stream() = nullptr;
        res.status = status_of(e);
        res.error = e.what();
        res.finished = now();
        release(*m);
        Job *j = m->job;
        const int64_t off = m->arena_off, len = m->arena_len;
        m.reset();
        std::lock_guard<std::mutex> lk(m_);
        arena_.give(off, len);
        finish_locked(j);
        cv_done_.notify_all();
      }
    }
    if (changed) {
      std::vector<LaneInst> insts;
      for (auto &m : active) {
        LaneInst I{};
        I.t = tables_of(m->h);
        I.off = m->voff;
        I.D = m->sec.dim();
        I.slot = m->slot;
        I.k = m->k;
        insts.push_back(I);
      }
      lane_->set_members(insts, ls);
      changed = false;
    }
    if (!active.empty()) {
      if (ticks >= kInFlight) (void)hipEventSynchronize(tick_ev[(ticks - kInFlight) % (kInFlight + 1)]);
      lane_->tick(ls);
      (void)hipEventRecord(tick_ev[ticks % (kInFlight + 1)], ls);
      for (auto it = active.begin(); it != active.end();) {
        LaneMember &m = **it;
        if (m.last_tick == ticks) {
          lane_->fetch(m.slot, pin_alpha(m.slot), pin_beta(m.slot), pin_steps(m.slot), ls);
          (void)hipEventRecord(slot_ev[m.slot], ls);
          finishing.push_back(std::move(*it));
          it = active.erase(it);
          changed = true;
        } else {
          ++it;
        }
      }
      ++ticks;
    }
    // complete members whose results have reached the host
    for (auto it = finishing.begin(); it != finishing.end();) {
      LaneMember &m = **it;
      const hipError_t q = active.empty() ? hipEventSynchronize(slot_ev[m.slot]) : hipEventQuery(slot_ev[m.slot]);
      if (q == hipErrorNotReady) {
        ++it;
        continue;
      }
      ServeResult &res = m.job->res;
      if (q != hipSuccess) {
        res.status = MB_ERR_DEVICE;
        res.error = hipGetErrorString(q);
      } else {
        const int steps = *pin_steps(m.slot);
        res.e0 = steps > 0 ? tridiag_min_eig(steps, pin_alpha(m.slot), pin_beta(m.slot)) : 0.0;
        res.status = MB_OK;
      }
      res.peak = m.reg ? m.reg->peak() + 4 * 16 * res.dim : 0;
      res.finished = now();
      release(m);
      Job *j = m.job;
      const int64_t off = m.arena_off, len = m.arena_len;
      it = finishing.erase(it);  // frees the member's tables into its region
      {
        std::lock_guard<std::mutex> lk(m_);
        arena_.give(off, len);
        finish_locked(j);
      }
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }
}

}  // namespace mb
