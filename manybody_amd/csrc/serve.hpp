// Serving loop: ground-state, KPM density-of-states and Krylov time-evolution
// requests of mixed (lattice, N_up, N_dn, U), admitted as slots and device
// memory free up, executed concurrently on the device (see serve.cpp).
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mb_internal.hpp"

namespace mb {

enum ReqKind : int { kReqGround = 0, kReqKpm = 1, kReqEvolve = 2 };

struct ServeRequest {
  int kind = kReqGround;
  int ndim = 1;
  int sizes[3] = {1, 1, 1};
  int n_up = 0, n_dn = 0;
  double t = 1.0, u = 0.0;
  int steps = 150;              // ground: Lanczos steps; KPM: bound-estimate Lanczos steps;
                                // evolve: Krylov dimension per time step
  int moments = 256, nvec = 4;  // KPM
  double t_final = 0.0, dt = 0.05;  // evolve
  uint64_t seed = 0;
  double arrival = -1.0;  // seconds after the server started; < 0: on submission
};

struct ServeResult {
  int status = 0;  // MB_OK or an MB_ERR_* code
  std::string error;
  int64_t dim = 0, footprint = 0, peak = 0;  // D, admitted bytes, bytes actually used
  double e0 = 0.0;                           // ground: lowest Ritz value
  double e_min = 0.0, e_max = 0.0;           // KPM: spectral bounds used for the rescaling
  double ov_re = 0.0, ov_im = 0.0, norm = 0.0;  // evolve: <psi0|psi(t)>, |psi(t)|
  std::vector<double> mu;                    // KPM: moments 0 .. moments-1
  double arrival = 0.0, admitted = 0.0, finished = 0.0;  // seconds after server start
  int slot = -1;
};

// Device bytes a request may use (tables + vectors + workspace): admission
// reserves exactly this much; run_request fails with MB_ERR_NOMEM beyond it.
// `tables` is the part for the basis strings and hop tables.
struct Footprint {
  int64_t dim, tables, work;
};
Footprint footprint_parts(const ServeRequest &r);
int64_t request_footprint(const ServeRequest &r);

// First-fit pool over one device allocation (offsets in bytes, 256-aligned).
// Not synchronised: the server's mutex protects it.
class Arena {
 public:
  explicit Arena(int64_t bytes);
  ~Arena();
  Arena(const Arena &) = delete;
  Arena &operator=(const Arena &) = delete;
  int64_t take(int64_t n);  // offset, or -1 if no free range is large enough
  void give(int64_t off, int64_t n);
  char *base() const { return base_; }
  int64_t capacity() const { return cap_; }
  int64_t used() const { return used_; }

 private:
  char *base_ = nullptr;
  int64_t cap_ = 0, used_ = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
};

// Stack allocator over one admitted range; buffers are freed in any order and
// the top is popped past every freed buffer.  Used by one worker thread.
class Region final : public DevAlloc {
 public:
  Region(char *base, int64_t size) : base_(base), size_(size) {}
  void *alloc(size_t n) override;
  void free(void *p) override;
  int64_t peak() const { return peak_; }

 private:
  struct Item {
    int64_t off, len;
    bool freed;
  };
  char *base_;
  int64_t size_, top_ = 0, peak_ = 0;
  std::vector<Item> stack_;
};

// Runs one request on stream s; device allocations come from the calling
// thread's DevAlloc.  Fills the result fields of its kind and dim.
void run_request(const ServeRequest &r, ServeResult &res, hipStream_t s);

class LaneEngine;

class Server {
 public:
  // lane_bytes > 0 reserves that much of the arena for the Lanczos lane: ground-
  // state requests of dimension <= lane_max_dim join it at step boundaries and
  // share its launches instead of taking a slot.
  Server(int slots, int64_t arena_bytes, int64_t lane_bytes = 0, int64_t lane_max_dim = 1 << 20);
  ~Server();  // finishes every submitted request, then stops the workers
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  int64_t submit(const ServeRequest &r);
  const ServeResult &wait(int64_t id);  // blocks until request id has finished
  void drain();                         // blocks until every submitted request has finished
  double now() const;
  int64_t arena_bytes() const { return arena_.capacity(); }
  int slots() const { return n_slots_; }

 private:
  struct Job {
    int64_t id;
    ServeRequest req;
    ServeResult res;
    Footprint fp{};
    double cost = 0.0;  // expected service time alone (seconds), for the admission order
    bool lane = false;  // runs in the Lanczos lane
    bool done = false;
  };
  struct LaneMember;
  void worker(int slot);
  void lane_worker();
  void finish_locked(Job *j);
  Job *pick_locked(double t, double *next_arrival, int64_t *off, bool lane);

  std::chrono::steady_clock::time_point t0_;
  int device_ = 0;
  Arena arena_;
  std::mutex m_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Job *> waiting_;  // by arrival time
  std::unordered_map<int64_t, std::unique_ptr<Job>> jobs_;
  int64_t next_id_ = 0, pending_ = 0;
  bool stop_ = false;
  int n_slots_ = 0;
  // Lanczos lane
  int64_t lane_max_dim_ = 0, lane_off_ = -1, lane_bytes_ = 0;
  std::unique_ptr<Region> lane_region_;
  std::unique_ptr<LaneEngine> lane_;
  std::vector<std::thread> threads_;
};

}  // namespace mb
