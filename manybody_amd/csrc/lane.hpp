// Continuous-batching Lanczos lane (see lane.hip).
#pragma once

#include <vector>

#include "batch.hpp"
#include "mb_internal.hpp"

namespace mb {

// One member of the lane as the kernels see it.
struct LaneInst {
  HvTables t;
  int64_t off, D;  // offset (doubles) of its vectors in each role buffer; plane stride D
  int tile0, tile1;
  int rows;        // 1: tiles are whole rows staged in LDS
  int slot;        // row of the per-slot state arrays
  int k;           // Lanczos steps requested
};

// Device side of the lane: three role buffers (x, x_prev, w) whose roles rotate
// every tick, per-slot Lanczos state (alpha, beta, step, done, start norm),
// member and tile tables.  Members join and leave between ticks; every member
// does one Lanczos step per tick, counted by its own step index, so members at
// different points of their recurrence share the same launches.  Driven by one
// host thread; every call is asynchronous on the stream it is given.
class LaneEngine {
 public:
  LaneEngine(int64_t buf_doubles, int nslots, int kmax, int tile = 1024);
  ~LaneEngine();
  LaneEngine(const LaneEngine &) = delete;
  LaneEngine &operator=(const LaneEngine &) = delete;
  static int64_t bytes_needed(int64_t buf_doubles, int nslots, int kmax, int tile = 1024);
  int64_t buf_doubles() const { return buf_; }
  int kmax() const { return kmax_; }

  int64_t vec_take(int64_t n);  // a range of n doubles in every role buffer, or -1
  void vec_give(int64_t off, int64_t n);
  int slot_take();  // -1 if all slots are in use
  void slot_give(int slot);

  // Seeded N(0,1) start vector (per reference index) of a new member in the
  // current input role, its norm, step 0.  Call before the member's first tick.
  void init_member(const LaneInst &I, const PermTables &p, uint64_t seed, hipStream_t s);
  // Replace the member list (tiles rebuilt and uploaded on s).
  void set_members(const std::vector<LaneInst> &members, hipStream_t s);
  // One Lanczos step for every member.
  void tick(hipStream_t s);
  // Copies slot's alpha[0..kmax), beta[0..kmax) and step count into host
  // (pinned) memory on s.
  void fetch(int slot, double *alpha, double *beta, int *steps, hipStream_t s);
  bool empty() const { return n_members_ == 0; }

 private:
  int64_t buf_;
  int nslots_, kmax_, tile_;
  int64_t max_tiles_;
  DevBuf mem_;
  double *role_[3];                    // role_[0] = x (step input), [1] = x_prev, [2] = w
  double *alpha_, *beta_, *bnorm_, *partial_;
  int *done_, *step_[2];
  int cur_ = 0;                        // step_[cur_] is read by the next tick
  LaneInst *d_inst_;
  BTile *d_tiles_;
  int n_members_ = 0, n_tiles_ = 0;
  size_t lds_bytes_ = 0;
  std::vector<char *> staging_;        // pinned upload buffers (ring)
  std::vector<hipEvent_t> staging_ev_;
  int staging_next_ = 0;
  std::vector<int> free_slots_;
  std::vector<std::pair<int64_t, int64_t>> free_vec_;  // (offset, length), sorted
};

}  // namespace mb
