// Batched Lanczos over many Hubbard instances (see batch.hip).
#pragma once

#include <utility>
#include <vector>

#include "krylov.hpp"
#include "mb_internal.hpp"

namespace mb {

// One instance of the batch: its tables, dimension, offset of its planar
// vector in the concatenated batch buffers, and its range of tiles.
struct BInst {
  HvTables t;
  int64_t off;  // in doubles; plane 0 at off, plane 1 at off + D
  int64_t D;
  int tile0, tile1;
  int rows;     // 1: tiles are whole rows, down hops read the rows from LDS
  int small;    // 1: all k steps in the persistent per-instance kernel (no tiles)
};

// Instances up to this dimension run every Lanczos step inside one persistent
// launch, one workgroup per instance; larger ones are tiled over the per-phase
// kernels.  Measured on 512 1D rings (L 8..12, 150 steps): 0.1421 s per batch
// with every instance tiled, 0.1427 s with the 200 instances up to 2048 in the
// persistent launch, 0.1442 / 0.1452 / 0.1621 s at 4096 / 8192 / 16384 (one
// workgroup per instance is latency-bound beyond a few thousand elements).
constexpr int64_t kPersistMaxDim = 2048;

// Krylov time step of one small real-space sector in one workgroup (batch.hip):
// k Lanczos steps from b with the vectors kept in basis ((k + 1) * 2 D doubles)
// and scal = alpha[k], beta[k], |b|, steps.  D <= kKrylovSmallMaxDim.
constexpr int64_t kKrylovSmallMaxDim = 2048;
void launch_krylov_small(const HvTables &t, int64_t D, const double *b, double *basis, double *scal,
                         int k, hipStream_t s);

// Longest row (D_dn) handled by row tiles.
constexpr int64_t kMaxRowsTileDn = 2048;
// LDS of a row tile of n elements (whole rows of D_dn): both planes at row
// stride D_dn + 1, then the up-hop lists of its rows (amplitudes, targets,
// counts); a row tile may use at most kRowsTileLdsMax bytes.
constexpr size_t kRowsTileLdsMax = 64 * 1024;
inline size_t rows_tile_lds(int64_t Ddn, int64_t n, int max_up) {
  const int64_t R = n / Ddn, S = Ddn + 1;
  return (size_t)(8 * (2 * R * S + R * max_up) + 4 * (R * max_up + R) + 64);
}

struct BTile {
  int inst;
  int n;
  int64_t e0;
};

class Batch {
 public:
  explicit Batch(std::vector<const Hubbard *> models);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  size_t size() const { return host_inst_.size(); }
  int64_t footprint_bytes() const;
  void set_mapping(int tile, int block);
  std::pair<int, int> mapping() const { return {tile_, block_}; }
  void fill_gaussian(const uint64_t *seeds, hipStream_t s);
  void set_start(size_t i, const void *b_dev, hipStream_t s);
  // k Lanczos steps for every instance from the start vectors; optionally as one hipGraph.
  void run(int k, bool use_graph, hipStream_t s);
  void results(int k, std::vector<double> &alphas, std::vector<double> &betas,
               std::vector<int> &steps, std::vector<double> &bnorms, hipStream_t s) const;
  std::vector<double> ground_energies(int k, bool use_graph, hipStream_t s);
  // time k steps (ms) for the current mapping
  double time_steps(int k, hipStream_t s);
  // try tile x block candidates on k_probe steps; keep and return the fastest
  std::pair<int, int> autotune(int k_probe, hipStream_t s);

  // launch helpers (used by the templates in batch.hip)
  void hv_step(const double *X, double *V, const double *VP, double *W, int k, int step,
               hipStream_t s) const;
  void vec_step(double *V, double *VP, double *W, int k, int step, int mode, hipStream_t s) const;
  void reduce(int k, int step, int what, hipStream_t s) const;
  // every step of every small instance, one workgroup each (after the reset)
  void persistent_run(int k, hipStream_t s) const;
  // instances up to max_dim go to the persistent kernel (0: none); re-tiles
  void set_persist_max_dim(int64_t max_dim);

  std::vector<const Hubbard *> models_;
  std::vector<BInst> host_inst_;
  std::vector<BTile> tiles_;
  // B_: the start vectors (never overwritten); W_, V_, VP_: the steps'
  // unnormalised outputs rotate through them (the persistent kernel uses the
  // same buffers in its own instances' ranges)
  DevBuf d_inst_, d_tiles_, V_, VP_, W_, B_, partial_, alpha_, beta_, done_, steps_, bnorm_;
  DevBuf d_small_;      // indices of the persistent (small) instances
  int n_small_ = 0;
  hipStream_t side_ = nullptr;  // the persistent kernel's branch of a run
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  int64_t persist_max_dim_ = kPersistMaxDim;
  int64_t total_ = 0;
  size_t lds_bytes_ = 0;
  int tile_ = 1024, block_ = 256;
  int kcap_ = 0;
  int graph_k_ = -1;
  hipGraphExec_t graph_exec_ = nullptr;

 private:
  void ensure_k(int k);
  void record_run(int k, hipStream_t s);
};

}  // namespace mb
