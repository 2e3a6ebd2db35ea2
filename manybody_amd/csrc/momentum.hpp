// Fixed-total-momentum sectors and the factorized momentum-space Hubbard
// operator (see momentum.cpp).
#pragma once

#include <vector>

#include "krylov.hpp"
#include "mb_internal.hpp"

namespace mb {

// Basis of (N_up, N_dn, total momentum K) on a periodic lattice, orbital k =
// the momentum k_x + L_x k_y + L_x L_y k_z
// (Basis::with_fixed_particle_number_spin_momentum, basis.cpp:128-185, 304-356).
// Device order: blocks p = 0 .. N-1 of the up strings with momentum p times the
// down strings with momentum K - p, each block row-major; within a momentum
// group the strings are in colex order.  Reference order: the real-space
// DFS order restricted to the sector.
struct MomentumSector {
  int L = 0, n_up = 0, n_dn = 0, ndim = 0, K = 0;
  int dims[3] = {1, 1, 1};
  std::vector<uint32_t> up_str, dn_str;  // grouped by momentum
  std::vector<int32_t> up_off, dn_off;   // N + 1 group offsets
  std::vector<int64_t> blk_off;          // N + 1 element offsets of the blocks
  std::vector<int32_t> ncols;            // N: columns of block p = |dn group K - p|
  std::vector<int64_t> ref_of_pos;       // device position -> reference index
  std::vector<int64_t> pos_of_ref;       // reference index -> device position
  int64_t D = 0;
  Sector real;                           // the real-space sector (reference ranks)
  DevBuf d_ref_of_pos, d_pos_of_ref;
  int device = -1;

  int msub(int p, int q) const;  // p - q, component-wise periodic (flat)
  int madd(int p, int q) const;  // p + q
  int momentum_of(uint32_t bits) const;
  // (up, dn) of reference index i; reference index of (up, dn) or -1
  void ref_state(int64_t i, uint32_t *up, uint32_t *dn) const;
  int64_t ref_index(uint32_t up, uint32_t dn) const;
  void ensure_device();
};

void build_momentum_sector(MomentumSector &m, int ndim, const int *sizes, int n_up, int n_dn,
                           const int *momentum);

// Device tables of the factorized operator
//   H = t sum_{k,s} eps(k) n_{k,s} + (U/N) sum_q rho_{q,up} rho_{-q,dn}
// (hubbard_momentum_factorized_fixed_k.h:92-117): per output up string and
// q != 0 its source strings under rho_{q,up}, per output down string and q its
// source strings under rho_{-q,dn}, each as a 15-bit index inside the source
// momentum group with the fermionic sign in bit 15; q = 0 is N_up N_dn.  Lists
// of at most 8 entries are padded to 8 (one 16-byte load per string and q).
struct MomTables {
  int N, K, mu, md;  // sites, total momentum, strides of the source lists
  int64_t D, nrows;  // dimension, up strings (output rows)
  double t, U, q0;                          // q0 = (U/N) N_up N_dn, the q = 0 term
  const double *eps_up, *eps_dn;            // sum of eps(k) over each string
  const uint16_t *up_src, *dn_src;          // [string][q][mu or md]
  const uint8_t *up_cnt, *dn_cnt;           // [string][q]
  const int32_t *up_grp, *up_loc;           // momentum group and index inside it
  const int32_t *up_off;                    // N + 1 up group offsets (block rows)
  const int32_t *dn_off, *ncols, *msub;     // group offsets, block columns, p - q table
  const int64_t *blk_off;
  int max_cols, max_rows;                   // largest down / up momentum group
};

constexpr int kMomLockMaxW = 64;  // workgroups per XCD of the persistent momentum kernel

class MomentumHubbard final : public HvOp {
 public:
  MomentumHubbard(const MomentumSector &m, double t, double U);
  int64_t dim() const override { return m_.D; }
  // per plane one slot per output row (per-row kernel) or per (block, workgroup)
  // of the persistent kernel, at most kMomLockMaxW workgroups per XCD
  int64_t partials() const override {
    return 2 * ((int64_t)m_.up_str.size() + (int64_t)m_.L * kMomLockMaxW) + 16;
  }
  int64_t hv(const double *x, double *y, Epi e, const EpiArgs &a, hipStream_t s) const override;
  int64_t hv(const float *x, float *y, Epi e, const EpiArgs &a, hipStream_t s) const override;
  const MomentumSector &sector() const { return m_; }
  MomTables tables() const { return tb_; }

 private:
  const MomentumSector &m_;
  MomTables tb_{};
  DevBuf d_eps_up_, d_eps_dn_, d_up_src_, d_dn_src_, d_up_cnt_, d_dn_cnt_, d_up_grp_, d_up_loc_,
      d_up_off_, d_dn_off_, d_ncols_, d_msub_, d_blk_off_;
};

// kernels (kernels.hip)
int64_t launch_mom_hv(const MomTables &t, const double *x, double *y, Epi epi, const EpiArgs &a,
                      hipStream_t s);
void launch_mom_perm(const int64_t *map, int64_t D, const double *src, double *dst, bool to_device,
                     hipStream_t s);
void launch_mom_fill_gaussian(const int64_t *ref_of_pos, int64_t D, uint64_t seed, double *x,
                              hipStream_t s);

}  // namespace mb
