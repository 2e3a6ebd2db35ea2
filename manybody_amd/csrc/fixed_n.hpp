// Basis of all states with N particles on L orbitals, any N_up
// (Basis::with_fixed_particle_number, basis.cpp:70-89, 240-265), see fixed_n.cpp.
#pragma once

#include <vector>

#include "mb_internal.hpp"

namespace mb {

// Device order: blocks k = N_up from k_lo to k_hi, block k = (up strings with k
// particles) x (down strings with N - k), row-major, planar complex; strings in
// colex order inside each popcount group.  Reference order: the reference's
// depth-first enumeration.
struct FixedNBasis {
  int L = 0, N = 0, k_lo = 0, k_hi = -1;
  std::vector<uint32_t> up_str, dn_str;  // grouped by popcount 0 .. L
  std::vector<int32_t> up_off, dn_off;   // L + 2 group offsets
  std::vector<int64_t> blk_off;          // L + 2 element offsets of the blocks (by N_up)
  std::vector<int32_t> ncols;            // L + 1: columns of block k = C(L, N - k)
  std::vector<int64_t> ref_of_pos, pos_of_ref;
  int64_t D = 0;
  DevBuf d_ref_of_pos;
  int device = -1;

  // device position of (up, dn), or -1 outside the basis
  int64_t pos_of(uint32_t up, uint32_t dn) const;
  void ref_state(int64_t i, uint32_t *up, uint32_t *dn) const;
  int64_t ref_index(uint32_t up, uint32_t dn) const;  // -1 outside the basis
  void ensure_device();
};

void build_fixed_n_basis(FixedNBasis &b, int L, int N);

}  // namespace mb
