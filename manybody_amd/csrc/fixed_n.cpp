// Fixed-particle-number basis (any N_up): all states with N particles on L
// orbitals.  The reference enumerates them depth first over the modes
// (spin, orbital): at each level the up candidates from the current orbital
// on, then the down ones, a mode allowed when its orbital is above the last
// mode's or equal with a higher spin (basis.cpp:240-265); its index is the
// position in that enumeration.  Here the enumeration runs once on the host and
// records, for each reference index, the device position in blocks of equal
// N_up (block k = up strings of k particles x down strings of N - k, each in
// colex order).  The Hamiltonian of the Hubbard family is block diagonal in
// that layout; spin-flip terms move between neighbouring blocks.
#include "fixed_n.hpp"

#include <algorithm>
#include <mutex>
#include <stdexcept>

namespace mb {

namespace {

// strings of L bits with k set, ascending (Gosper's hack)
void strings_with(int L, int k, std::vector<uint32_t> &out) {
  if (k == 0) {
    out.push_back(0u);
    return;
  }
  const uint64_t end = uint64_t{1} << L;
  for (uint64_t v = (uint64_t{1} << k) - 1; v < end;) {
    out.push_back((uint32_t)v);
    const uint64_t c = v & (~v + 1), r = v + c;
    v = (((r ^ v) >> 2) / c) | r;
  }
}

}  // namespace

int64_t FixedNBasis::pos_of(uint32_t up, uint32_t dn) const {
  const uint64_t lim = uint64_t{1} << L;
  if ((uint64_t)up >= lim || (uint64_t)dn >= lim) return -1;
  const int k = __builtin_popcount(up);
  if (k + __builtin_popcount(dn) != N || k < k_lo || k > k_hi) return -1;
  return blk_off[k] + Sector::colex_rank(up) * ncols[k] + Sector::colex_rank(dn);
}

int64_t FixedNBasis::ref_index(uint32_t up, uint32_t dn) const {
  const int64_t p = pos_of(up, dn);
  return p < 0 ? -1 : ref_of_pos[p];
}

void FixedNBasis::ref_state(int64_t i, uint32_t *up, uint32_t *dn) const {
  if (i < 0 || i >= D) throw std::out_of_range("basis index out of range");
  const int64_t p = pos_of_ref[i];
  int k = k_lo;
  while (k < k_hi && blk_off[k + 1] <= p) ++k;
  const int64_t r = p - blk_off[k];
  *up = up_str[up_off[k] + r / ncols[k]];
  *dn = dn_str[dn_off[N - k] + r % ncols[k]];
}

void FixedNBasis::ensure_device() {
  static std::mutex mu;  // the first upload is serialised across host threads
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  MB_HIP(hipGetDevice(&dev));
  if (device == dev) return;
  d_ref_of_pos = upload(ref_of_pos);
  device = dev;
}

void build_fixed_n_basis(FixedNBasis &b, int L, int N) {
  if (L < 1 || L > kMaxSites) throw std::invalid_argument("orbitals must be in [1, 32]");
  if (N < 0 || N > 2 * L) throw std::invalid_argument("particles must be in [0, 2 orbitals]");
  b = FixedNBasis{};
  b.L = L;
  b.N = N;
  b.k_lo = std::max(0, N - L);
  b.k_hi = std::min(N, L);
  const Binom &C = binom();
  int64_t D = 0;
  for (int k = b.k_lo; k <= b.k_hi; ++k) D += (int64_t)(C(L, k) * C(L, N - k));
  if (D >= (int64_t{1} << 31)) throw std::invalid_argument("basis too large");
  b.D = D;
  // strings grouped by popcount (only the groups the blocks use)
  b.up_off.assign(L + 2, 0);
  b.dn_off.assign(L + 2, 0);
  for (int g = 0; g <= L; ++g) {
    b.up_off[g] = (int32_t)b.up_str.size();
    if (g >= b.k_lo && g <= b.k_hi) strings_with(L, g, b.up_str);
    b.dn_off[g] = (int32_t)b.dn_str.size();
    if (N - g >= b.k_lo && N - g <= b.k_hi) strings_with(L, g, b.dn_str);
  }
  b.up_off[L + 1] = (int32_t)b.up_str.size();
  b.dn_off[L + 1] = (int32_t)b.dn_str.size();
  b.blk_off.assign(L + 2, 0);
  b.ncols.assign(L + 1, 0);
  int64_t off = 0;
  for (int k = 0; k <= L; ++k) {
    b.blk_off[k] = off;
    if (k >= b.k_lo && k <= b.k_hi) {
      b.ncols[k] = (int32_t)C(L, N - k);
      off += (int64_t)C(L, k) * b.ncols[k];
    }
  }
  b.blk_off[L + 1] = off;
  // the reference's depth-first enumeration (dead branches pruned: the order of
  // the leaves is unchanged)
  b.ref_of_pos.assign(D, -1);
  b.pos_of_ref.assign(D, -1);
  int64_t idx = 0;
  std::vector<int> ms, mo;  // spin, orbital of the chosen modes
  ms.reserve(N);
  mo.reserve(N);
  uint32_t up = 0, dn = 0;
  auto rec = [&](auto &&self, int first) -> void {
    const int size = (int)ms.size();
    if (size == N) {
      const int64_t p = b.pos_of(up, dn);
      b.ref_of_pos[p] = idx;
      b.pos_of_ref[idx] = p;
      ++idx;
      return;
    }
    for (int s = 0; s < 2; ++s) {
      for (int i = first; i < L; ++i) {
        if (size > 0 && !(mo.back() < i || (mo.back() == i && s > ms.back()))) continue;
        // modes after (s, i): (i, down) if s is up, then 2 per later orbital
        const int after = 2 * (L - 1 - i) + (s == 0 ? 1 : 0);
        if (after < N - size - 1) continue;
        ms.push_back(s);
        mo.push_back(i);
        (s == 0 ? up : dn) |= 1u << i;
        self(self, i);
        (s == 0 ? up : dn) &= ~(1u << i);
        ms.pop_back();
        mo.pop_back();
      }
    }
  };
  rec(rec, 0);
  if (idx != D) throw std::runtime_error("fixed-N enumeration count mismatch");
}

}  // namespace mb
