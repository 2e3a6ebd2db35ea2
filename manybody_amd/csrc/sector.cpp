// Fixed-(N_up, N_dn) Fock sector: species strings in combinatorial rank order
// and the closed-form map to the reference's basis index.
//
// Reference: FermionBasis::with_fixed_particle_number_and_spin
// (src/algebra/basis.cpp:91-126) enumerates states by a depth-first search
// (basis.cpp:267-302) and indexes them by DFS position (IndexedHashSet,
// src/utils/indexed_hash_set.h:82-106).  Here the DFS position is computed by
// counting, without enumerating or hashing: the DFS appends modes in increasing
// (orbital, spin) order and, at each level, tries all up-spin candidates before
// all down-spin candidates, so a state's index is the number of leaves in the
// subtrees of the candidates tried before each of its own modes.
#include <algorithm>

#include "mb_internal.hpp"

#include <mutex>

namespace mb {

DevAlloc *&thread_dev_alloc() {
  static thread_local DevAlloc *a = nullptr;
  return a;
}

hipStream_t &thread_upload_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

void DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 1;
  if (DevAlloc *a = thread_dev_alloc()) {
    p = a->alloc(n);
    if (!p) throw NoMemError("request region exhausted (" + std::to_string(n) + " bytes)");
    owner = a;
  } else if (hipMalloc(&p, n) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    throw NoMemError("hipMalloc of " + std::to_string(n) + " bytes failed");
  }
  bytes = n;
}

void DevBuf::release() {
  if (p) {
    if (owner)
      owner->free(p);
    else
      (void)hipFree(p);
  }
  p = nullptr;
  bytes = 0;
  owner = nullptr;
}

Binom::Binom() {
  for (int n = 0; n <= kMaxSites; ++n)
    for (int k = 0; k <= kMaxSites; ++k) c[n][k] = 0;
  for (int n = 0; n <= kMaxSites; ++n) {
    c[n][0] = 1;
    for (int k = 1; k <= n; ++k) c[n][k] = c[n - 1][k - 1] + (k <= n - 1 ? c[n - 1][k] : 0);
  }
}

void build_sector(Sector &s, int L, int n_up, int n_dn) {
  if (L < 1 || L > kMaxSites) throw std::invalid_argument("orbitals must be in [1, 32]");
  if (n_up < 0 || n_dn < 0 || n_up > L || n_dn > L)
    throw std::invalid_argument("spin sector exceeds the orbital count");
  s.L = L;
  s.n_up = n_up;
  s.n_dn = n_dn;
  const Binom &C = binom();
  s.d_up = (int64_t)C(L, n_up);
  s.d_dn = (int64_t)C(L, n_dn);
  if (s.d_up * s.d_dn >= ((int64_t)1 << 40)) throw std::invalid_argument("basis too large");
  if (s.d_up >= ((int64_t)1 << 31) || s.d_dn >= ((int64_t)1 << 31))
    throw std::invalid_argument("species sector too large");
  for (int n = 0; n < 2; ++n) {
    const int k = n == 0 ? n_up : n_dn;
    auto &v = n == 0 ? s.s_up : s.s_dn;
    v.clear();
    v.reserve(n == 0 ? s.d_up : s.d_dn);
    // ascending bitstrings with popcount k (Gosper's hack)
    if (k == 0) {
      v.push_back(0u);
    } else {
      uint64_t x = (1ull << k) - 1;
      while (x < (1ull << L)) {
        v.push_back((uint32_t)x);
        const uint64_t c = x & (~x + 1), r = x + c;
        x = (((r ^ x) >> 2) / c) | r;
      }
    }
  }
}

const Binom &binom() {
  static const Binom b;
  return b;
}

int64_t Sector::colex_rank(uint32_t bits) {
  const Binom &C = binom();
  int64_t r = 0;
  int k = 0;
  while (bits) {
    const int p = __builtin_ctz(bits);
    r += (int64_t)C(p, k + 1);
    ++k;
    bits &= bits - 1;
  }
  return r;
}

namespace {

// Leaves below a just-placed up mode at orbital o with remaining budgets (u, d):
// up modes above it are orbitals o+1.., down modes above it are orbitals o..
inline uint64_t count_after_up(const Binom &C, int L, int o, int u, int d) {
  return C(L - 1 - o, u) * C(L - o, d);
}
// ... and below a just-placed down mode at orbital o: both species use o+1..
inline uint64_t count_after_dn(const Binom &C, int L, int o, int u, int d) {
  return C(L - 1 - o, u) * C(L - 1 - o, d);
}

}  // namespace

int64_t Sector::ref_index(uint32_t up, uint32_t dn) const {
  const Binom &C = binom();
  int64_t rank = 0;
  int U = n_up, D = n_dn;
  int prev_o = -1, prev_s = -1;  // last placed mode (none yet)
  for (int o = 0; o < L; ++o) {
    for (int s = 0; s < 2; ++s) {
      const bool occ = s == 0 ? (up >> o & 1u) : (dn >> o & 1u);
      if (!occ) continue;
      // up candidates: orbital > prev_o (any orbital at the first level)
      if (U > 0) {
        const int lo = prev_o + 1;
        const int hi = (s == 0) ? o - 1 : L - 1;
        for (int i = lo; i <= hi; ++i) rank += (int64_t)count_after_up(C, L, i, U - 1, D);
      }
      // down candidates tried before a down mode: orbital >= prev_o when the
      // last mode was up at prev_o, > prev_o otherwise
      if (s == 1 && D > 0) {
        const int lo = (prev_s == 0) ? prev_o : prev_o + 1;
        for (int i = std::max(lo, 0); i < o; ++i) rank += (int64_t)count_after_dn(C, L, i, U, D - 1);
      }
      if (s == 0) --U; else --D;
      prev_o = o;
      prev_s = s;
    }
  }
  return rank;
}

void Sector::ref_state(int64_t index, uint32_t *up_out, uint32_t *dn_out) const {
  const Binom &C = binom();
  uint32_t up = 0, dn = 0;
  int U = n_up, D = n_dn;
  int prev_o = -1, prev_s = -1;
  int64_t rem = index;
  while (U > 0 || D > 0) {
    bool placed = false;
    // children in DFS order: up candidates ascending, then down candidates
    if (U > 0) {
      for (int i = prev_o + 1; i < L && !placed; ++i) {
        const int64_t c = (int64_t)count_after_up(C, L, i, U - 1, D);
        if (rem < c) {
          up |= 1u << i; --U; prev_o = i; prev_s = 0; placed = true;
        } else {
          rem -= c;
        }
      }
    }
    if (!placed && D > 0) {
      const int lo = (prev_s == 0) ? prev_o : prev_o + 1;
      for (int i = std::max(lo, 0); i < L && !placed; ++i) {
        const int64_t c = (int64_t)count_after_dn(C, L, i, U, D - 1);
        if (rem < c) {
          dn |= 1u << i; --D; prev_o = i; prev_s = 1; placed = true;
        } else {
          rem -= c;
        }
      }
    }
    if (!placed) throw RangeError("basis index out of range");
  }
  *up_out = up;
  *dn_out = dn;
}

void Sector::ensure_device() {
  // operators share a basis across host threads: the first upload is serialised
  static std::mutex m;
  std::lock_guard<std::mutex> lock(m);
  int dev = 0;
  MB_HIP(hipGetDevice(&dev));
  if (device == dev && d_s_up.p) return;
  d_s_up = upload(s_up);
  d_s_dn = upload(s_dn);
  std::vector<uint64_t> bt((kMaxSites + 1) * (kMaxSites + 1));
  const Binom &C = binom();
  for (int n = 0; n <= kMaxSites; ++n)
    for (int k = 0; k <= kMaxSites; ++k) bt[n * (kMaxSites + 1) + k] = C.c[n][k];
  d_binom = upload(bt);
  device = dev;
}

PermTables perm_of(const Sector &s) {
  PermTables p;
  p.L = s.L;
  p.n_up = s.n_up;
  p.n_dn = s.n_dn;
  p.d_up = s.d_up;
  p.d_dn = s.d_dn;
  p.s_up = s.d_s_up.as<uint32_t>();
  p.s_dn = s.d_s_dn.as<uint32_t>();
  p.binom = s.d_binom.as<uint64_t>();
  return p;
}

}  // namespace mb
