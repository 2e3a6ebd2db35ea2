// Hubbard model tables: bonds in the reference's loop order, merged hopping
// monomials, per-species hop tables with fermionic signs, and the interaction
// table.
//
// Reference: HubbardModel / HubbardModel2D / HubbardModel3D
// (src/algebra/fermion/model/hubbard_model.h:10-212) build
//   H = sum_bonds sum_spin -t (c+_i c_j + c+_j c_i) + U sum_i n_i,up n_i,dn
// as a FermionExpression, where identical monomials merge by coefficient
// addition (ExpressionBase::add_to, expression_base.h:85-97): a linear size of
// 2 yields -2t on the repeated pair, a linear size of 1 yields the self bond
// c+_i c_i + c+_i c_i = 2 n_i (hopping_generic, expression_base.h:394-403).
//
// Sign: a basis key orders all up creators (ascending orbital) before all down
// creators (operator.h:69-79), so moving c_b to its partner and inserting c+_a
// in canonical position passes, in total, the same-spin creators strictly
// between a and b: <f| c+_a c_b |s> = (-1)^popcount(s & between(a, b)).  The
// up-spin block in front of the down creators is passed twice by a down hop
// and drops out, so each species' sign depends on its own string only.
#include <algorithm>
#include <cmath>

#include "mb_internal.hpp"

namespace mb {

namespace {

struct Mono {
  int a, b;  // coefficient * c+_a c_b
  double c;
};

std::vector<Mono> merge_monomials(const std::vector<int> &from, const std::vector<int> &to,
                                  double t) {
  std::vector<Mono> m;
  const double coeff = -t;
  for (size_t k = 0; k < from.size(); ++k) {
    const int pairs[2][2] = {{from[k], to[k]}, {to[k], from[k]}};
    for (const auto &p : pairs) {
      auto it = std::find_if(m.begin(), m.end(),
                             [&](const Mono &x) { return x.a == p[0] && x.b == p[1]; });
      if (it != m.end())
        it->c += coeff;
      else
        m.push_back({p[0], p[1], coeff});
    }
  }
  // ExpressionBase::is_zero (expression_base.h:79-83)
  m.erase(std::remove_if(m.begin(), m.end(), [](const Mono &x) { return x.c * x.c < kTol * kTol; }),
          m.end());
  return m;
}

void species_hops(const std::vector<uint32_t> &strings, const std::vector<Mono> &mono,
                  SpeciesHops &out) {
  const size_t n = strings.size();
  out.row_ptr.assign(n + 1, 0);
  out.tgt.clear();
  out.amp.clear();
  out.onsite.assign(n, 0.0);
  out.max_row = 0;
  for (size_t r = 0; r < n; ++r) {
    const uint32_t s = strings[r];
    double onsite = 0.0;
    for (const Mono &m : mono) {
      const uint32_t ba = 1u << m.a, bb = 1u << m.b;
      if (!(s & bb)) continue;
      if (m.a == m.b) {  // c+_i c_i = n_i
        onsite += m.c;
        continue;
      }
      if (s & ba) continue;
      const int lo = std::min(m.a, m.b), hi = std::max(m.a, m.b);
      const uint32_t between = (hi - lo > 1) ? (((1u << hi) - 1u) & ~((2u << lo) - 1u)) : 0u;
      const int sign = (__builtin_popcount(s & between) & 1) ? -1 : 1;
      const uint32_t f = s ^ ba ^ bb;
      out.tgt.push_back((int32_t)Sector::colex_rank(f));
      out.amp.push_back(sign * m.c);
    }
    out.onsite[r] = onsite;
    out.row_ptr[r + 1] = (int32_t)out.tgt.size();
    out.max_row = std::max(out.max_row, out.row_ptr[r + 1] - out.row_ptr[r]);
  }
}

// Position of entry e (< 32) of string or slot d in a compact table: groups
// of 64 strings, each [chunk of 8 entries][string][8], so that a wave reads
// each chunk of its 64 strings as one contiguous KB.
size_t comp_pos(size_t d, int e) {
  return (d / 64) * 2048 + (size_t)(e / 8) * 512 + (d % 64) * 8 + (size_t)(e % 8);
}

// Orders the targets of one sign class of a wave's 64 slots over the columns
// [0, C) so that in a column the lanes of each half-wave read distinct LDS
// bank pairs where the targets allow it: the LDS serves a wave's 8-byte reads
// as two halves of 32 lanes over 32 bank pairs (bank pair = element index mod
// 32), and a column costs its largest number of distinct elements on one pair.
// Per half-wave this is an edge colouring of the bipartite multigraph lanes x
// residues with C colours (a residue with more than C targets is split into
// copies of at most C, which may then share a column); Koenig's alternating
// path recolouring finds one.  Columns a lane has no target for get the pad.
// Measured on the 4 x 4, 7 + 7 sector: 2.0 elements per bank pair and column
// instead of 3.4 in hop order.
void colour_columns(const std::vector<std::vector<uint16_t>> &lists, int C, uint16_t pad,
                    std::vector<uint16_t> &out) {
  out.assign(64 * (size_t)C, pad);
  if (C == 0) return;
  struct Edge {
    int lane, rv;
    uint16_t t;
    int col;
  };
  constexpr int kCopies = 32;
  std::vector<int> lane_c(32 * (size_t)C), rv_c(32 * kCopies * (size_t)C);
  for (int h0 = 0; h0 < 64; h0 += 32) {
    std::vector<Edge> edges;
    int cnt[32] = {0};
    for (int i = 0; i < 32; ++i)
      for (uint16_t t : lists[h0 + i]) {
        const int r = t % 32, k = cnt[r]++ / C;
        edges.push_back({i, r * kCopies + k, t, -1});
      }
    std::fill(lane_c.begin(), lane_c.end(), -1);
    std::fill(rv_c.begin(), rv_c.end(), -1);
    auto at = [&](bool lane_side, int x, int c) -> int & {
      return lane_side ? lane_c[(size_t)x * C + c] : rv_c[(size_t)x * C + c];
    };
    for (size_t e = 0; e < edges.size(); ++e) {
      const int u = edges[e].lane, v = edges[e].rv;
      int a = 0, b = 0;
      while (at(true, u, a) >= 0) ++a;
      while (at(false, v, b) >= 0) ++b;
      int c = a;
      if (at(false, v, a) >= 0) {
        if (at(true, u, b) < 0) {
          c = b;
        } else {  // swap a and b along the path from v that starts with colour a
          std::vector<int> path;
          bool lane_side = false;
          int x = v, want = a;
          while (at(lane_side, x, want) >= 0) {
            const int f = at(lane_side, x, want);
            path.push_back(f);
            x = lane_side ? edges[f].rv : edges[f].lane;
            lane_side = !lane_side;
            want = want == a ? b : a;
          }
          for (int f : path) {
            at(true, edges[f].lane, edges[f].col) = -1;
            at(false, edges[f].rv, edges[f].col) = -1;
          }
          for (int f : path) {
            edges[f].col = edges[f].col == a ? b : a;
            at(true, edges[f].lane, edges[f].col) = f;
            at(false, edges[f].rv, edges[f].col) = f;
          }
        }
      }
      edges[e].col = c;
      at(true, u, c) = (int)e;
      at(false, v, c) = (int)e;
    }
    for (const Edge &e : edges) out[(size_t)(h0 + e.lane) * C + e.col] = e.t;
  }
}

// The compact table in sorted slot order with one sign pattern per wave.
// Strings are sorted (stable) by hop count, then by positive hop count, cut
// into groups of 64 (one wave), and the groups of each round of kSortedRound
// slots are dealt to the waves in serpentine order (round j gives wave w the
// sorted group 16 j + w, or 16 j + 15 - w for odd j), which evens out the
// total work per wave.  Per group P = the most positive and N = the most
// negative hops of its strings; a slot lists its positive targets first.
void build_sorted(SpeciesHops &h) {
  const size_t n = h.row_ptr.size() - 1;
  h.srt_ok = false;
  h.sperm.clear();
  h.scomp.clear();
  h.spn.clear();
  const size_t np = (n + kSortedRound - 1) / kSortedRound * kSortedRound;
  if (!h.comp_ok || n < (size_t)kSortedRound || np > (size_t)kSortedRound * kSortedRounds) return;
  std::vector<int> cpos(n, 0), cneg(n, 0);
  for (size_t d = 0; d < n; ++d)
    for (int e = h.row_ptr[d]; e < h.row_ptr[d + 1]; ++e) (h.amp[e] < 0 ? cneg : cpos)[d]++;
  std::vector<uint32_t> order(n);
  for (size_t d = 0; d < n; ++d) order[d] = (uint32_t)d;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    const int ta = cpos[a] + cneg[a], tb = cpos[b] + cneg[b];
    return ta != tb ? ta < tb : cpos[a] < cpos[b];
  });
  order.resize(np, (uint32_t)n);  // pad slots at the end
  constexpr int kGroups = kSortedRound / 64;
  h.sperm.assign(np, (uint16_t)n);
  h.scomp.assign(np * 32, (uint16_t)n);
  h.spn.assign(np / 64, 0);
  for (size_t j = 0; j < np / kSortedRound; ++j)
    for (int w = 0; w < kGroups; ++w) {
      const size_t src = j * kGroups + ((j & 1) ? kGroups - 1 - w : w);
      int P = 0, N = 0;
      for (int l = 0; l < 64; ++l) {
        const uint32_t d = order[src * 64 + l];
        if (d < n) {
          P = std::max(P, cpos[d]);
          N = std::max(N, cneg[d]);
        }
      }
      if (P + N > 32) return;  // srt_ok stays false
      std::vector<std::vector<uint16_t>> lp(64), ln(64);
      for (int l = 0; l < 64; ++l) {
        const size_t slot = j * kSortedRound + (size_t)w * 64 + l;
        const uint32_t d = order[src * 64 + l];
        h.sperm[slot] = (uint16_t)d;
        if (d >= n) continue;
        for (int e = h.row_ptr[d]; e < h.row_ptr[d + 1]; ++e)
          (h.amp[e] < 0 ? ln : lp)[l].push_back((uint16_t)h.tgt[e]);
      }
      std::vector<uint16_t> cp_, cn_;
      colour_columns(lp, P, (uint16_t)n, cp_);
      colour_columns(ln, N, (uint16_t)n, cn_);
      for (int l = 0; l < 64; ++l) {
        const size_t slot = j * kSortedRound + (size_t)w * 64 + l;
        for (int k = 0; k < P; ++k) h.scomp[comp_pos(slot, k)] = cp_[(size_t)l * P + k];
        for (int k = 0; k < N; ++k) h.scomp[comp_pos(slot, P + k)] = cn_[(size_t)l * N + k];
      }
      h.spn[(j * kSortedRound + (size_t)w * 64) / 64] = (uint16_t)(P | N << 8);
    }
  h.srt_np = (int64_t)np;
  h.srt_ok = true;
}

// Column-major packed copy of the hop table for the row kernel: entry k of
// string d at ell[k * D + d] = target << 8 | (amplitude < 0) << 7 | magnitude
// index.  Needs D < 2^24 and at most 128 distinct magnitudes.
void pack_ell(SpeciesHops &h) {
  const size_t n = h.row_ptr.size() - 1;
  h.ell_ok = false;
  h.mags.clear();
  if (n >= (1u << 24) || h.max_row > 255) return;
  std::vector<uint32_t> code(h.tgt.size());
  for (size_t e = 0; e < h.tgt.size(); ++e) {
    const double m = std::fabs(h.amp[e]);
    auto it = std::find(h.mags.begin(), h.mags.end(), m);
    int idx = (int)(it - h.mags.begin());
    if (it == h.mags.end()) {
      if (h.mags.size() >= 128) return;
      h.mags.push_back(m);
    }
    code[e] = ((uint32_t)h.tgt[e] << 8) | (h.amp[e] < 0 ? 0x80u : 0u) | (uint32_t)idx;
  }
  h.ell_k = h.max_row;
  h.ell.assign((size_t)std::max(h.ell_k, 1) * n, 0u);
  h.cnt.assign(n, 0);
  for (size_t d = 0; d < n; ++d) {
    const int c = h.row_ptr[d + 1] - h.row_ptr[d];
    h.cnt[d] = (uint8_t)c;
    for (int k = 0; k < c; ++k) h.ell[(size_t)k * n + d] = code[h.row_ptr[d] + k];
  }
  if (h.mags.empty()) h.mags.push_back(0.0);
  h.ell_ok = true;
  // compact copy: one magnitude, at most 32 entries per string, indices up to
  // 65535 (the zero slot n included).  Strings padded to a multiple of 64 (one
  // wave); unused slots point at the zero slot n; bit k of sign[d] is the sign
  // of slot k; gmax[g] is the largest hop count among strings 64 g .. 64 g + 63.
  h.comp_ok = h.mags.size() == 1 && n <= 65535 && h.max_row <= 32;
  h.comp.clear();
  h.csign.clear();
  h.cgmax.clear();
  if (h.comp_ok) {
    const size_t np = (n + 63) / 64 * 64;
    h.comp.assign(np * 32, (uint16_t)n);
    h.csign.assign(np, 0u);
    h.cgmax.assign(np / 64, 0);
    for (size_t d = 0; d < n; ++d) {
      const int c = h.row_ptr[d + 1] - h.row_ptr[d];
      for (int k = 0; k < c; ++k) {
        const int e = h.row_ptr[d] + k;
        h.comp[comp_pos(d, k)] = (uint16_t)h.tgt[e];
        if (h.amp[e] < 0) h.csign[d] |= 1u << k;
      }
      h.cgmax[d / 64] = std::max<uint8_t>(h.cgmax[d / 64], (uint8_t)c);
    }
  }
  build_sorted(h);
}

// Blocks of up strings for the block-staged up kernel.  Strings are in colex
// (numeric) order, so the strings that agree from bit k upward are a
// contiguous run of rows, and a hop between two sites below k stays inside
// the run.  For each k the run sizes bound the LDS a workgroup needs (256
// bytes of staged strip and 4 bytes of metadata per row, 2 bytes per hop
// entry); the k with the most in-block hops within kUpBlockLds wins.  Needs
// fewer than 2^15 strings, at most 64 hops per string, and a fifth of the hops
// inside blocks (otherwise the per-row strip kernel stays).
void build_up_blocks(const std::vector<uint32_t> &strings, int L, SpeciesHops &h) {
  h.blk_ok = false;
  h.blk.clear();
  h.nin.clear();
  const size_t n = strings.size();
  if (n == 0 || n >= 32768 || h.max_row > 64 || h.tgt.empty()) return;
  auto runs = [&](int k, std::vector<std::pair<int, int>> *out, int &maxr, int &maxe, int64_t &in) {
    maxr = maxe = 0;
    in = 0;
    size_t r = 0;
    while (r < n) {
      size_t r1 = r + 1;
      while (r1 < n && ((uint64_t)strings[r1] >> k) == ((uint64_t)strings[r] >> k)) ++r1;
      maxr = std::max(maxr, (int)(r1 - r));
      maxe = std::max(maxe, h.row_ptr[r1] - h.row_ptr[r]);
      for (int e = h.row_ptr[r]; e < h.row_ptr[r1]; ++e)
        in += (h.tgt[e] >= (int32_t)r && h.tgt[e] < (int32_t)r1);
      if (out) out->push_back({(int)r, (int)(r1 - r)});
      r = r1;
    }
  };
  int best_k = -1;
  int64_t best_in = 0;
  for (int k = 1; k <= L; ++k) {
    int maxr, maxe;
    int64_t in;
    runs(k, nullptr, maxr, maxe, in);
    const int64_t bytes = (int64_t)maxr * (kUpBlockRowBytes + 4) + (int64_t)maxe * 2;
    if (bytes > kUpBlockLds || maxe > 65535) continue;
    if (in > best_in) {
      best_in = in;
      best_k = k;
    }
  }
  if (best_k < 0 || best_in * 5 < (int64_t)h.tgt.size()) return;
  std::vector<std::pair<int, int>> blocks;
  int64_t in;
  runs(best_k, &blocks, h.blk_rows, h.blk_ent, in);
  // out-of-block entries first, in-block after, each part in its old order
  h.nin.assign(n, 0);
  std::vector<int32_t> tgt(h.tgt.size());
  std::vector<double> amp(h.amp.size());
  for (const auto &b : blocks)
    for (int r = b.first; r < b.first + b.second; ++r) {
      int w = h.row_ptr[r];
      for (int pass = 0; pass < 2; ++pass)
        for (int e = h.row_ptr[r]; e < h.row_ptr[r + 1]; ++e) {
          const bool inside = h.tgt[e] >= b.first && h.tgt[e] < b.first + b.second;
          if (inside != (pass == 1)) continue;
          tgt[w] = h.tgt[e];
          amp[w] = h.amp[e];
          ++w;
          h.nin[r] += (uint8_t)inside;
        }
    }
  h.tgt.swap(tgt);
  h.amp.swap(amp);
  std::stable_sort(blocks.begin(), blocks.end(),
                   [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
                     return a.second > b.second;
                   });
  for (const auto &b : blocks) {
    h.blk.push_back(b.first);
    h.blk.push_back(b.second);
  }
  h.blk_bits = best_k;
  h.blk_ok = true;
}

void pack_c16(SpeciesHops &h) {
  h.c16_ok = false;
  h.c16.clear();
  if (h.tgt.empty() || h.row_ptr.size() - 1 > 32768) return;
  h.mag16 = std::fabs(h.amp[0]);
  for (double a : h.amp)
    if (std::fabs(a) != h.mag16) return;
  h.c16.resize(h.tgt.size());
  for (size_t e = 0; e < h.tgt.size(); ++e)
    h.c16[e] = (uint16_t)((uint32_t)h.tgt[e] | (h.amp[e] < 0 ? 0x8000u : 0u));
  h.c16_ok = true;
}

void upload_species(SpeciesHops &h) {
  h.d_row_ptr = upload(h.row_ptr);
  h.d_tgt = upload(h.tgt);
  h.d_amp = upload(h.amp);
  h.d_onsite = upload(h.onsite);
  std::vector<float> af(h.amp.begin(), h.amp.end()), of(h.onsite.begin(), h.onsite.end());
  h.d_ampf = upload(af);
  h.d_onsitef = upload(of);
  if (h.ell_ok) {
    h.d_ell = upload(h.ell);
    h.d_cnt = upload(h.cnt);
    h.d_mags = upload(h.mags);
    std::vector<float> mf(h.mags.begin(), h.mags.end());
    h.d_magsf = upload(mf);
  }
  if (h.comp_ok) {
    h.d_comp = upload(h.comp);
    h.d_csign = upload(h.csign);
    h.d_cgmax = upload(h.cgmax);
  }
  if (h.c16_ok) h.d_c16 = upload(h.c16);
  if (h.blk_ok) {
    h.d_blk = upload(h.blk);
    h.d_nin = upload(h.nin);
  }
  if (h.srt_ok) {
    h.d_sperm = upload(h.sperm);
    h.d_scomp = upload(h.scomp);
    h.d_spn = upload(h.spn);
  }
}

}  // namespace

void lattice_bonds(int ndim, const int *sizes, std::vector<int> &from, std::vector<int> &to) {
  from.clear();
  to.clear();
  if (ndim == 1) {  // hubbard_model.h:16-20
    const int L = sizes[0];
    for (int i = 0; i < L; ++i) { from.push_back(i); to.push_back((i + 1) % L); }
  } else if (ndim == 2) {  // hubbard_model.h:72-82: x outer, y inner; x-bond then y-bond
    const int Lx = sizes[0], Ly = sizes[1];
    for (int x = 0; x < Lx; ++x)
      for (int y = 0; y < Ly; ++y) {
        const int s = x + Lx * y;
        from.push_back(s); to.push_back(((x + 1) % Lx) + Lx * y);
        from.push_back(s); to.push_back(x + Lx * ((y + 1) % Ly));
      }
  } else if (ndim == 3) {  // hubbard_model.h:144-159
    const int Lx = sizes[0], Ly = sizes[1], Lz = sizes[2];
    for (int x = 0; x < Lx; ++x)
      for (int y = 0; y < Ly; ++y)
        for (int z = 0; z < Lz; ++z) {
          const int s = x + Lx * y + Lx * Ly * z;
          from.push_back(s); to.push_back(((x + 1) % Lx) + Lx * y + Lx * Ly * z);
          from.push_back(s); to.push_back(x + Lx * ((y + 1) % Ly) + Lx * Ly * z);
          from.push_back(s); to.push_back(x + Lx * y + Lx * Ly * ((z + 1) % Lz));
        }
  } else {
    throw std::invalid_argument("ndim must be 1, 2 or 3");
  }
}

void build_hubbard(Hubbard &h) {
  const Sector &s = *h.sec;
  const auto mono = merge_monomials(h.bond_from, h.bond_to, h.t);
  species_hops(s.s_up, mono, h.up);
  species_hops(s.s_dn, mono, h.dn);
  finish_hubbard(h);
}

void finish_hubbard(Hubbard &h) {
  const Sector &s = *h.sec;
  pack_ell(h.dn);
  build_up_blocks(s.s_up, s.L, h.up);  // reorders the up entries: before pack_c16
  pack_c16(h.up);
  h.has_onsite = false;
  for (double v : h.up.onsite) h.has_onsite |= (v != 0.0);
  for (double v : h.dn.onsite) h.has_onsite |= (v != 0.0);
  // U n_up n_dn per doubly occupied site, summed term by term
  h.diagU.assign(s.L + 1, 0.0);
  if (!(h.U * h.U < kTol * kTol))
    for (int k = 0; k <= s.L; ++k) {
      double acc = 0.0;
      for (int i = 0; i < k; ++i) acc += h.U;
      h.diagU[k] = acc;
    }
  int dev = 0;
  MB_HIP(hipGetDevice(&dev));
  upload_species(h.up);
  upload_species(h.dn);
  h.d_diagU = upload(h.diagU);
  std::vector<float> df(h.diagU.begin(), h.diagU.end());
  h.d_diagUf = upload(df);
  h.device = dev;
}

HvTables tables_of(const Hubbard &h) {
  HvTables t;
  t.d_up = h.sec->d_up;
  t.d_dn = h.sec->d_dn;
  t.s_up = h.sec->d_s_up.as<uint32_t>();
  t.s_dn = h.sec->d_s_dn.as<uint32_t>();
  t.up_ptr = h.up.d_row_ptr.as<int32_t>();
  t.up_tgt = h.up.d_tgt.as<int32_t>();
  t.dn_ptr = h.dn.d_row_ptr.as<int32_t>();
  t.dn_tgt = h.dn.d_tgt.as<int32_t>();
  t.up_amp = h.up.d_amp.as<double>();
  t.dn_amp = h.dn.d_amp.as<double>();
  t.up_on = h.up.d_onsite.as<double>();
  t.dn_on = h.dn.d_onsite.as<double>();
  t.diagU = h.d_diagU.as<double>();
  t.up_ampf = h.up.d_ampf.as<float>();
  t.dn_ampf = h.dn.d_ampf.as<float>();
  t.up_onf = h.up.d_onsitef.as<float>();
  t.dn_onf = h.dn.d_onsitef.as<float>();
  t.diagUf = h.d_diagUf.as<float>();
  t.has_onsite = h.has_onsite ? 1 : 0;
  t.ell_ok = h.dn.ell_ok ? 1 : 0;
  t.ell_k = h.dn.ell_k;
  t.n_mags = (int)h.dn.mags.size();
  t.dn_ell = h.dn.d_ell.as<uint32_t>();
  t.dn_cnt = h.dn.d_cnt.as<uint8_t>();
  t.mags = h.dn.d_mags.as<double>();
  t.magsf = h.dn.d_magsf.as<float>();
  t.comp_ok = h.dn.comp_ok ? 1 : 0;
  t.dn_comp = h.dn.d_comp.as<uint16_t>();
  t.dn_csign = h.dn.d_csign.as<uint32_t>();
  t.dn_cgmax = h.dn.d_cgmax.as<uint8_t>();
  t.srt_ok = h.dn.srt_ok ? 1 : 0;
  t.srt_np = h.dn.srt_np;
  t.srt_perm = h.dn.d_sperm.as<uint16_t>();
  t.srt_comp = h.dn.d_scomp.as<uint16_t>();
  t.srt_pn = h.dn.d_spn.as<uint16_t>();
  t.n_diag = std::min(h.sec->n_up, h.sec->n_dn) + 1;
  t.max_up_row = h.up.max_row;
  t.up_c16_ok = h.up.c16_ok ? 1 : 0;
  t.up_c16 = h.up.d_c16.as<uint16_t>();
  t.up_mag = h.up.mag16;
  t.ub_ok = (h.up.blk_ok && h.up.c16_ok) ? 1 : 0;
  t.ub_nb = (int)(h.up.blk.size() / 2);
  t.ub_rows = h.up.blk_rows;
  t.ub_ent = h.up.blk_ent;
  t.ub_blk = h.up.d_blk.as<int32_t>();
  t.ub_nin = h.up.d_nin.as<uint8_t>();
  return t;
}

}  // namespace mb
