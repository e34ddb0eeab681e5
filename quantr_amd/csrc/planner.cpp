// planner.cpp — groups a gate sequence into fused HBM sweeps ("stages") for gfx950.
//
// Replaces the one-pass-per-gate loop of Circuit::simulate_with_register
// (simulation.rs:37-56). Gates are applied in an order that respects every dependency of
// the reference's sequence: two gates commute when neither changes (targets) a bit the other
// touches, so diagonal gates and control bits never order each other.
//
// A tile stage owns k index bits B (always including the c lowest bits, so every tile is
// 2^(k-c) contiguous segments of 2^c amplitudes). A workgroup holds one tile: 2^r amplitudes
// per thread in registers, the remaining k-r tile bits spread over threads. Gates whose
// target bits lie in the current register bits are applied in registers; "remaps" move
// amplitudes through LDS so other tile bits become register bits. Diagonal gates and
// controls may sit on any bit, in or out of the tile.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>

#include "engine.hpp"

namespace qr {

namespace {

bool is_swap_perm(const FastOp& op) {
  if (op.kind != FK_DENSEK || op.k != 2 || op.cmask) return false;
  static const int perm[4] = {0, 2, 1, 3};
  for (int r = 0; r < 4; ++r)
    for (int j = 0; j < 4; ++j) {
      const cd v = op.m[r * 4 + j];
      const bool want = perm[j] == r;
      if (want ? !(v.real() == 1.0 && v.imag() == 0.0) : !(v.real() == 0.0 && v.imag() == 0.0))
        return false;
    }
  return true;
}

bool tile_eligible(const FastOp& op, int k) {
  return op.kind == FK_DENSE1 || op.kind == FK_DIAG1 || op.kind == FK_GRAD || (k >= 2 && is_swap_perm(op));
}

// Ops that need their target bit in a register slot (a 2x2 across it).
bool needs_reg_target(const FastOp& op) {
  return op.kind == FK_DENSE1 || (op.kind == FK_GRAD && (op.gen == 1 || op.gen == 2));
}

struct Masks {
  uint64_t t, r;
};

// Scans `remaining` in order; an op is taken if it is eligible, its targets lie in `allowed`
// (a set of global bits) and it does not depend on an op left behind.
void absorb(const std::vector<int>& remaining, const std::vector<Masks>& mk,
            const std::vector<uint8_t>& eligible, uint64_t allowed, std::vector<int>* taken,
            std::vector<int>* left) {
  uint64_t bT = 0, bR = 0;
  for (int i : remaining) {
    const uint64_t t = mk[i].t, r = mk[i].r;
    const bool dep = (t & (bT | bR)) || (r & bT);
    if (dep || !eligible[i] || (t & ~allowed)) {
      if (left) left->push_back(i);
      bT |= t;
      bR |= r;
    } else if (taken) {
      taken->push_back(i);
    }
  }
}

int count_absorb(const std::vector<int>& remaining, const std::vector<Masks>& mk,
                 const std::vector<uint8_t>& eligible, uint64_t allowed) {
  uint64_t bT = 0, bR = 0;
  int c = 0;
  for (int i : remaining) {
    const uint64_t t = mk[i].t, r = mk[i].r;
    const bool dep = (t & (bT | bR)) || (r & bT);
    if (dep || !eligible[i] || (t & ~allowed)) {
      bT |= t;
      bR |= r;
    } else {
      ++c;
    }
  }
  return c;
}

// Score of a register set for one phase: ops absorbed (x4), plus one for each absorbed
// controlled op whose in-tile controls are all register bits (those controls then cost
// nothing; thread-held ones cost a per-lane select of the whole 2x2).
int phase_score(const std::vector<int>& remaining, const std::vector<Masks>& mk, const std::vector<uint8_t>& eligible,
                const std::vector<uint64_t>& cmask, uint64_t regs, uint64_t tile) {
  uint64_t bT = 0, bR = 0;
  int score = 0;
  for (int i : remaining) {
    const uint64_t t = mk[i].t, r = mk[i].r;
    const bool dep = (t & (bT | bR)) || (r & bT);
    if (dep || !eligible[i] || (t & ~regs)) {
      bT |= t;
      bR |= r;
    } else {
      score += 4;
      const uint64_t c = cmask[i] & tile;
      if (c && (c & ~regs) == 0) score += 1;
    }
  }
  return score;
}

uint64_t greedy_phase_bits(int want, const std::vector<int>& cands, const std::vector<int>& remaining,
                           const std::vector<Masks>& mk, const std::vector<uint8_t>& eligible,
                           const std::vector<uint64_t>& cmask, uint64_t tile) {
  uint64_t cur = 0;
  for (int step = 0; step < want; ++step) {
    int best = -1, best_bit = -1;
    for (int b : cands) {
      if (cur & (1ull << b)) continue;
      const int c = phase_score(remaining, mk, eligible, cmask, cur | (1ull << b), tile);
      if (c > best) {
        best = c;
        best_bit = b;
      }
    }
    if (best_bit < 0) break;
    cur |= 1ull << best_bit;
  }
  return cur;
}

// Register set for the next phase, with one phase of lookahead: every `want`-subset of
// `cands` is scored by what it absorbs (phase_score); the best few are scored again by what
// the best following subset of `next_cands` then absorbs. A set that absorbs everything left
// wins outright, preferring one without bits of `low` (the stage then ends without a remap
// back to the coalesced store layout). Each remap costs a round trip of the tile through LDS,
// so fewer phases is the objective; the lookahead avoids greedy sets that strand ops.
uint64_t lookahead_phase_bits(int want, const std::vector<int>& cands, const std::vector<int>& next_cands,
                              const std::vector<int>& remaining, const std::vector<Masks>& mk,
                              const std::vector<uint8_t>& eligible, const std::vector<uint64_t>& cmask,
                              uint64_t tile, uint64_t low) {
  const int nc = (int)cands.size();
  want = std::min(want, nc);
  auto subsets = [](const std::vector<int>& cs, int w) {
    std::vector<uint64_t> out;
    const int m = (int)cs.size();
    if (w <= 0 || w > m || m > 20) return out;
    for (uint32_t s = 0; s < (1u << m); ++s) {
      if (__builtin_popcount(s) != w) continue;
      uint64_t b = 0;
      for (int j = 0; j < m; ++j)
        if (s & (1u << j)) b |= 1ull << cs[j];
      out.push_back(b);
    }
    return out;
  };
  const std::vector<uint64_t> first = subsets(cands, want);
  const std::vector<uint64_t> second = subsets(next_cands, std::min<int>(want, (int)next_cands.size()));
  if (first.empty()) return greedy_phase_bits(want, cands, remaining, mk, eligible, cmask, tile);
  const int total = (int)remaining.size();
  std::vector<std::pair<int, uint64_t>> scored;
  scored.reserve(first.size());
  for (uint64_t s : first) scored.push_back({phase_score(remaining, mk, eligible, cmask, s, tile), s});
  std::stable_sort(scored.begin(), scored.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  // a set that finishes the stage
  uint64_t finish = 0;
  bool have_finish = false;
  for (const auto& sc : scored) {
    if (count_absorb(remaining, mk, eligible, sc.second) != total) continue;
    if (!have_finish || ((finish & low) && !(sc.second & low))) finish = sc.second;
    have_finish = true;
  }
  if (have_finish) return finish;
  const int keep = std::min<int>((int)scored.size(), 8);  // wider gave no fewer remaps on the
                                                          // workloads measured, at twice the time
  int best = -1;
  uint64_t best_set = scored[0].second;
  std::vector<int> left;
  for (int i = 0; i < keep; ++i) {
    if (scored[i].first <= 0) break;
    left.clear();
    absorb(remaining, mk, eligible, scored[i].second, nullptr, &left);
    int s2 = 0;
    for (uint64_t t : second) {
      int v = phase_score(left, mk, eligible, cmask, t, tile);
      // finishing after the next phase, without a final remap, is worth one more op
      if (count_absorb(left, mk, eligible, t) == (int)left.size()) v += (t & low) ? 2 : 6;
      s2 = std::max(s2, v);
    }
    const int v = scored[i].first + s2;
    if (v > best) {
      best = v;
      best_set = scored[i].second;
    }
  }
  return best_set;
}

// Greedy choice of `want` extra bits from `cands` on top of `base` maximising absorption.
uint64_t greedy_bits(uint64_t base, int want, const std::vector<int>& cands,
                     const std::vector<int>& remaining, const std::vector<Masks>& mk,
                     const std::vector<uint8_t>& eligible) {
  uint64_t cur = base;
  for (int step = 0; step < want; ++step) {
    int best = -1, best_bit = -1;
    for (int b : cands) {
      if (cur & (1ull << b)) continue;
      const int c = count_absorb(remaining, mk, eligible, cur | (1ull << b));
      if (c > best) {
        best = c;
        best_bit = b;
      }
    }
    if (best_bit < 0) break;
    cur |= 1ull << best_bit;
  }
  return cur;
}

// True if the low `bits` bits of the patterns of the first `m` thread slots are linearly
// independent over GF(2).
bool independent(const std::vector<uint16_t>& pat, int m, int bits) {
  const uint16_t mask = (uint16_t)((1u << bits) - 1);
  uint16_t basis[16] = {0};
  int rank = 0;
  for (int i = 0; i < m; ++i) {
    uint16_t v = pat[i] & mask;
    for (int b = bits - 1; b >= 0 && v; --b) {
      if (!((v >> b) & 1)) continue;
      if (!basis[b]) {
        basis[b] = v;
        ++rank;
        v = 0;
        break;
      }
      v ^= basis[b];
    }
  }
  return rank == m;
}

// True if the rows r[0..q) (bits 0..q-1) are linearly independent over GF(2).
bool invertible_low(const std::vector<uint16_t>& S, int q) {
  std::vector<uint16_t> rows(S.begin(), S.begin() + q);
  for (auto& r : rows) r &= (uint16_t)((1u << q) - 1);
  return independent(rows, q, q);
}

// LDS address pattern per tile position (the LDS slot of an amplitude is the XOR of the
// patterns of its set tile coordinates). 16-byte accesses are served eight lanes at a time:
// the lanes of each group of eight must hit distinct slots mod 8, i.e. the patterns of thread
// slots 0..2 (lane bits 0..2) must be independent in the low 3 bits, for the map written
// (old) and the map read (new). Patterns are S[p] = (1 << p) ^ e_p for p >= 4 and free rows
// of the low 4 x 4 block for p < 4 (kept invertible, so the map stays a bijection); only the
// positions held by those lane slots are changed from the identity.
std::vector<uint16_t> choose_swizzle(int k, int nts, const std::vector<int>& oldm, const std::vector<int>& newm) {
  std::vector<uint16_t> S(k);
  for (int p = 0; p < k; ++p) S[p] = (uint16_t)(1u << p);
  const int lanes = std::min(nts, 3);
  if (lanes == 0 || k < 3) return S;
  const int q = std::min(k, 4);
  auto ok = [&]() {
    for (const std::vector<int>* m : {&oldm, &newm}) {
      std::vector<uint16_t> pat(lanes);
      for (int i = 0; i < lanes; ++i) pat[i] = S[(*m)[i]];
      if (!independent(pat, lanes, 3)) return false;
    }
    return true;
  };
  if (ok()) return S;
  std::vector<int> free_pos;
  for (const std::vector<int>* m : {&oldm, &newm})
    for (int i = 0; i < lanes; ++i)
      if (std::find(free_pos.begin(), free_pos.end(), (*m)[i]) == free_pos.end()) free_pos.push_back((*m)[i]);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (int attempt = 0; attempt < 8192; ++attempt) {
    for (int p : free_pos) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      const uint16_t e = (uint16_t)(x & ((1u << q) - 1));
      S[p] = p >= q ? (uint16_t)((1u << p) ^ e) : e;
    }
    if (!invertible_low(S, q)) continue;
    if (ok()) return S;
  }
  for (int p = 0; p < k; ++p) S[p] = (uint16_t)(1u << p);
  return S;
}

}  // namespace

void plan_key_append(const FastOp& op, std::vector<uint64_t>& key) {
  uint64_t cls = 0;
  if (op.kind == FK_DENSE1) {
    cls = (uint64_t)op.form;
    if (op.form == U1_PERM && op.m[1] == cd(1, 0) && op.m[2] == cd(1, 0)) cls = 3;
  } else if (op.kind == FK_DIAG1) {
    if (op.m[0] == cd(1, 0)) {
      cls = 1;
      if (op.m[1] == cd(-1, 0)) cls |= 1u << 4;
      else if (op.m[1] == cd(0, 1)) cls |= 2u << 4;
      else if (op.m[1] == cd(0, -1)) cls |= 3u << 4;
    }
  } else if (op.kind == FK_DENSEK) {
    cls = is_swap_perm(op) ? 1 : 0;
  }
  uint64_t bits = 0;
  for (int q = 0; q < op.k && q < 8; ++q) bits |= (uint64_t)(op.bits[q] & 0xFF) << (8 * q);
  key.push_back((uint64_t)op.kind | ((uint64_t)op.k << 8) | ((uint64_t)op.gen << 16) | (cls << 24) |
                ((uint64_t)op.sub << 40));
  key.push_back(op.cmask);
  key.push_back(bits);
}

std::vector<Stage> plan_fused(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                              int bytes_per_amp, bool dual) {
  std::vector<Stage> stages;
  const int N = (int)ops.size();
  std::vector<Masks> mk(N);
  std::vector<uint8_t> eligible(N);
  const int k = std::min<int>(pp.tile_bits, (int)n);
  std::vector<uint64_t> cmask(N);
  for (int i = 0; i < N; ++i) cmask[i] = ops[i].cmask;
  for (int i = 0; i < N; ++i) {
    mk[i] = {ops[i].target_mask(), ops[i].read_mask()};
    eligible[i] = tile_eligible(ops[i], k) ? 1 : 0;
  }
  const int r = std::min(pp.reg_bits, k);
  const int nts = k - r;  // thread slots
  const int c = std::min(pp.low_bits, nts);
  const double sweep_bytes = 2.0 * (double)(1ull << n) * bytes_per_amp;
  uint64_t low_mask = (c > 0) ? ((1ull << c) - 1) : 0;

  std::vector<int> remaining(N);
  for (int i = 0; i < N; ++i) remaining[i] = i;

  while (!remaining.empty()) {
    const int front = remaining[0];
    if (!eligible[front]) {
      Stage s;
      s.tile = false;
      s.standalone = front;
      s.bytes = sweep_bytes * ops[front].f_touch;
      s.ngates = 1;
      stages.push_back(std::move(s));
      remaining.erase(remaining.begin());
      continue;
    }
    // ---- choose the tile bits B ------------------------------------------------------
    std::vector<int> all_bits(n);
    for (uint32_t b = 0; b < n; ++b) all_bits[b] = (int)b;
    // the front op's targets, then the low bits (coalescing), then greedy
    uint64_t B = mk[front].t;
    for (int b = 0; b < c && __builtin_popcountll(B) < k; ++b) B |= 1ull << b;
    int have = __builtin_popcountll(B);
    if (have < k) B = greedy_bits(B, k - have, all_bits, remaining, mk, eligible);
    // fill up if greedy ran out of useful bits (keeps k fixed for the kernel variant)
    for (uint32_t b = 0; b < n && __builtin_popcountll(B) < k; ++b) B |= 1ull << b;

    std::vector<int> taken, left;
    absorb(remaining, mk, eligible, B, &taken, &left);
    if (pp.max_ops > 0 && (int)taken.size() > pp.max_ops) {
      // defer the tail; the remaining list stays in program order
      left.insert(left.end(), taken.begin() + pp.max_ops, taken.end());
      taken.resize(pp.max_ops);
      std::sort(left.begin(), left.end());
    }

    Stage s;
    s.tile = true;
    s.k = k;
    for (uint32_t b = 0; b < n; ++b)
      if (B & (1ull << b)) s.tbits.push_back((int)b);
    int pos_of_bit[64];
    for (int i = 0; i < 64; ++i) pos_of_bit[i] = -1;
    for (int p = 0; p < k; ++p) pos_of_bit[s.tbits[p]] = p;
    s.bytes = sweep_bytes;
    s.ngates = (int)taken.size();

    // ---- phases: register bit sets and remaps ------------------------------------------
    // map[slot] = tile position; slots [0, nts) are thread bits, [nts, k) register bits.
    std::vector<int> map(k);
    auto make_map = [&](uint64_t regbits_global) {
      std::vector<int> m;
      // thread slots: low positions first (coalescing / conflict-free LDS), then the rest
      for (int p = 0; p < k; ++p)
        if (!(regbits_global & (1ull << s.tbits[p]))) m.push_back(p);
      for (int p = 0; p < k; ++p)
        if (regbits_global & (1ull << s.tbits[p])) m.push_back(p);
      return m;
    };
    auto push_map = [&](const std::vector<int>& m) {
      const int idx = (int)(s.maps.size() / k);
      s.maps.insert(s.maps.end(), m.begin(), m.end());
      return idx;
    };
    // candidates for register bits
    std::vector<int> tile_bits_vec(s.tbits.begin(), s.tbits.end());
    std::vector<int> nonlow_bits;
    for (int b : s.tbits)
      if (!(low_mask & (1ull << b))) nonlow_bits.push_back(b);

    std::vector<int> rem = taken;
    // first register set: from the non-low tile bits so the load stays coalesced
    uint64_t R = lookahead_phase_bits(r, nonlow_bits, tile_bits_vec, rem, mk, eligible, cmask, B, low_mask);
    for (int b : nonlow_bits)
      if (__builtin_popcountll(R) < r) R |= 1ull << b;
    map = make_map(R);
    push_map(map);  // map 0 = load map

    auto slot_of_pos = [&](int p) {
      for (int sl = 0; sl < k; ++sl)
        if (map[sl] == p) return sl;
      return -1;
    };
    // Control bits held in register slots become a uniform slot mask; the rest (thread slots
    // and bits outside the tile) a per-lane mask on the global index.
    auto split_controls = [&](uint64_t cmask, TileOpHost& t) {
      for (int b = 0; b < 64; ++b) {
        if (!(cmask & (1ull << b))) continue;
        const int p = pos_of_bit[b];
        const int sl = p >= 0 ? slot_of_pos(p) : -1;
        if (sl >= nts) t.ns |= 1u << (sl - nts);
        else t.lane_cmask |= 1ull << b;
      }
    };
    // A remap moves data through LDS; its swizzle keeps the 16 lanes of a quarter-wave on
    // distinct 16-byte LDS slots for both the write (old map) and the read (new map).
    auto emit_remap = [&](const std::vector<int>& nm) {
      TileOpHost t{};
      t.code = TC_REMAP;
      t.swz = (uint32_t)(s.swz.size() / (2 * k));
      std::vector<uint16_t> table = choose_swizzle(k, nts, map, nm);
      // per-slot LDS patterns: old map (write) then new map (read)
      for (int sl = 0; sl < k; ++sl) s.swz.push_back(table[map[sl]]);
      for (int sl = 0; sl < k; ++sl) s.swz.push_back(table[nm[sl]]);
      t.a = (uint32_t)push_map(nm);
      s.ops.push_back(t);
      map = nm;
    };

    while (!rem.empty()) {
      // global bits currently in register slots
      uint64_t regs = 0;
      for (int j = 0; j < r; ++j) regs |= 1ull << s.tbits[map[nts + j]];
      // ops allowed now: DENSE1 with target in regs; DIAG1 anywhere; SWAP within the tile.
      // A swap relabels slots, so the scan stops after the first one.
      std::vector<int> now, later;
      bool hit_swap = false;
      {
        uint64_t bT = 0, bR = 0;
        for (int i : rem) {
          if (hit_swap) {
            later.push_back(i);
            continue;
          }
          const FastOp& op = ops[i];
          const uint64_t t = mk[i].t, rr = mk[i].r;
          const bool dep = (t & (bT | bR)) || (rr & bT);
          bool ok = !dep;
          if (ok && needs_reg_target(op)) ok = (t & ~regs) == 0;
          if (ok && op.kind == FK_DENSEK) ok = (t & ~B) == 0;
          if (ok) {
            now.push_back(i);
            if (op.kind == FK_DENSEK) hit_swap = true;
          } else {
            later.push_back(i);
            bT |= t;
            bR |= rr;
          }
        }
      }
      for (int i : now) {
        const FastOp& op = ops[i];
        TileOpHost t{};
        split_controls(op.cmask, t);
        if (op.kind == FK_DENSE1) {
          t.code = TC_U1;
          const int p = pos_of_bit[op.bits[0]];
          t.a = (uint32_t)(slot_of_pos(p) - nts);
          for (int q = 0; q < 4; ++q) t.coef[q] = op.m[q];
          int form = op.form;
          const bool x = form == U1_PERM && op.m[1] == cd(1, 0) && op.m[2] == cd(1, 0);
          const bool free = !dual && t.ns == 0 && t.lane_cmask == 0;  // uncontrolled: scalar is global
          if (x) {
            form = TF_X;
          } else if (free && op.sub == SUB_RROT) {
            form = TF_RROT;
            t.fold = FOLD_RROT;
          } else if (free && op.sub == SUB_XROT) {
            form = TF_XROT;
            t.fold = FOLD_XROT;
          } else if (op.sub == SUB_RX || op.sub == SUB_XROT) {
            form = TF_RX;
          } else if (free && op.sub == SUB_HAD) {
            form = TF_HAD;
            t.fold = FOLD_M00;
          } else if (free && op.sub == SUB_XLIKE) {
            form = TF_X;
            t.fold = FOLD_M10;
          } else if (free && op.sub == SUB_YLIKE) {
            form = TF_NEGSWAP;
            t.fold = FOLD_M10;
          }
          t.b = (uint32_t)form;
          t.src = i;
          s.ops.push_back(t);
        } else if (op.kind == FK_DIAG1) {
          t.code = TC_D1;
          t.a = (uint32_t)op.bits[0];
          const int p = pos_of_bit[op.bits[0]];
          const int sl = p >= 0 ? slot_of_pos(p) : -1;
          t.b = (sl >= nts) ? (uint32_t)(sl - nts) : 0xFFu;
          t.coef[0] = op.m[0];
          t.coef[1] = op.m[1];
          t.src = i;
          if (op.m[0] == cd(1, 0)) {
            t.flags = 1;  // only the |1> half changes
            if (op.m[1] == cd(-1, 0)) t.flags |= 1u << 4;
            else if (op.m[1] == cd(0, 1)) t.flags |= 2u << 4;
            else if (op.m[1] == cd(0, -1)) t.flags |= 3u << 4;
          } else if (!dual && t.ns == 0 && t.lane_cmask == 0 && op.sub == SUB_D0) {
            t.flags = 1;  // d0 factored out: diag(1, d1/d0)
            t.fold = FOLD_D0;
          }
          s.ops.push_back(t);
        } else if (op.kind == FK_GRAD) {
          t.code = TC_GRAD;
          t.flags = (uint32_t)op.gen;
          t.param = (uint32_t)op.param;
          if (op.gen == 1 || op.gen == 2) {
            t.a = (uint32_t)(slot_of_pos(pos_of_bit[op.bits[0]]) - nts);
          } else {
            t.a = (uint32_t)op.bits[0];
            const int p = op.gen == 3 ? pos_of_bit[op.bits[0]] : -1;
            const int sl = p >= 0 ? slot_of_pos(p) : -1;
            t.b = (sl >= nts) ? (uint32_t)(sl - nts) : 0xFFu;
          }
          s.ops.push_back(t);
        } else {  // swap: exchange the tile positions held by the two slots
          const int p0 = pos_of_bit[op.bits[0]], p1 = pos_of_bit[op.bits[1]];
          const int s0 = slot_of_pos(p0), s1 = slot_of_pos(p1);
          std::swap(map[s0], map[s1]);
          t.code = TC_SWAPSLOTS;
          t.a = (uint32_t)push_map(map);
          s.ops.push_back(t);
        }
      }
      rem = later;
      if (rem.empty() || hit_swap) continue;
      // next register set: any tile bits
      uint64_t R2 = lookahead_phase_bits(r, tile_bits_vec, tile_bits_vec, rem, mk, eligible, cmask, B, low_mask);
      for (int b : tile_bits_vec)
        if (__builtin_popcountll(R2) < r) R2 |= 1ull << b;
      if (now.empty() && count_absorb(rem, mk, eligible, R2) == 0)
        throw Error{QR_ERR_UNSUPPORTED, "internal planner error: no progress inside a tile"};
      emit_remap(make_map(R2));
    }
    // ---- folded scalars: one general-form uncontrolled op absorbs their product, else a
    // final scale applies it
    bool folded = false;
    for (const auto& t : s.ops) folded = folded || (t.fold != FOLD_NONE);
    if (folded) {
      bool absorbed = false;
      for (auto& t : s.ops)
        if (t.code == TC_U1 && t.fold == FOLD_NONE && t.b == U1_GENERAL && t.ns == 0 && t.lane_cmask == 0) {
          t.fold = FOLD_ABSORB;
          absorbed = true;
          break;
        }
      if (!absorbed) {
        TileOpHost t{};
        t.code = TC_SCALE;
        t.fold = FOLD_SCALE;
        s.ops.push_back(t);
      }
    }
    // ---- final layout: low tile positions on the low thread slots for a coalesced store --
    bool canon = true;
    for (int sl = 0; sl < c; ++sl)
      if (map[sl] != sl) canon = false;
    if (!canon) {
      uint64_t Rf = 0;
      for (int b : nonlow_bits)
        if (__builtin_popcountll(Rf) < r) Rf |= 1ull << b;
      emit_remap(make_map(Rf));
    }
    {  // plan-time values of the coefficients (the CPU model reads them)
      std::vector<cd> c;
      stage_coefficients(s, ops, c);
      for (size_t i = 0; i < s.ops.size(); ++i)
        for (int q = 0; q < 4; ++q) s.ops[i].coef[q] = c[4 * i + q];
    }
    stages.push_back(std::move(s));
    remaining = left;
  }
  return stages;
}

void stage_coefficients(const Stage& s, const std::vector<FastOp>& ops, std::vector<cd>& c) {
  c.assign(4 * s.ops.size(), cd(0, 0));
  cd total(1, 0);
  int absorber = -1, scaler = -1;
  for (size_t i = 0; i < s.ops.size(); ++i) {
    const TileOpHost& t = s.ops[i];
    cd* o = &c[4 * i];
    if (t.src < 0) {
      for (int q = 0; q < 4; ++q) o[q] = t.coef[q];
      if (t.code == TC_GRAD) o[0] = cd((double)t.param, 0.0);  // the tap's row
      if (t.fold == FOLD_SCALE) scaler = (int)i;
      continue;
    }
    const std::vector<cd>& m = ops[t.src].m;
    for (size_t q = 0; q < 4 && q < m.size(); ++q) o[q] = m[q];
    cd f(1, 0);
    switch (t.fold) {
      case FOLD_M00: f = m[0]; break;
      case FOLD_M10: f = m[2]; break;
      case FOLD_D0: f = m[0]; break;
      case FOLD_ABSORB: absorber = (int)i; break;
      default: break;
    }
    if (t.fold == FOLD_M00 || t.fold == FOLD_M10) {  // the residual matrix is exact
      static const cd had[4] = {cd(1, 0), cd(1, 0), cd(1, 0), cd(-1, 0)};
      static const cd xs[4] = {cd(0, 0), cd(1, 0), cd(1, 0), cd(0, 0)};
      static const cd ns[4] = {cd(0, 0), cd(-1, 0), cd(1, 0), cd(0, 0)};
      const cd* r = t.b == TF_HAD ? had : (t.b == TF_X ? xs : ns);
      for (int q = 0; q < 4; ++q) o[q] = r[q];
      total *= f;
    } else if (t.fold == FOLD_D0) {
      o[1] = m[1] / f;
      o[0] = cd(1, 0);
      total *= f;
    } else if (t.fold == FOLD_RROT) {  // [[a,-b],[b,a]]
      const double a = m[0].real(), b = m[2].real();
      if (std::fabs(a) >= std::fabs(b)) {
        const double tq = b / a;
        o[0] = cd(1, 0); o[1] = cd(-tq, 0); o[2] = cd(tq, 0); o[3] = cd(1, 0);
        total *= a;
      } else {
        const double u = a / b;
        o[0] = cd(u, 0); o[1] = cd(-1, 0); o[2] = cd(1, 0); o[3] = cd(u, 0);
        total *= b;
      }
    } else if (t.fold == FOLD_XROT) {  // [[a,-ib],[-ib,a]]
      const double a = m[0].real(), b = -m[1].imag();
      if (std::fabs(a) >= std::fabs(b)) {
        const double tq = b / a;
        o[0] = cd(1, 0); o[1] = cd(0, -tq); o[2] = cd(0, -tq); o[3] = cd(1, 0);
        total *= a;
      } else {
        const double u = a / b;
        o[0] = cd(0, u); o[1] = cd(1, 0); o[2] = cd(1, 0); o[3] = cd(0, u);
        total *= cd(0, -b);
      }
    }
  }
  if (absorber >= 0)
    for (int q = 0; q < 4; ++q) c[4 * absorber + q] *= total;
  if (scaler >= 0) c[4 * scaler] = total;
}

// ---- stage balancing ----------------------------------------------------------------------
//
// A tile sweep takes the longer of its HBM time and its ALU time. The greedy planner fills
// each stage with every op it can take, so some stages carry 170 ops (ALU-bound, twice the
// HBM time) while others carry 20. Capping the ops per stage moves work into later, lighter
// sweeps; whether a cap adds sweeps or even removes some depends on the circuit, so a few
// caps are planned and the one with the lowest estimated duration is kept.
namespace {

// Ops (plus LDS remaps) whose ALU time over a complex128 state equals one sweep of it at copy
// speed: measured on MI355X, where a plain copy moves about 5.4 TB/s and a 172-op stage of
// the 30-qubit brickwork takes 2.1 times a copy-speed sweep. Proportional to the bytes per
// amplitude, since the ALU work per amplitude does not depend on the precision.
constexpr double kSweepOpsC128 = 87.0;

double plan_cost(const std::vector<Stage>& st, int bytes_per_amp) {
  const double sweep = kSweepOpsC128 * bytes_per_amp / 16.0;
  double cost = 0.0;
  for (const Stage& s : st) {
    if (!s.tile) {
      cost += sweep;
      continue;
    }
    double units = s.ngates;
    for (const auto& t : s.ops)
      if (t.code == TC_REMAP) units += 1.0;
    cost += std::max(sweep, units);
  }
  return cost;
}

}  // namespace

std::vector<Stage> plan_fused_balanced(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                                       int bytes_per_amp, bool dual) {
  if (pp.max_ops > 0 || dual || n < kBalanceMinQubits || pp.tile_bits >= (int)n)
    return plan_fused(n, ops, pp, bytes_per_amp, dual);
  static const int caps[] = {0, 160, 140, 120, 100, 90, 80};
  constexpr int nc = (int)(sizeof(caps) / sizeof(caps[0]));
  std::vector<std::vector<Stage>> plans(nc);
  std::vector<std::exception_ptr> errs(nc);
  std::vector<std::thread> th;
  for (int i = 0; i < nc; ++i)
    th.emplace_back([&, i] {
      try {
        PlanParams q = pp;
        q.max_ops = caps[i];
        plans[i] = plan_fused(n, ops, q, bytes_per_amp, dual);
      } catch (...) {
        errs[i] = std::current_exception();
      }
    });
  for (auto& t : th) t.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
  int best = 0;
  double best_cost = plan_cost(plans[0], bytes_per_amp);
  for (int i = 1; i < nc; ++i) {
    const double c = plan_cost(plans[i], bytes_per_amp);
    if (c < best_cost) {
      best_cost = c;
      best = i;
    }
  }
  return std::move(plans[best]);
}

}  // namespace qr
