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
#include <atomic>
#include <cmath>
#include <cstdlib>
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

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Greedy choice of `want` extra bits from `cands` on top of `base` maximising absorption.
// Ties go to the first candidate, or with a nonzero `salt` to the one whose salted hash is
// lowest (another plan of the same quality for the same program); with `slack`, bits that
// absorb up to `slack` ops fewer than the best also tie (plans greedy would not make).
uint64_t greedy_bits(uint64_t base, int want, const std::vector<int>& cands,
                     const std::vector<int>& remaining, const std::vector<Masks>& mk,
                     const std::vector<uint8_t>& eligible, uint64_t salt = 0, int slack = 0) {
  uint64_t cur = base;
  std::vector<std::pair<int, int>> score;
  for (int step = 0; step < want; ++step) {
    int best = -1, best_bit = -1;
    score.clear();
    for (int b : cands) {
      if (cur & (1ull << b)) continue;
      const int c = count_absorb(remaining, mk, eligible, cur | (1ull << b));
      score.push_back({b, c});
      best = std::max(best, c);
    }
    const int lowest = salt ? best - slack : best;
    const uint64_t h = salt ^ ((uint64_t)step << 40);
    for (const auto& bc : score) {
      if (bc.second < lowest) continue;
      if (best_bit < 0 || (salt && mix64(h ^ (uint64_t)bc.first) < mix64(h ^ (uint64_t)best_bit))) best_bit = bc.first;
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

// Pattern table of a remap done in two halves split by tile position h: the other positions
// get a conflict-free table over k-1 bits (choose_swizzle on the maps without h) and h gets
// the top bit, so the table is still a bijection onto 2^k slots (a kernel with the whole
// tile in LDS can run the remap in one round) while each round of the split exchange uses
// the low 2^(k-1) slots.
std::vector<uint16_t> choose_swizzle_split(int k, int nts, const std::vector<int>& oldm, const std::vector<int>& newm,
                                           int h) {
  auto reduce = [&](const std::vector<int>& m) {
    std::vector<int> r;
    for (int p : m)
      if (p != h) r.push_back(p > h ? p - 1 : p);
    return r;
  };
  const std::vector<uint16_t> t = choose_swizzle(k - 1, nts, reduce(oldm), reduce(newm));
  std::vector<uint16_t> S(k);
  for (int p = 0; p < k; ++p)
    if (p != h) S[p] = t[p > h ? p - 1 : p];
  S[h] = (uint16_t)(1u << (k - 1));
  return S;
}

// The tile position that splits a two-round remap from map `om` to map `nm`, best first:
// a register position in both maps at the same register slot (no temporaries), a register
// position in both maps, the same wavefront-index thread slot in both maps (each wavefront
// takes part in one round), a register position on one side and a wavefront-index slot on
// the other, then any position off lane slots 0-2 (their patterns carry the LDS swizzle).
int choose_split_pos(int k, int nts, const std::vector<int>& om, const std::vector<int>& nm) {
  std::vector<int> os(k), ns(k);
  for (int sl = 0; sl < k; ++sl) {
    os[om[sl]] = sl;
    ns[nm[sl]] = sl;
  }
  constexpr int kWaveSlot = 6;  // thread slots from 6 on index whole 64-lane wavefronts
  int best = -1, best_score = -1;
  for (int p = 0; p < k; ++p) {
    const bool oreg = os[p] >= nts, nreg = ns[p] >= nts;
    int score;
    if (oreg && nreg) score = os[p] == ns[p] ? 6 : 5;
    else if (!oreg && !nreg && os[p] == ns[p] && os[p] >= kWaveSlot) score = 4;
    else if ((oreg && ns[p] >= kWaveSlot) || (nreg && os[p] >= kWaveSlot)) score = 3;
    else if (os[p] >= kWaveSlot && ns[p] >= kWaveSlot) score = 2;
    else if (os[p] >= 3 && ns[p] >= 3) score = 1;
    else continue;
    if (score > best_score) {
      best_score = score;
      best = p;
    }
  }
  return best;
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
  // Uncontrolled Ry and Rx run as three shears (6 fused multiply-adds per amplitude pair, no
  // branch) instead of plain 2x2 forms (8); QR_ROT_FORMS=0 turns this off.
  const char* rot_env = std::getenv("QR_ROT_FORMS");
  const bool rot = rot_env ? rot_env[0] != '0' : true;
  std::vector<uint64_t> cmask(N);
  for (int i = 0; i < N; ++i) cmask[i] = ops[i].cmask;
  for (int i = 0; i < N; ++i) {
    mk[i] = {ops[i].target_mask(), ops[i].read_mask()};
    eligible[i] = tile_eligible(ops[i], k) ? 1 : 0;
  }
  const int r = std::min(pp.reg_bits, k);
  const int nts = k - r;  // thread slots
  const int c = std::min(pp.low_bits, nts);
  const bool split = pp.lds_split != 0 && nts >= 4;
  const double sweep_bytes = 2.0 * (double)(1ull << n) * bytes_per_amp;
  // Layout: logical index bit l (wire n-1-l) is stored at physical bit phys[l]. Without a
  // moving layout it is the identity. With one (dyn), each stage may store some of its tile
  // bits at the c physical low bits, so that the next stage's contiguous segments hold bits it
  // uses. The layout stays a product of at most c transpositions between a low physical bit j
  // and the home of the logical bit it holds, and is the identity again before a standalone op
  // and after the last stage (by an op-free restoring sweep if no stage can do it).
  const bool dyn = pp.dyn_low != 0 && !dual && c > 0 && k < (int)n;
  std::vector<int> phys(n), logi(n);
  for (uint32_t b = 0; b < n; ++b) phys[b] = logi[b] = (int)b;
  auto identity_layout = [&]() {
    for (uint32_t b = 0; b < n; ++b)
      if (phys[b] != (int)b) return false;
    return true;
  };
  std::vector<int> all_bits(n);
  for (uint32_t b = 0; b < n; ++b) all_bits[b] = (int)b;

  std::vector<int> remaining(N);
  for (int i = 0; i < N; ++i) remaining[i] = i;

  while (!remaining.empty() || (dyn && !identity_layout())) {
    // logical bits held by the physical low bits (the contiguous segment of every tile)
    uint64_t low_mask = 0;
    for (int j = 0; j < c; ++j) low_mask |= 1ull << logi[j];
    const bool restore_only = remaining.empty() || (dyn && !identity_layout() && !eligible[remaining[0]]);
    const int front = restore_only ? -1 : remaining[0];
    if (front >= 0 && !eligible[front]) {
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
    // the front op's targets, then the logical bits at the physical low bits (coalescing),
    // then greedy; an op-free restoring sweep takes every displaced bit and its partner
    uint64_t B = front >= 0 ? mk[front].t : 0;
    B |= low_mask;
    if (restore_only)
      for (uint32_t b = 0; b < n; ++b)
        if (phys[b] != (int)b || (int)b < c) B |= 1ull << b;
    int have = __builtin_popcountll(B);
    const uint64_t salt = pp.seed ? mix64(((uint64_t)pp.seed << 32) ^ stages.size()) | 1 : 0;
    if (have < k && !restore_only)
      B = greedy_bits(B, k - have, all_bits, remaining, mk, eligible, salt, pp.seed ? pp.tie_slack : 0);
    // fill up if greedy ran out of useful bits (keeps k fixed for the kernel variant)
    for (uint32_t b = 0; b < n && __builtin_popcountll(B) < k; ++b) B |= 1ull << b;
    if (dyn && !restore_only && !identity_layout()) {
      // when every op left fits in one tile that also holds every displaced bit, that tile is
      // the last stage and restores the layout (no op-free restoring sweep after it)
      uint64_t Bl = mk[front].t | low_mask;
      for (uint32_t b = 0; b < n; ++b)
        if (phys[b] != (int)b) Bl |= 1ull << b;
      const bool standalone_later = std::any_of(remaining.begin(), remaining.end(), [&](int i) { return !eligible[i]; });
      if (!standalone_later && __builtin_popcountll(Bl) <= k) {
        const int hv = __builtin_popcountll(Bl);
        if (hv < k) Bl = greedy_bits(Bl, k - hv, all_bits, remaining, mk, eligible);
        for (uint32_t b = 0; b < n && __builtin_popcountll(Bl) < k; ++b) Bl |= 1ull << b;
        if (count_absorb(remaining, mk, eligible, Bl) == (int)remaining.size() &&
            (pp.max_ops <= 0 || (int)remaining.size() <= pp.max_ops))
          B = Bl;
      }
    }

    std::vector<int> taken, left;
    if (!restore_only) {
      absorb(remaining, mk, eligible, B, &taken, &left);
      if (pp.max_ops > 0 && (int)taken.size() > pp.max_ops) {
        // defer the tail; the remaining list stays in program order
        left.insert(left.end(), taken.begin() + pp.max_ops, taken.end());
        taken.resize(pp.max_ops);
        std::sort(left.begin(), left.end());
      }
    } else {
      left = remaining;
    }
    if (taken.empty() && !restore_only) throw Error{QR_ERR_UNSUPPORTED, "internal planner error: a stage takes no op"};

    // ---- the layout this stage stores (moving layout only) ----------------------------
    std::vector<int> nphys = phys, nlogi = logi;
    if (dyn) {
      // the bits the next stage would choose if c of them must come from this tile: its front
      // op's targets, its best free picks, then (if the picks hold fewer than c of this tile's
      // bits) the best of this tile's bits
      uint64_t Bn = 0;
      const bool next_tile = !left.empty() && eligible[left[0]];
      if (next_tile) {
        Bn = mk[left[0]].t;
        const int need = std::max(0, c - __builtin_popcountll(Bn & B));
        const int hv = __builtin_popcountll(Bn);
        if (hv + need < k) Bn = greedy_bits(Bn, k - hv - need, all_bits, left, mk, eligible);
        const int have_mine = __builtin_popcountll(Bn & B);
        if (have_mine < c) {
          std::vector<int> mine;
          for (uint32_t b = 0; b < n; ++b)
            if ((B & (1ull << b)) && !(Bn & (1ull << b))) mine.push_back((int)b);
          Bn = greedy_bits(Bn, std::min(c - have_mine, k - __builtin_popcountll(Bn)), mine, left, mk, eligible);
        }
        const int hv2 = __builtin_popcountll(Bn);
        if (hv2 < k) Bn = greedy_bits(Bn, k - hv2, all_bits, left, mk, eligible);
      }
      // first use of each bit among the ops left (sooner first)
      std::vector<int> first_use(n, 1 << 30);
      for (size_t q = 0; q < left.size(); ++q) {
        uint64_t m = mk[left[q]].t | mk[left[q]].r;
        while (m) {
          const int b = __builtin_ctzll(m);
          m &= m - 1;
          if (first_use[b] > (int)q) first_use[b] = (int)q;
        }
      }
      // A candidate layout: for each low slot j, keep its bit when `keep` says so, else lower
      // the soonest-used bit of this tile that `want` contains. The bit leaving slot j goes
      // home when its home is in the tile, else to the place of the bit that replaces it. A
      // move is kept only while the displaced bits and the low bits fit in one tile, so one
      // op-free sweep can always restore the identity. Before a standalone op and at the end
      // every displaced bit goes home when all of them are in this tile.
      auto displaced_fit = [&](const std::vector<int>& cp) {
        uint64_t m = 0;
        for (uint32_t b = 0; b < n; ++b)
          if (cp[b] != (int)b) m |= 1ull << b;
        for (int j = 0; j < c; ++j) m |= 1ull << j;  // the low bits' own logical bits
        return __builtin_popcountll(m) <= k;
      };
      // when a tile with the identity's low bits takes every op left, the next stage is the last:
      // restore the identity now, so that it needs no op-free restoring sweep after it
      bool final_next = false;
      if (next_tile) {
        uint64_t Bi = mk[left[0]].t;
        for (int j = 0; j < c; ++j) Bi |= 1ull << j;
        if (__builtin_popcountll(Bi) <= k) {
          const int hv = __builtin_popcountll(Bi);
          if (hv < k) Bi = greedy_bits(Bi, k - hv, all_bits, left, mk, eligible);
          final_next = count_absorb(left, mk, eligible, Bi) == (int)left.size() &&
                       (pp.max_ops <= 0 || (int)left.size() <= pp.max_ops);
        }
      }
      auto candidate = [&](uint64_t want, bool keep_useful, std::vector<int>& cp, std::vector<int>& cl) {
        cp = phys;
        cl = logi;
        if (!next_tile || final_next) {
          // every displaced bit of this tile whose home is in the tile goes home (the bit there
          // takes its place); each move fixes one bit more, so this ends
          bool moved = true;
          while (moved) {
            moved = false;
            for (uint32_t b = 0; b < n; ++b) {
              if (cp[b] == (int)b || !(B & (1ull << b))) continue;
              const int z = cl[b];  // the bit at b's home
              if (!(B & (1ull << z))) continue;
              const int pb = cp[b];
              cp[b] = (int)b;
              cp[z] = pb;
              cl[b] = (int)b;
              cl[pb] = z;
              moved = true;
            }
          }
          return;
        }
        for (int j = 0; j < c; ++j) {
          const int x = cl[j];
          if (keep_useful && (want & (1ull << x))) continue;
          int y = -1, best = 1 << 30;
          for (uint32_t b = 0; b < n; ++b)
            if ((B & (1ull << b)) && (want & (1ull << b)) && cp[b] >= c && first_use[b] < best) {
              best = first_use[b];
              y = (int)b;
            }
          if (y < 0 || y == x) continue;
          std::vector<int> tp = cp;
          const int py = tp[y];
          tp[y] = j;
          const bool home_in_tile = (B & (1ull << cl[x])) != 0;  // the bit now at x's home
          if (home_in_tile && x != cl[x] && cl[x] != y) {
            const int z = cl[x];  // x home, z to y's place
            tp[x] = x;
            tp[z] = py;
          } else {
            tp[x] = py;  // x to y's place (x's home when y sat there)
          }
          if (!displaced_fit(tp)) continue;
          cp = tp;
          for (uint32_t b = 0; b < n; ++b) cl[cp[b]] = (int)b;
        }
      };
      // ops the next stage absorbs when its low bits are those of layout cl
      auto next_absorbs = [&](const std::vector<int>& cl) {
        uint64_t Bx = mk[left[0]].t;
        for (int j = 0; j < c; ++j) Bx |= 1ull << cl[j];
        const int hv = __builtin_popcountll(Bx);
        if (hv < k) Bx = greedy_bits(Bx, k - hv, all_bits, left, mk, eligible);
        return count_absorb(left, mk, eligible, Bx);
      };
      candidate(Bn, true, nphys, nlogi);
      if (next_tile && !final_next) {
        // other candidates: the soonest-used tile bits, keeping useful low bits or not, and the
        // unconstrained pick without keeping; the one whose next stage absorbs most wins
        uint64_t soon = 0;
        for (uint32_t b = 0; b < n; ++b)
          if (first_use[b] < (1 << 30)) soon |= 1ull << b;
        int best = next_absorbs(nlogi);
        const std::pair<uint64_t, bool> alts[] = {{soon, false}, {soon, true}, {Bn, false}};
        for (const auto& a : alts) {
          std::vector<int> p2, l2;
          candidate(a.first, a.second, p2, l2);
          const int v = next_absorbs(l2);
          if (v > best) {
            best = v;
            nphys = p2;
            nlogi = l2;
          }
        }
      }
    }
    if (dyn) {  // the stored layout is a permutation of the tile's physical bits
      std::vector<uint8_t> seen(n, 0);
      for (uint32_t b = 0; b < n; ++b) {
        const int q = nphys[b];
        if (q < 0 || q >= (int)n || seen[q] || nlogi[q] != (int)b || ((B >> b) & 1) != ((B >> logi[q]) & 1))
          throw Error{QR_ERR_UNSUPPORTED, "internal planner error: invalid stored layout"};
        seen[q] = 1;
      }
    }
    uint64_t nlow_mask = 0;
    for (int j = 0; j < c; ++j) nlow_mask |= 1ull << nlogi[j];

    Stage s;
    s.tile = true;
    s.k = k;
    // tile positions: the logical bits at the physical low bits (in physical order), then the
    // rest by physical position, so that thread slots in position order keep every segment
    // contiguous
    for (int j = 0; j < c; ++j) s.tbits.push_back(logi[j]);
    {
      std::vector<int> rest;
      for (uint32_t b = 0; b < n; ++b)
        if ((B & (1ull << b)) && !(low_mask & (1ull << b))) rest.push_back((int)b);
      std::sort(rest.begin(), rest.end(), [&](int a, int b2) { return phys[a] < phys[b2]; });
      s.tbits.insert(s.tbits.end(), rest.begin(), rest.end());
    }
    int pos_of_bit[64];
    for (int i = 0; i < 64; ++i) pos_of_bit[i] = -1;
    for (int p = 0; p < k; ++p) pos_of_bit[s.tbits[p]] = p;
    for (int p = 0; p < k; ++p) {
      s.pload.push_back(phys[s.tbits[p]]);
      s.pstore.push_back(nphys[s.tbits[p]]);
    }
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
    uint64_t R = lookahead_phase_bits(r, nonlow_bits, tile_bits_vec, rem, mk, eligible, cmask, B, nlow_mask);
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
        else t.lane_cmask |= 1ull << phys[b];  // tested on the physical index
      }
    };
    // A remap moves data through LDS; its swizzle keeps the 16 lanes of a quarter-wave on
    // distinct 16-byte LDS slots for both the write (old map) and the read (new map).
    auto emit_remap = [&](const std::vector<int>& nm_in) {
      TileOpHost t{};
      t.code = TC_REMAP;
      t.swz = (uint32_t)(s.swz.size() / (2 * k));
      std::vector<int> nm = nm_in;
      const int h = split ? choose_split_pos(k, nts, map, nm) : -1;
      if (h >= 0) {  // keep a register-held split position in the same register slot (free relabel)
        int os = -1, ns = -1;
        for (int sl = 0; sl < k; ++sl) {
          if (map[sl] == h) os = sl;
          if (nm[sl] == h) ns = sl;
        }
        if (os >= nts && ns >= nts && os != ns) std::swap(nm[os], nm[ns]);
      }
      std::vector<uint16_t> table;
      if (h >= 0) {
        table = choose_swizzle_split(k, nts, map, nm, h);
        t.flags = 1;
        t.b = (uint32_t)h;
      } else {
        table = choose_swizzle(k, nts, map, nm);
      }
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
          } else if (rot && free && op.sub == SUB_RROT) {
            form = TF_RROT;
            t.fold = FOLD_RROT;
          } else if (rot && free && op.sub == SUB_XROT) {
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
          t.a = (uint32_t)phys[op.bits[0]];
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
            t.a = (uint32_t)phys[op.bits[0]];
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
      uint64_t R2 = lookahead_phase_bits(r, tile_bits_vec, tile_bits_vec, rem, mk, eligible, cmask, B, nlow_mask);
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
    // ---- final layout: the positions stored at the physical low bits on thread slots 0..c-1
    // (in physical order) for a coalesced store --------------------------------------------
    bool canon = true;
    for (int sl = 0; sl < c; ++sl)
      if (map[sl] != pos_of_bit[nlogi[sl]]) canon = false;
    bool remapped_last = false;
    if (!canon) {
      uint64_t Rf = 0;
      for (int b : s.tbits)
        if (!(nlow_mask & (1ull << b)) && __builtin_popcountll(Rf) < r) Rf |= 1ull << b;
      std::vector<int> m;
      for (int sl = 0; sl < c; ++sl) m.push_back(pos_of_bit[nlogi[sl]]);
      for (int p = 0; p < k; ++p)
        if (!(Rf & (1ull << s.tbits[p])) && !(nlow_mask & (1ull << s.tbits[p]))) m.push_back(p);
      for (int p = 0; p < k; ++p)
        if (Rf & (1ull << s.tbits[p])) m.push_back(p);
      emit_remap(m);
      remapped_last = true;
    }
    if (s.pstore != s.pload) {
      // The last map addresses the store layout; ops before it use the load layout. A thread
      // then stores where other threads of the tile loaded, so every load must be done first:
      // a remap's barrier orders them (loads all precede a stage's first remap). A stage with
      // no remap gets one that keeps the map (a round trip through LDS); otherwise a relabel.
      bool any_remap = false;
      for (const auto& t : s.ops) any_remap = any_remap || t.code == TC_REMAP;
      if (!any_remap) {
        emit_remap(map);
      } else if (!remapped_last) {
        TileOpHost t{};
        t.code = TC_SWAPSLOTS;
        t.a = (uint32_t)push_map(map);
        s.ops.push_back(t);
      }
      s.store_map = (int)(s.maps.size() / k) - 1;
    }
    {  // plan-time values of the coefficients (the CPU model reads them)
      std::vector<cd> c;
      stage_coefficients(s, ops, c);
      for (size_t i = 0; i < s.ops.size(); ++i)
        for (int q = 0; q < 4; ++q) s.ops[i].coef[q] = c[4 * i + q];
    }
    stages.push_back(std::move(s));
    remaining = left;
    if (restore_only && nphys != phys) {
      for (uint32_t b = 0; b < n; ++b)
        if (nphys[b] != (int)b) throw Error{QR_ERR_UNSUPPORTED, "internal planner error: the layout was not restored"};
    } else if (restore_only) {
      throw Error{QR_ERR_UNSUPPORTED, "internal planner error: a restoring sweep changed nothing"};
    }
    phys = nphys;
    logi = nlogi;
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
    } else if (t.fold == FOLD_RROT || t.fold == FOLD_XROT) {
      // [[a,-b],[b,a]] (RROT), or [[a,-ib],[-ib,a]] (XROT: the same rotation of (x, i y)),
      // = r * rotation(phi). phi is reduced to [-pi/2, pi/2] (a >= 0, the sign folded out),
      // then al = -tan(phi/2) = -b / (r + a) and be = sin(phi) = b / r, both in [-1, 1].
      double a = m[0].real(), b = t.fold == FOLD_RROT ? m[2].real() : -m[1].imag();
      double sgn = 1.0;
      if (a < 0) {
        a = -a;
        b = -b;
        sgn = -1.0;
      }
      const double r = std::hypot(a, b);
      const double al = r > 0 ? -b / (r + a) : 0.0, be = r > 0 ? b / r : 0.0;
      o[0] = cd(al, 0); o[1] = cd(be, 0); o[2] = cd(0, 0); o[3] = cd(0, 0);
      total *= sgn * r;
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

std::vector<std::vector<Stage>> plan_fused_candidates(uint32_t n, const std::vector<FastOp>& ops,
                                                     const PlanParams& pp, int bytes_per_amp, bool dual,
                                                     int count, std::vector<PlanParams>* chosen) {
  std::vector<std::vector<Stage>> out;
  if (pp.max_ops > 0 || dual || n < kBalanceMinQubits || pp.tile_bits >= (int)n) {
    out.push_back(plan_fused(n, ops, pp, bytes_per_amp, dual));
    if (chosen) chosen->assign(1, pp);
    return out;
  }
  // every cap with the first-candidate tie rule; the caps that balance best with seeded tie
  // rules (greedy choices that tie lead to plans of different sweep counts: 23 to 28 for the
  // 30-qubit brickwork over seeds); and, for large states, seeds that also let near-ties
  // (one op short of the best) tie, whose rare short plans pay for their planning time
  // (the 30-qubit brickwork in 20 sweeps, against 23)
  static const int caps[] = {0, 160, 140, 120, 100, 90, 80};
  static const int seeded_caps[] = {0, 160, 140, 120, 100};
  static const int slack_caps[] = {0, 160};
  constexpr uint32_t kSeeds = 8;
  // (QR_PLAN_NEAR_TIES=0 turns the near-tie search off; complex64, 30-qubit brickwork: 96.2 ms
  // in 23 sweeps with it, 102.2 ms in 25 without)
  bool near_ties = true;
  if (const char* e = std::getenv("QR_PLAN_NEAR_TIES")) near_ties = e[0] != '0';
  const uint32_t slack_seeds = near_ties ? (n >= 28 ? 64u : 16u) : 0u;
  std::vector<PlanParams> jobs;
  for (int cap : caps) {
    PlanParams q = pp;
    q.max_ops = cap;
    jobs.push_back(q);
  }
  for (uint32_t seed = 1; seed < kSeeds; ++seed)
    for (int cap : seeded_caps) {
      PlanParams q = pp;
      q.max_ops = cap;
      q.seed = seed;
      jobs.push_back(q);
    }
  for (uint32_t seed = 1; seed <= slack_seeds; ++seed)
    for (int cap : slack_caps) {
      PlanParams q = pp;
      q.max_ops = cap;
      q.seed = seed;
      q.tie_slack = 1;
      jobs.push_back(q);
    }
  const int nc = (int)jobs.size();
  std::vector<std::vector<Stage>> plans(nc);
  std::vector<std::exception_ptr> errs(nc);
  std::atomic<int> next{0};
  // at most 16 planning threads (a host share per GPU), fewer on smaller hosts
  const int nthreads = std::max(1, std::min<int>({nc, 16, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&] {
      for (int i = next++; i < nc; i = next++) {
        try {
          plans[i] = plan_fused(n, ops, jobs[i], bytes_per_amp, dual);
        } catch (...) {
          errs[i] = std::current_exception();
        }
      }
    });
  for (auto& t : th) t.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
  // ascending estimate; the first of equal estimates (the earlier job) first
  std::vector<int> idx(nc);
  std::vector<double> cost(nc);
  for (int i = 0; i < nc; ++i) {
    idx[i] = i;
    cost[i] = plan_cost(plans[i], bytes_per_amp);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] < cost[b]; });
  // caps that change nothing give the same plan: keep one of each (same gates per stage)
  auto shape = [](const std::vector<Stage>& st) {
    std::vector<int> g;
    for (const Stage& s : st) g.push_back(s.tile ? s.ngates : -1 - s.standalone);
    return g;
  };
  std::vector<std::vector<int>> seen;
  if (chosen) chosen->clear();
  for (int i : idx) {
    if ((int)out.size() >= std::max(1, count)) break;
    std::vector<int> sh = shape(plans[i]);
    if (std::find(seen.begin(), seen.end(), sh) != seen.end()) continue;
    seen.push_back(std::move(sh));
    out.push_back(std::move(plans[i]));
    if (chosen) chosen->push_back(jobs[i]);
  }
  return out;
}

std::vector<Stage> plan_fused_balanced(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                                       int bytes_per_amp, bool dual) {
  return std::move(plan_fused_candidates(n, ops, pp, bytes_per_amp, dual, 1)[0]);
}

}  // namespace qr
