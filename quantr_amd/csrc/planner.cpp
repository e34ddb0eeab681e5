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
#include <cstring>

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
  return op.kind == FK_DENSE1 || op.kind == FK_DIAG1 || (k >= 2 && is_swap_perm(op));
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

}  // namespace

std::vector<Stage> plan_fused(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                              int bytes_per_amp, bool /*dual*/) {
  std::vector<Stage> stages;
  const int N = (int)ops.size();
  std::vector<Masks> mk(N);
  std::vector<uint8_t> eligible(N);
  const int k = std::min<int>(pp.tile_bits, (int)n);
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
    uint64_t R = greedy_bits(0, r, nonlow_bits, rem, mk, eligible);
    for (int b : nonlow_bits)
      if (__builtin_popcountll(R) < r) R |= 1ull << b;
    map = make_map(R);
    push_map(map);  // map 0 = load map

    auto slot_of_pos = [&](int p) {
      for (int sl = 0; sl < k; ++sl)
        if (map[sl] == p) return sl;
      return -1;
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
          if (ok && op.kind == FK_DENSE1) ok = (t & ~regs) == 0;
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
        t.cmask = op.cmask;
        if (op.kind == FK_DENSE1) {
          t.code = TC_U1;
          const int p = pos_of_bit[op.bits[0]];
          t.a = (uint32_t)(slot_of_pos(p) - nts);
          t.coef = (uint32_t)s.coefs.size();
          s.coefs.insert(s.coefs.end(), op.m.begin(), op.m.begin() + 4);
          int form = op.form;
          if (form == U1_PERM && op.m[1] == cd(1, 0) && op.m[2] == cd(1, 0)) form = 3;  // X
          t.b = (uint32_t)form;
          s.ops.push_back(t);
        } else if (op.kind == FK_DIAG1) {
          t.code = TC_D1;
          t.a = (uint32_t)op.bits[0];
          const int p = pos_of_bit[op.bits[0]];
          const int sl = p >= 0 ? slot_of_pos(p) : -1;
          t.b = (sl >= nts) ? (uint32_t)(sl - nts) : 0xFFu;
          t.coef = (uint32_t)s.coefs.size();
          s.coefs.push_back(op.m[0]);
          s.coefs.push_back(op.m[1]);
          t.c = (op.m[0] == cd(1, 0)) ? 1u : 0u;  // only the |1> half changes
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
      uint64_t R2 = greedy_bits(0, r, tile_bits_vec, rem, mk, eligible);
      for (int b : tile_bits_vec)
        if (__builtin_popcountll(R2) < r) R2 |= 1ull << b;
      if (now.empty() && count_absorb(rem, mk, eligible, R2) == 0)
        throw Error{QR_ERR_UNSUPPORTED, "internal planner error: no progress inside a tile"};
      std::vector<int> nm = make_map(R2);
      TileOpHost t{};
      t.code = TC_REMAP;
      t.a = (uint32_t)push_map(nm);
      s.ops.push_back(t);
      map = nm;
    }
    // ---- final layout: low tile positions on the low thread slots for a coalesced store --
    bool canon = true;
    for (int sl = 0; sl < c; ++sl)
      if (map[sl] != sl) canon = false;
    if (!canon) {
      uint64_t Rf = 0;
      for (int b : nonlow_bits)
        if (__builtin_popcountll(Rf) < r) Rf |= 1ull << b;
      std::vector<int> nm = make_map(Rf);
      TileOpHost t{};
      t.code = TC_REMAP;
      t.a = (uint32_t)push_map(nm);
      s.ops.push_back(t);
      map = nm;
    }
    stages.push_back(std::move(s));
    remaining = left;
  }
  return stages;
}

}  // namespace qr
