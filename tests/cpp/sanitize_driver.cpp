// Host-code driver for AddressSanitizer / UndefinedBehaviorSanitizer builds of the engine's
// validation, lowering, observable parsing and planning (lower.cpp, planner.cpp). No GPU is
// touched. Valid random programs go through every planner mode; malformed programs and
// observables (NULL tables, offsets that overflow, out-of-range wires, oversized custom gates)
// must be rejected with an Error, never read out of bounds.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <random>
#include <vector>

#include "engine.hpp"

using namespace qr;

namespace {

int failures = 0;

void expect_error(const char* what, void (*f)(void*), void* arg) {
  try {
    f(arg);
    std::printf("FAIL: %s was accepted\n", what);
    ++failures;
  } catch (const Error&) {
  }
}

struct Prog {
  std::vector<qr_op> ops;
  std::vector<uint32_t> ctl;
  std::vector<double> img;
  std::vector<uint8_t> some;
  qr_program view() const {
    return qr_program{ops.empty() ? nullptr : ops.data(), ops.size(), ctl.empty() ? nullptr : ctl.data(), ctl.size(),
                      img.empty() ? nullptr : img.data(), img.size(), some.empty() ? nullptr : some.data(),
                      some.size()};
  }
};

Prog random_program(uint32_t n, int count, std::mt19937_64& rng) {
  Prog p;
  std::uniform_real_distribution<double> ang(-3.2, 3.2);
  for (int i = 0; i < count; ++i) {
    qr_op op{};
    op.position = (uint32_t)(rng() % n);
    const int pick = (int)(rng() % 26);
    op.kind = (uint32_t)std::min(pick, (int)QR_GATE_CUSTOM);
    op.arg = ang(rng);
    op.arg_int = 1 + (int)(rng() % 5);
    int nc = 0;
    switch (op.kind) {
      case QR_GATE_CR: case QR_GATE_CRK: case QR_GATE_CZ: case QR_GATE_CY: case QR_GATE_CNOT: case QR_GATE_SWAP:
        nc = 1; break;
      case QR_GATE_TOFFOLI: nc = 2; break;
      case QR_GATE_CUSTOM: nc = (int)(rng() % 3); break;
      default: break;
    }
    if (nc + 1 > (int)n) { op.kind = QR_GATE_H; nc = 0; }
    op.ncontrols = (uint32_t)nc;
    op.controls_offset = p.ctl.size();
    std::vector<uint32_t> used{op.position};
    for (int c = 0; c < nc; ++c) {
      uint32_t w;
      do { w = (uint32_t)(rng() % n); } while (std::find(used.begin(), used.end(), w) != used.end());
      used.push_back(w);
      p.ctl.push_back(w);
    }
    if (op.kind == QR_GATE_CUSTOM) {
      const uint64_t dim = 1ull << (nc + 1);
      op.image_offset = p.img.size();
      op.some_offset = p.some.size();
      std::normal_distribution<double> g;
      for (uint64_t q = 0; q < dim * dim * 2; ++q) p.img.push_back(g(rng));
      for (uint64_t q = 0; q < dim; ++q) p.some.push_back((uint8_t)(rng() % 4 != 0));
    }
    p.ops.push_back(op);
  }
  return p;
}

void run_all_planners(uint32_t n, const Prog& p) {
  const qr_program v = p.view();
  validate_program(n, &v);
  auto gates = lower_exact(n, &v);
  auto ops = lower_fast(n, gates, true);
  for (int tb : {2, 5, 10, 12, 13})
    for (int lb : {0, 3, 4})
      for (int split : {0, 1})
        for (int dyn : {0, 1}) {
          PlanParams pp;
          pp.seed = (uint32_t)(dyn * 7 + split);
          pp.tie_slack = split;
          pp.tile_bits = std::min<int>(tb, (int)n);
          pp.reg_bits = std::min(4, pp.tile_bits);
          pp.low_bits = lb;
          pp.lds_split = split;
          pp.dyn_low = dyn;
          auto st = plan_fused(n, ops, pp, 16, false);
          for (const Stage& s : st) {
            std::vector<cd> c;
            if (s.tile) stage_coefficients(s, ops, c);
          }
        }
  auto rev = adjoint_sweep_ops(ops);
  PlanParams pd;
  pd.tile_bits = std::min<int>(10, (int)n);
  pd.reg_bits = std::min(4, pd.tile_bits);
  pd.low_bits = 3;
  (void)plan_fused(n, rev, pd, 16, true);
}

// pauli_passes: every group lands in exactly one pass; a tiled pass spans K bits including
// the low segment bits and its groups' X masks, with lxm the masks in tile coordinates; only
// groups too wide for a tile are read directly.
void check_pauli_passes(std::mt19937_64& rng) {
  for (int trial = 0; trial < 200; ++trial) {
    const uint32_t n = 1 + (uint32_t)(rng() % 34);
    const int prec = (int)(rng() % 2);
    const int G = 1 + (int)(rng() % 120);
    const uint64_t dimmask = n >= 64 ? ~0ull : (1ull << n) - 1;
    std::vector<PauliGroup> g(G);
    for (int i = 0; i < G; ++i) {
      uint64_t xm = 0;
      const int kind = (int)(rng() % 4);
      if (kind == 1) xm = 1ull << (rng() % n);
      if (kind == 2) xm = (1ull << (rng() % n)) | (1ull << (rng() % n));
      if (kind == 3) xm = rng() & dimmask;
      g[i] = PauliGroup{xm, (uint32_t)i, 1u};
    }
    std::vector<PauliGroup> orig = g;
    std::vector<PauliPass> passes(G + 1);
    std::vector<uint32_t> lxm(G);
    const int np = pauli_passes(n, prec, g.data(), G, passes.data(), lxm.data());
    const int K = pauli_tile_bits(prec), c = pauli_low_bits(prec);
    const uint64_t low = (1ull << c) - 1;
    uint32_t next = 0;
    bool ok = np >= 1 && np <= G + 1;
    for (int p = 0; ok && p < np; ++p) {
      ok = passes[p].first == next && passes[p].count > 0;
      next += passes[p].count;
      const uint64_t t = passes[p].tmask;
      if (t != 0) {
        ok = ok && __builtin_popcountll(t) == K && (t & low) == low && (t & ~dimmask) == 0;
        for (uint32_t i = passes[p].first; ok && i < passes[p].first + passes[p].count; ++i) {
          ok = (g[i].xm & ~t) == 0;
          uint32_t lx = 0;
          uint64_t m = t;
          for (int b = 0; b < K; ++b) {
            const uint64_t bit = m & (0ull - m);
            m ^= bit;
            if (g[i].xm & bit) lx |= 1u << b;
          }
          ok = ok && lx == lxm[i];
        }
      } else if ((int)n >= K + 3) {
        for (uint32_t i = passes[p].first; ok && i < passes[p].first + passes[p].count; ++i)
          ok = __builtin_popcountll(g[i].xm & ~low) > K - c;
      }
    }
    ok = ok && next == (uint32_t)G;
    auto key = [](const PauliGroup& a, const PauliGroup& b) { return a.first < b.first; };
    std::sort(g.begin(), g.end(), key);
    for (int i = 0; ok && i < G; ++i) ok = g[i].xm == orig[i].xm;
    if (!ok) {
      std::printf("FAIL: pauli_passes n=%u prec=%d G=%d\n", n, prec, G);
      ++failures;
      return;
    }
  }
}

}  // namespace

int main() {
  std::mt19937_64 rng(20261203);
  // valid programs through every planner mode
  for (int trial = 0; trial < 40; ++trial) {
    const uint32_t n = 1 + (uint32_t)(rng() % 15);
    run_all_planners(n, random_program(n, 60, rng));
  }
  {  // balanced planning of a large state (host only)
    Prog p = random_program(26, 400, rng);
    const qr_program v = p.view();
    auto ops = lower_fast(26, lower_exact(26, &v), false);
    PlanParams pp;
    pp.tile_bits = 12;
    pp.reg_bits = 4;
    pp.low_bits = 3;
    pp.lds_split = 1;
    pp.dyn_low = 1;
    (void)plan_fused_balanced(26, ops, pp, 16, false);
  }
  // malformed programs
  static Prog base;
  base = random_program(5, 10, rng);
  struct Case {
    const char* name;
    void (*mutate)(Prog&, qr_program&);
  };
  const Case cases[] = {
      {"ops NULL with nops", [](Prog&, qr_program& v) { v.ops = nullptr; }},
      {"unknown kind", [](Prog& p, qr_program& v) { p.ops[0].kind = 99; v.ops = p.ops.data(); }},
      {"position out of range", [](Prog& p, qr_program& v) { p.ops[1].position = 5; v.ops = p.ops.data(); }},
      {"controls offset overflow",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_CNOT;
         p.ops[0].ncontrols = 1;
         p.ops[0].controls_offset = ~0ull;
         v.ops = p.ops.data();
       }},
      {"controls NULL",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_TOFFOLI;
         p.ops[0].ncontrols = 2;
         p.ops[0].controls_offset = 0;
         v.ops = p.ops.data();
         v.controls = nullptr;
       }},
      {"custom some offset overflow",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_CUSTOM;
         p.ops[0].ncontrols = 0;
         p.ops[0].some_offset = ~0ull - 1;
         p.ops[0].image_offset = 0;
         p.img.assign(8, 0.0);
         p.some.assign(2, 1);
         v = p.view();
       }},
      {"custom image offset overflow",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_CUSTOM;
         p.ops[0].ncontrols = 0;
         p.ops[0].some_offset = 0;
         p.ops[0].image_offset = ~0ull - 3;
         p.img.assign(8, 0.0);
         p.some.assign(2, 1);
         v = p.view();
       }},
      {"custom too wide",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_CUSTOM;
         p.ops[0].ncontrols = 40;
         p.ops[0].controls_offset = 0;
         p.ctl.assign(40, 1);
         v = p.view();
       }},
      {"custom tables NULL",
       [](Prog& p, qr_program& v) {
         p.ops[0].kind = QR_GATE_CUSTOM;
         p.ops[0].ncontrols = 0;
         v = p.view();
         v.images = nullptr;
         v.some = nullptr;
       }},
  };
  for (const Case& c : cases) {
    static Prog work;
    static qr_program view;
    static const Case* cur;
    work = base;
    view = work.view();
    cur = &c;
    cur->mutate(work, view);
    expect_error(c.name, [](void*) { validate_program(5, &view); (void)lower_exact(5, &view); }, nullptr);
  }
  expect_error("NULL program", [](void*) { validate_program(3, nullptr); }, nullptr);
  expect_error("zero qubits", [](void*) {
    qr_program v{};
    validate_program(0, &v);
  }, nullptr);
  // malformed observables
  static std::vector<qr_pauli_term> terms;
  static std::vector<uint32_t> wires;
  static std::vector<uint8_t> paulis;
  auto obs = []() {
    return qr_observable{terms.data(), terms.size(), wires.empty() ? nullptr : wires.data(),
                         paulis.empty() ? nullptr : paulis.data(), wires.size()};
  };
  static qr_observable o;
  auto reset = [&]() {
    terms = {qr_pauli_term{1.0, 2, 0, 0}, qr_pauli_term{0.5, 1, 0, 2}};
    wires = {0, 1, 2};
    paulis = {'Z', 'Z', 'X'};
    o = obs();
  };
  reset();
  (void)parse_observable(3, &o);
  expect_error("observable NULL", [](void*) { (void)parse_observable(3, nullptr); }, nullptr);
  reset();
  o.terms = nullptr;
  expect_error("terms NULL", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  reset();
  terms[1].offset = ~0ull;
  o = obs();
  expect_error("factor offset overflow", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  reset();
  wires[2] = 3;
  expect_error("wire out of range", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  reset();
  wires[1] = 0;
  expect_error("repeated wire", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  reset();
  paulis[0] = 'Q';
  expect_error("bad pauli", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  reset();
  o.wires = nullptr;
  expect_error("wires NULL", [](void*) { (void)parse_observable(3, &o); }, nullptr);
  check_pauli_passes(rng);
  if (failures) return 1;
  std::printf("sanitize driver: ok\n");
  return 0;
}
