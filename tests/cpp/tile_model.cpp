// tile_model.cpp — CPU model of the fused tile kernel (test infrastructure only).
//
// Runs the engine's own lowering and planner (lower.cpp, planner.cpp) and executes the
// resulting stages with the same per-thread indexing as k_tile in kernels.hip: register
// slots, thread slots, LDS remaps and slot relabels. Lets the CPU suite check the planner
// and the kernel's index arithmetic against the oracle before any GPU run.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "engine.hpp"

using namespace qr;
using cdv = std::complex<double>;

namespace qr {
const char* last_error_cstr();
}

static void apply_standalone(uint32_t n, std::vector<cdv>& psi, const FastOp& op) {
  const uint64_t dim = 1ull << n;
  if (op.kind == FK_DIAG1 || op.kind == FK_DIAGK) {
    for (uint64_t i = 0; i < dim; ++i) {
      if ((i & op.cmask) != op.cmask) continue;
      int l = 0;
      for (int q = 0; q < op.k; ++q) l = (l << 1) | (int)((i >> op.bits[q]) & 1);
      psi[i] *= op.m[l];
    }
    return;
  }
  const int k = op.k, d = 1 << k;
  uint64_t amask = 0;
  for (int q = 0; q < k; ++q) amask |= 1ull << op.bits[q];
  std::vector<cdv> in(d), out(d);
  for (uint64_t base = 0; base < dim; ++base) {
    if (base & amask) continue;
    if ((base & op.cmask) != op.cmask) continue;
    std::vector<uint64_t> off(d);
    for (int j = 0; j < d; ++j) {
      off[j] = base;
      for (int q = 0; q < k; ++q)
        if ((j >> (k - 1 - q)) & 1) off[j] |= 1ull << op.bits[q];
      in[j] = psi[off[j]];
    }
    for (int r = 0; r < d; ++r) {
      out[r] = 0;
      for (int j = 0; j < d; ++j) out[r] += op.m[r * d + j] * in[j];
    }
    for (int j = 0; j < d; ++j) psi[off[j]] = out[j];
  }
}

// One tile stage, modelled thread by thread exactly as k_tile computes indices.
static void run_tile(uint32_t n, std::vector<cdv>& psi, const Stage& s, int RB) {
  const int K = s.k, NT = K - RB;
  const uint64_t ntiles = 1ull << (n - K);
  uint64_t tm = 0;
  for (int b : s.tbits) tm |= 1ull << b;
  std::vector<int> nontile;
  for (uint32_t b = 0; b < n; ++b)
    if (!(tm & (1ull << b))) nontile.push_back((int)b);
  const int T = 1 << NT, R = 1 << RB;
  std::vector<cdv> lds(1u << K);
  struct Ctx {
    uint64_t g0;
    uint32_t l0;
    uint64_t gb[8];
    uint32_t lb[8];
    uint64_t regmask;
  };
  auto setup = [&](const int* mp, uint64_t base, uint32_t tid, Ctx& c) {
    c.g0 = base;
    c.l0 = 0;
    for (int sl = 0; sl < NT; ++sl)
      if ((tid >> sl) & 1) {
        c.g0 |= 1ull << s.tbits[mp[sl]];
        c.l0 |= 1u << mp[sl];
      }
    c.regmask = 0;
    for (int j = 0; j < RB; ++j) {
      c.gb[j] = 1ull << s.tbits[mp[NT + j]];
      c.lb[j] = 1u << mp[NT + j];
      c.regmask |= c.gb[j];
    }
  };
  auto gidx = [&](const Ctx& c, int m) {
    uint64_t x = c.g0;
    for (int j = 0; j < RB; ++j)
      if ((m >> j) & 1) x |= c.gb[j];
    return x;
  };
  auto lidx = [&](const Ctx& c, int m) {
    uint32_t x = c.l0;
    for (int j = 0; j < RB; ++j)
      if ((m >> j) & 1) x |= c.lb[j];
    return x;
  };
  for (uint64_t blk = 0; blk < ntiles; ++blk) {
    uint64_t base = 0;
    for (size_t i = 0; i < nontile.size(); ++i) base |= ((blk >> i) & 1ull) << nontile[i];
    std::vector<Ctx> ctx(T);
    std::vector<std::vector<cdv>> reg(T, std::vector<cdv>(R));
    for (int t = 0; t < T; ++t) {
      setup(&s.maps[0], base, t, ctx[t]);
      for (int m = 0; m < R; ++m) reg[t][m] = psi[gidx(ctx[t], m)];
    }
    for (const TileOpHost& op : s.ops) {
      switch (op.code) {
        case TC_U1: {
          const cdv* u = &s.coefs[op.coef];
          for (int t = 0; t < T; ++t) {
            const uint64_t need = op.cmask & ~ctx[t].g0;
            const bool ok = (need & ~ctx[t].regmask) == 0;
            uint32_t ns = 0;
            for (int j = 0; j < RB; ++j)
              if (need & ctx[t].gb[j]) ns |= 1u << j;
            const int J = (int)op.a;
            for (int m0 = 0; m0 < R; ++m0) {
              if (m0 & (1 << J)) continue;
              if (!ok || ((uint32_t)m0 & ns) != ns) continue;
              const int m1 = m0 | (1 << J);
              const cdv x = reg[t][m0], y = reg[t][m1];
              reg[t][m0] = u[0] * x + u[1] * y;
              reg[t][m1] = u[2] * x + u[3] * y;
            }
          }
          break;
        }
        case TC_D1: {
          const cdv* d = &s.coefs[op.coef];
          for (int t = 0; t < T; ++t) {
            const uint64_t need = op.cmask & ~ctx[t].g0;
            const bool ok = (need & ~ctx[t].regmask) == 0;
            uint32_t ns = 0;
            for (int j = 0; j < RB; ++j)
              if (need & ctx[t].gb[j]) ns |= 1u << j;
            const bool lane_bit = (ctx[t].g0 >> op.a) & 1;
            for (int m = 0; m < R; ++m) {
              if (((uint32_t)m & ns) != ns || !ok) continue;
              const bool bit = (op.b != 0xFFu) ? ((m >> op.b) & 1) : lane_bit;
              reg[t][m] *= bit ? d[1] : d[0];
            }
          }
          break;
        }
        case TC_REMAP: {
          for (int t = 0; t < T; ++t)
            for (int m = 0; m < R; ++m) lds[lidx(ctx[t], m)] = reg[t][m];
          for (int t = 0; t < T; ++t) {
            setup(&s.maps[(size_t)K * op.a], base, t, ctx[t]);
            for (int m = 0; m < R; ++m) reg[t][m] = lds[lidx(ctx[t], m)];
          }
          break;
        }
        case TC_SWAPSLOTS:
          for (int t = 0; t < T; ++t) setup(&s.maps[(size_t)K * op.a], base, t, ctx[t]);
          break;
      }
    }
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < R; ++m) psi[gidx(ctx[t], m)] = reg[t][m];
  }
}

extern "C" {

// Plans `prog` with the engine's planner (tile width k, register bits, low bits) and executes
// the plan on the CPU model. state: 2^n complex128, in/out. Returns 0, or 1 with a message.
int model_run(uint32_t n, double* state, const qr_program* prog, int tile_bits, int low_bits,
              uint64_t* sweeps_out) {
  try {
    auto gates = lower_exact(n, prog);
    auto ops = lower_fast(n, gates, false);
    PlanParams pp;
    pp.tile_bits = std::min<int>(tile_bits, (int)n);
    pp.reg_bits = pp.tile_bits < 4 ? pp.tile_bits : 4;
    pp.low_bits = low_bits;
    auto stages = plan_fused(n, ops, pp, 16, false);
    std::vector<cdv> psi(1ull << n);
    std::memcpy(psi.data(), state, psi.size() * 16);
    for (const Stage& s : stages) {
      if (s.tile) run_tile(n, psi, s, pp.reg_bits);
      else apply_standalone(n, psi, ops[s.standalone]);
    }
    std::memcpy(state, psi.data(), psi.size() * 16);
    if (sweeps_out) *sweeps_out = stages.size();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return 1;
  }
}

const char* model_error() { return last_error_cstr(); }
}
