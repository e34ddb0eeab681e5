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

static double imc(cdv l, cdv p) { return l.real() * p.imag() - l.imag() * p.real(); }
static double rec(cdv l, cdv p) { return l.real() * p.real() + l.imag() * p.imag(); }

// One tile stage, modelled thread by thread exactly as k_tile computes indices (maps with
// global bits and tile positions, per-lane / register-slot control split, swizzled LDS).
// With `lam` (the adjoint sweep) both states go through every op and gradient taps add
// sum_j Im(conj(lam_j) (G psi)_j) into grads[param].
// With `cf` (the launch coefficients of stage_coefficients: folded residual matrices, the
// stage's scalar moved onto one op) ops take their coefficients from there, as the kernels do.
static void run_tile(uint32_t n, std::vector<cdv>& psi, const Stage& s, int RB, std::vector<cdv>* lam = nullptr,
                     std::vector<double>* grads = nullptr, const std::vector<cd>* cf = nullptr) {
  const int K = s.k, NT = K - RB;
  const uint64_t ntiles = 1ull << (n - K);
  uint64_t tm = 0;
  for (int b : s.pload) tm |= 1ull << b;
  std::vector<int> nontile;
  for (uint32_t b = 0; b < n; ++b)
    if (!(tm & (1ull << b))) nontile.push_back((int)b);
  const int T = 1 << NT, R = 1 << RB;
  // a stage that stores in another layout than it loads needs a barrier (a remap) between
  // every load and the stores, since threads then store where other threads loaded
  if (s.pstore != s.pload) {
    bool any_remap = false;
    for (const TileOpHost& op : s.ops) any_remap = any_remap || op.code == TC_REMAP;
    if (!any_remap) throw Error{QR_ERR_HIP, "model: layout moves in a stage without a barrier"};
  }
  std::vector<cdv> lds(1u << K);
  std::vector<uint8_t> lds_written(1u << K);
  auto gbit = [&](int mapi, int slot) { return s.phys_bit(mapi, s.maps[(size_t)mapi * K + slot]); };
  for (uint64_t blk = 0; blk < ntiles; ++blk) {
    uint64_t base = 0;
    for (size_t i = 0; i < nontile.size(); ++i) base |= ((blk >> i) & 1ull) << nontile[i];
    int cur_map = 0;
    auto gidx = [&](int t, int m) {
      uint64_t x = base;
      for (int sl = 0; sl < NT; ++sl)
        if ((t >> sl) & 1) x |= 1ull << gbit(cur_map, sl);
      for (int j = 0; j < RB; ++j)
        if ((m >> j) & 1) x |= 1ull << gbit(cur_map, NT + j);
      return x;
    };
    std::vector<std::vector<cdv>> reg(T, std::vector<cdv>(R)), lreg(T, std::vector<cdv>(R));
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < R; ++m) {
        reg[t][m] = psi[gidx(t, m)];
        if (lam) lreg[t][m] = (*lam)[gidx(t, m)];
      }
    std::vector<std::vector<std::vector<cdv>>*> sets{&reg};
    if (lam) sets.push_back(&lreg);
    for (size_t oi = 0; oi < s.ops.size(); ++oi) {
      const TileOpHost& op = s.ops[oi];
      const cd* coef = cf ? &(*cf)[4 * oi] : op.coef;
      if (op.code == TC_GRAD) {
        if (!lam || !grads) throw Error{QR_ERR_HIP, "model: gradient tap in a single-state stage"};
        const int gen = (int)op.flags;
        for (int t = 0; t < T; ++t) {
          const std::vector<cdv>& p = reg[t];
          const std::vector<cdv>& l = lreg[t];
          double acc = 0.0;
          if (gen == 1 || gen == 2) {
            const int J = (int)op.a;
            for (int m0 = 0; m0 < R; ++m0) {
              if (m0 & (1 << J)) continue;
              const int m1 = m0 | (1 << J);
              acc += gen == 1 ? imc(l[m0], p[m1]) + imc(l[m1], p[m0]) : rec(l[m1], p[m0]) - rec(l[m0], p[m1]);
            }
          } else {
            const uint64_t g0 = gidx(t, 0);
            for (int m = 0; m < R; ++m) {
              double v = imc(l[m], p[m]);
              if (gen == 3) {
                const bool bit = (op.b != 0xFFu) ? ((m >> op.b) & 1) : ((g0 >> op.a) & 1);
                if (bit) v = -v;
              }
              acc += v;
            }
          }
          (*grads)[op.param] += acc;
        }
        continue;
      }
      for (auto* rs : sets) {
        std::vector<std::vector<cdv>>& reg = *rs;
        for (int t = 0; t < T && (op.code == TC_U1 || op.code == TC_D1); ++t) {
          const uint64_t g0 = gidx(t, 0);
          const bool lane_ok = (g0 & op.lane_cmask) == op.lane_cmask;
          if (!lane_ok) continue;
          if (op.code == TC_U1) {
            const int J = (int)op.a;
            // folded rotations hold the kernel's shear factors: x += al y, y += be x, x += al y
            // on (x, y) (TF_RROT) or on (x, i y) (TF_XROT)
            const bool shear = cf && (op.b == TF_RROT || op.b == TF_XROT);
            const cdv w = op.b == TF_XROT ? cdv(0, 1) : cdv(1, 0);
            for (int m0 = 0; m0 < R; ++m0) {
              if (m0 & (1 << J)) continue;
              if (((uint32_t)m0 & op.ns) != op.ns) continue;
              const int m1 = m0 | (1 << J);
              const cdv x = reg[t][m0], y = reg[t][m1];
              if (shear) {
                const double al = coef[0].real(), be = coef[1].real();
                cdv xs = x, ys = w * y;
                xs += al * ys;
                ys += be * xs;
                xs += al * ys;
                reg[t][m0] = xs;
                reg[t][m1] = ys / w;
                continue;
              }
              reg[t][m0] = coef[0] * x + coef[1] * y;
              reg[t][m1] = coef[2] * x + coef[3] * y;
            }
          } else {
            for (int m = 0; m < R; ++m) {
              if (((uint32_t)m & op.ns) != op.ns) continue;
              const bool bit = (op.b != 0xFFu) ? ((m >> op.b) & 1) : ((g0 >> op.a) & 1);
              if ((op.flags & 1) && !bit) continue;
              reg[t][m] *= bit ? coef[1] : coef[0];
            }
          }
        }
      }
      if (op.code == TC_SCALE) {
        for (auto* rs : sets)
          for (auto& th : *rs)
            for (auto& v : th) v *= coef[0];
      }
      if (op.code == TC_REMAP) {
        const uint16_t* P = &s.swz[(size_t)op.swz * 2 * K];
        auto lofs = [&](int which, int t, int m) {
          const uint16_t* Q = P + which * K;
          uint32_t x = 0;
          for (int sl = 0; sl < NT; ++sl)
            if ((t >> sl) & 1) x ^= Q[sl];
          for (int j = 0; j < RB; ++j)
            if ((m >> j) & 1) x ^= Q[NT + j];
          return x;
        };
        if (op.flags & 1) {
          // split remap: two rounds through half the LDS (the compiled kernels' exchange)
          const int h = (int)op.b;
          int osl = -1, nsl = -1;
          for (int sl = 0; sl < K; ++sl) {
            if (s.maps[(size_t)cur_map * K + sl] == h) osl = sl;
            if (s.maps[(size_t)op.a * K + sl] == h) nsl = sl;
          }
          if (osl < 0 || nsl < 0) throw Error{QR_ERR_HIP, "model: split position not in the maps"};
          const uint32_t hm = (1u << (K - 1)) - 1;
          auto side = [&](int slot, int t, int m) {  // which round amplitude (t, m) belongs to
            return slot < NT ? (int)((t >> slot) & 1) : (int)((m >> (slot - NT)) & 1);
          };
          for (auto* rs : sets) {
            std::vector<std::vector<cdv>>& reg = *rs;
            std::vector<std::vector<cdv>> tmpv(T, std::vector<cdv>(R));
            std::vector<std::vector<uint8_t>> used(T, std::vector<uint8_t>(R, 0));
            for (int rho = 0; rho < 2; ++rho) {
              std::fill(lds_written.begin(), lds_written.end(), 0);
              for (int t = 0; t < T; ++t)
                for (int m = 0; m < R; ++m) {
                  if (side(osl, t, m) != rho) continue;
                  const uint32_t at = lofs(0, t, m) & hm;
                  if (lds_written[at]) throw Error{QR_ERR_HIP, "model: LDS write collision (split remap)"};
                  lds_written[at] = 1;
                  lds[at] = reg[t][m];
                }
              for (int t = 0; t < T; ++t)
                for (int m = 0; m < R; ++m) {
                  if (side(nsl, t, m) != rho) continue;
                  const uint32_t at = lofs(1, t, m) & hm;
                  if (!lds_written[at]) throw Error{QR_ERR_HIP, "model: read of an unwritten LDS slot (split remap)"};
                  if (rho == 0 && split_read_needs_tmp(osl, nsl, NT, m)) {
                    tmpv[t][m] = lds[at];
                    used[t][m] = 1;
                  } else {
                    reg[t][m] = lds[at];
                  }
                }
            }
            for (int t = 0; t < T; ++t)
              for (int m = 0; m < R; ++m)
                if (used[t][m]) reg[t][m] = tmpv[t][m];
          }
          cur_map = (int)op.a;
          continue;
        }
        for (auto* rs : sets) {
          std::vector<std::vector<cdv>>& reg = *rs;
          std::fill(lds_written.begin(), lds_written.end(), 0);
          for (int t = 0; t < T; ++t)
            for (int m = 0; m < R; ++m) {
              const uint32_t at = lofs(0, t, m);
              if (at >= lds.size() || lds_written[at]) throw Error{QR_ERR_HIP, "model: LDS write collision"};
              lds_written[at] = 1;
              lds[at] = reg[t][m];
            }
          for (int t = 0; t < T; ++t)
            for (int m = 0; m < R; ++m) reg[t][m] = lds[lofs(1, t, m)];
        }
        cur_map = (int)op.a;
      } else if (op.code == TC_SWAPSLOTS) {
        cur_map = (int)op.a;
      }
    }
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < R; ++m) {
        psi[gidx(t, m)] = reg[t][m];
        if (lam) (*lam)[gidx(t, m)] = lreg[t][m];
      }
  }
}

extern "C" {

// Plans `prog` with the engine's planner (tile width k, register bits, low bits) and executes
// the plan on the CPU model. state: 2^n complex128, in/out. Returns 0, or 1 with a message.
int model_run(uint32_t n, double* state, const qr_program* prog, int tile_bits, int low_bits, int reg_bits,
              uint64_t* sweeps_out) {
  try {
    auto gates = lower_exact(n, prog);
    auto ops = lower_fast(n, gates, false);
    PlanParams pp;
    pp.tile_bits = std::min<int>(tile_bits, (int)n);
    pp.reg_bits = pp.tile_bits < 4 ? pp.tile_bits : 4;
    if ((reg_bits == 5 || reg_bits == 3) && pp.tile_bits >= 10) pp.reg_bits = reg_bits;
    pp.low_bits = low_bits;
    if (pp.tile_bits == (int)n) pp.max_ops = kWholeStateStageOps;
    auto stages = plan_fused(n, ops, pp, 16, false);
    std::vector<cdv> psi(1ull << n);
    std::memcpy(psi.data(), state, psi.size() * 16);
    std::vector<cd> cf;
    for (const Stage& s : stages) {
      if (s.tile) {
        stage_coefficients(s, ops, cf);
        run_tile(n, psi, s, pp.reg_bits, nullptr, nullptr, &cf);
      } else {
        apply_standalone(n, psi, ops[s.standalone]);
      }
    }
    std::memcpy(state, psi.data(), psi.size() * 16);
    if (sweeps_out) *sweeps_out = stages.size();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return 1;
  }
}

// model_run with every plan option: remaps split into two rounds through half the LDS
// (PlanParams::lds_split) and the moving layout (PlanParams::dyn_low).
int model_run_opts(uint32_t n, double* state, const qr_program* prog, int tile_bits, int low_bits, int reg_bits,
                   int max_ops, int lds_split, int dyn_low, uint32_t seed, int tie_slack, uint64_t* sweeps_out) {
  try {
    auto ops = lower_fast(n, lower_exact(n, prog), false);
    PlanParams pp;
    pp.tile_bits = std::min<int>(tile_bits, (int)n);
    pp.reg_bits = std::min(reg_bits, pp.tile_bits);
    pp.low_bits = low_bits;
    pp.max_ops = max_ops;
    pp.lds_split = lds_split;
    pp.dyn_low = dyn_low;
    pp.seed = seed;
    pp.tie_slack = tie_slack;
    auto stages = plan_fused(n, ops, pp, 16, false);
    std::vector<cdv> psi(1ull << n);
    std::memcpy(psi.data(), state, psi.size() * 16);
    std::vector<cd> cf;
    for (const Stage& s : stages) {
      if (s.tile) {
        stage_coefficients(s, ops, cf);
        run_tile(n, psi, s, pp.reg_bits, nullptr, nullptr, &cf);
      } else {
        apply_standalone(n, psi, ops[s.standalone]);
      }
    }
    std::memcpy(state, psi.data(), psi.size() * 16);
    if (sweeps_out) *sweeps_out = stages.size();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return 1;
  }
}

// model_run with a cap on the ops per tile stage (the planner's max_ops), register bits 4.
int model_run_capped(uint32_t n, double* state, const qr_program* prog, int tile_bits, int low_bits, int max_ops,
                     uint64_t* sweeps_out) {
  try {
    auto ops = lower_fast(n, lower_exact(n, prog), false);
    PlanParams pp;
    pp.tile_bits = std::min<int>(tile_bits, (int)n);
    pp.reg_bits = std::min(4, pp.tile_bits);
    pp.low_bits = low_bits;
    pp.max_ops = max_ops;
    auto stages = plan_fused(n, ops, pp, 16, false);
    std::vector<cdv> psi(1ull << n);
    std::memcpy(psi.data(), state, psi.size() * 16);
    std::vector<cd> cf;
    for (const Stage& s : stages) {
      if (s.tile) {
        stage_coefficients(s, ops, cf);
        run_tile(n, psi, s, pp.reg_bits, nullptr, nullptr, &cf);
      } else {
        apply_standalone(n, psi, ops[s.standalone]);
      }
    }
    std::memcpy(state, psi.data(), psi.size() * 16);
    if (sweeps_out) *sweeps_out = stages.size();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return 1;
  }
}

// The adjoint sweep on the CPU model: psi = U|0>, lam = O psi given; plans the reverse
// program with the engine's dual-state planner and replays it. grads: one per parameter.
int model_adjoint(uint32_t n, double* psi_io, double* lam_io, const qr_program* prog, int tile_bits, int low_bits,
                  double* grads_out, uint64_t nparams, uint64_t* sweeps_out) {
  try {
    auto gates = lower_exact(n, prog);
    auto ops = lower_fast(n, gates, true);
    uint64_t np = 0;
    for (const auto& op : ops)
      if (op.param >= 0) ++np;
    if (np != nparams) throw Error{QR_ERR_ARG, "model: parameter count mismatch"};
    auto rev = adjoint_sweep_ops(ops);
    PlanParams pp;
    pp.tile_bits = std::min<int>(tile_bits, (int)n);
    pp.reg_bits = std::min(4, pp.tile_bits);
    pp.low_bits = low_bits;
    if (pp.tile_bits == (int)n) pp.max_ops = kWholeStateStageOps;
    auto stages = plan_fused(n, rev, pp, 16, true);
    std::vector<cdv> psi(1ull << n), lam(1ull << n);
    std::memcpy(psi.data(), psi_io, psi.size() * 16);
    std::memcpy(lam.data(), lam_io, lam.size() * 16);
    std::vector<double> g(np, 0.0);
    for (const Stage& s : stages) {
      if (s.tile) {
        run_tile(n, psi, s, pp.reg_bits, &lam, &g);
      } else {
        if (rev[s.standalone].kind == FK_GRAD) throw Error{QR_ERR_HIP, "model: standalone gradient tap"};
        apply_standalone(n, psi, rev[s.standalone]);
        apply_standalone(n, lam, rev[s.standalone]);
      }
    }
    for (const auto& op : ops)
      if (op.param >= 0) grads_out[op.param] = op.gen == 4 ? -g[op.param] : g[op.param];
    std::memcpy(psi_io, psi.data(), psi.size() * 16);
    std::memcpy(lam_io, lam.data(), lam.size() * 16);
    if (sweeps_out) *sweeps_out = stages.size();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return 1;
  }
}

const char* model_error() { return last_error_cstr(); }
}
