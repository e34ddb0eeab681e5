// runtime.cpp — device states, fused-plan packing and caching, execution of programs,
// expectation values and the fused adjoint sweep. Entry points live in capi.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace qr {


// Default tile shape: 2^12 complex128 / 2^13 complex64 amplitudes (64 KB of LDS, two
// workgroups per CU), 16 amplitudes per thread.
int pick_tile_bits(const qr_state* st) {
  int k = st->prec == 0 ? 12 : 13;
  if (st->opt.tile_bits) k = std::min<int>(tile_max_bits(st->prec), (int)st->opt.tile_bits);
  return std::min<int>(k, (int)st->n);
}

// Ops per tile stage at most: the caller's stage_ops, else kWholeStateStageOps when one tile
// covers the whole state, else no limit.
static int stage_op_limit(const qr_state* st, int tile_bits) {
  int lim = tile_bits == (int)st->n ? kWholeStateStageOps : 0;
  if (st->opt.stage_ops) lim = lim ? std::min<int>(lim, (int)st->opt.stage_ops) : (int)st->opt.stage_ops;
  return lim;
}

PlanParams plan_params(const qr_state* st) { return plan_params(st->n, st->prec, st->opt); }

PlanParams plan_params(uint32_t n, int prec, const qr_options& opt) {
  qr_state shape;  // the fields the tile choices read (no device resources)
  shape.n = n;
  shape.prec = prec;
  shape.dim = 1ull << n;
  shape.opt = opt;
  const qr_state* st = &shape;
  PlanParams pp;
  pp.tile_bits = pick_tile_bits(st);
  pp.reg_bits = tile_reg_bits(pp.tile_bits);
  if ((st->opt.reg_bits == 5 || st->opt.reg_bits == 3) && pp.tile_bits >= 10) pp.reg_bits = (int)st->opt.reg_bits;
  pp.low_bits = st->opt.low_bits ? (int)st->opt.low_bits : (st->prec == 0 ? 3 : 4);
  pp.max_ops = stage_op_limit(st, pp.tile_bits);
  // tiles of 64 KB or more go through LDS in two halves: three complex128 workgroups of
  // 2^12 amplitudes then share a CU instead of two (measured 7.5 against 7.8 ms per 30-qubit
  // sweep); QR_LDS_SPLIT=0 or 1 overrides
  // (complex64 tiles of 2^13 run as fast without: 102.0 against 104.3 ms per 30-qubit step)
  pp.lds_split = (st->prec == 0 && (1ull << pp.tile_bits) * st->bytes_per_amp() >= 65536) ? 1 : 0;
  if (const char* e = std::getenv("QR_LDS_SPLIT")) pp.lds_split = e[0] == '1' ? 1 : 0;
  // moving layout: each stage stores at the physical low bits (the contiguous segments of
  // every tile) bits the next stage uses, instead of always the same wires; QR_DYN_LOW=0 keeps
  // the identity layout
  pp.dyn_low = 1;
  if (const char* e = std::getenv("QR_DYN_LOW"))
    if (e[0] == '0') pp.dyn_low = 0;
  return pp;
}

// Device copy of a plan: every tile stage's ops, coefficients, slot maps and LDS patterns,
// and the standalone ops' tables, in one upload.
std::vector<uint8_t> pack_blob(const std::vector<FastOp>& ops, const std::vector<Stage>& stages, std::vector<Loc>& loc) {
  std::vector<uint8_t> blob;
  auto align = [&](size_t a) {
    while (blob.size() % a) blob.push_back(0);
  };
  auto append = [&](const void* p, size_t bytes) {
    size_t off = blob.size();
    blob.resize(off + bytes);
    if (bytes) std::memcpy(blob.data() + off, p, bytes);
    return off;
  };
  loc.assign(stages.size(), Loc{});
  for (size_t si = 0; si < stages.size(); ++si) {
    const Stage& s = stages[si];
    if (s.tile) {
      align(64);
      std::vector<TileOpDev> d(s.ops.size());
      std::vector<double2> c(s.ops.size() * 4);
      for (size_t i = 0; i < s.ops.size(); ++i) {
        const TileOpHost& h = s.ops[i];
        TileOpDev t{};
        const bool maps_op = h.code == TC_REMAP || h.code == TC_SWAPSLOTS;
        t.w0 = (h.code & 0xFF) | ((maps_op ? 0u : (h.a & 0xFF)) << 8) | ((h.b & 0xFF) << 16) | ((h.flags & 0xFF) << 24);
        t.ns = h.ns;
        t.lane_cmask = h.lane_cmask;
        t.map = maps_op ? h.a : 0;
        t.swz = h.swz;
        d[i] = t;
      }
      // the current program's coefficients (plans are shared by structure), folded scalars
      // applied; gradient taps carry their parameter row
      std::vector<cd> cv;
      stage_coefficients(s, ops, cv);
      for (size_t i = 0; i < cv.size(); ++i) c[i] = make_double2(cv[i].real(), cv[i].imag());
      loc[si].ops = append(d.data(), d.size() * sizeof(TileOpDev));
      align(16);
      loc[si].coefs = append(c.data(), c.size() * sizeof(double2));
      align(4);
      std::vector<uint32_t> m(s.maps.size());
      for (size_t i = 0; i < m.size(); ++i)  // physical bit | tile position << 8
        m[i] = (uint32_t)s.phys_bit((int)(i / s.k), s.maps[i]) | ((uint32_t)s.maps[i] << 8);
      loc[si].maps = append(m.data(), m.size() * sizeof(uint32_t));
      std::vector<uint32_t> z(s.swz.begin(), s.swz.end());
      if (z.empty()) z.push_back(0);
      loc[si].off = append(z.data(), z.size() * sizeof(uint32_t));
    } else {
      const FastOp& op = ops[s.standalone];
      if (op.kind == FK_DENSEK || op.kind == FK_DIAGK) {
        align(16);
        std::vector<double2> c(op.m.size());
        for (size_t i = 0; i < c.size(); ++i) c[i] = make_double2(op.m[i].real(), op.m[i].imag());
        loc[si].coefs = append(c.data(), c.size() * sizeof(double2));
        align(8);
        const int dim = 1 << op.k;
        std::vector<uint64_t> off(dim, 0);
        for (int j = 0; j < dim; ++j)
          for (int q = 0; q < op.k; ++q)
            if ((j >> (op.k - 1 - q)) & 1) off[j] |= 1ull << op.bits[q];
        loc[si].off = append(off.data(), off.size() * sizeof(uint64_t));
      }
    }
  }
  return blob;
}

void pack_plan(qr_state* st, const std::vector<FastOp>& ops, PackedPlan& P) {
  const std::vector<uint8_t> blob = pack_blob(ops, P.stages, P.loc);
  P.dblob = upload(st, st->aux, blob);
}

TileArgs tile_args_at(uint32_t n, const Stage& s, const uint8_t* dblob, const Loc& l) {
  TileArgs a{};
  a.ops = (const TileOpDev*)(dblob + l.ops);
  a.coefs = (const double2*)(dblob + l.coefs);
  a.maps = (const uint32_t*)(dblob + l.maps);
  a.swz = (const uint32_t*)(dblob + l.off);
  a.nops = (uint32_t)s.ops.size();
  a.ntiles = 1ull << (n - s.k);
  uint64_t tm = 0;
  for (int b : s.pload) tm |= 1ull << b;  // physical bits of the tile (the same at the store)
  a.nontile_mask = (n >= 64 ? ~0ull : ((1ull << n) - 1)) & ~tm;
  return a;
}

TileArgs tile_args(const qr_state* st, const PackedPlan& P, size_t si) {
  return tile_args_at(st->n, P.stages[si], P.dblob, P.loc[si]);
}

void launch_standalone_raw(int prec, uint64_t dim, void* amp, DevBuf& tmp, const FastOp& op, const uint8_t* dblob,
                           const Loc& l, hipStream_t s) {
  switch (op.kind) {
    case FK_DENSE1: {
      double2 u[4];
      for (int i = 0; i < 4; ++i) u[i] = make_double2(op.m[i].real(), op.m[i].imag());
      hip_check(launch_dense1(prec, amp, dim / 2, op.bits[0], op.cmask, u, s), "dense1");
      break;
    }
    case FK_DIAG1:
      hip_check(launch_diag1(prec, amp, dim, op.bits[0], op.cmask, make_double2(op.m[0].real(), op.m[0].imag()),
                             make_double2(op.m[1].real(), op.m[1].imag()), s),
                "diag1");
      break;
    case FK_DENSEK: {
      DenseKArgs a{};
      a.k = op.k;
      a.cmask = op.cmask;
      std::vector<int> sb(op.bits, op.bits + op.k);
      std::sort(sb.begin(), sb.end());
      for (int q = 0; q < op.k; ++q) a.sorted_bits[q] = (uint8_t)sb[q];
      const uint64_t dimk = 1ull << op.k;
      const uint64_t ngroups = dim >> op.k;
      if (op.k >= 2 && op.k <= 4) {  // the group fits in registers
        hip_check(launch_densek_reg(prec, amp, ngroups, a, (const double2*)(dblob + l.coefs),
                                    (const uint64_t*)(dblob + l.off), s),
                  "densek");
        break;
      }
      uint64_t tg = std::min<uint64_t>(ngroups, std::max<uint64_t>(256, (64ull << 20) / (dimk * 16)));
      tg = std::max<uint64_t>(tg / 256, 1) * 256;
      if (ngroups < 256) tg = ngroups;
      tmp.reserve(tg * dimk * (prec == 0 ? 16 : 8));
      hip_check(launch_densek(prec, amp, tmp.p, tg, ngroups, a, (const double2*)(dblob + l.coefs),
                              (const uint64_t*)(dblob + l.off), s),
                "densek");
      break;
    }
    case FK_DIAGK: {
      DiagKArgs a{};
      a.k = op.k;
      a.cmask = op.cmask;
      for (int q = 0; q < op.k; ++q) a.bits[q] = (uint8_t)op.bits[q];
      hip_check(launch_diagk(prec, amp, dim, a, (const double2*)(dblob + l.coefs), s), "diagk");
      break;
    }
    default:
      throw Error{QR_ERR_UNSUPPORTED, "internal error: op kind has no standalone kernel"};
  }
}

// One standalone (unfused) op on `amp`, a state of st's size, on st's stream.
void launch_standalone(qr_state* st, void* amp, const FastOp& op, const PackedPlan& P, size_t si) {
  static const char* names[] = {"dense1", "diag1", "densek", "diagk"};
  ProfScope ps_(st, op.kind >= 0 && op.kind < 4 ? names[op.kind] : "standalone", P.stages[si].bytes);
  launch_standalone_raw(st->prec, st->dim, amp, st->tmp, op, P.dblob, P.loc[si], st->stream);
}

// ---- plan cache ---------------------------------------------------------------------------
// Plans depend on the program's structure only (plan_key_append); parameter sweeps, gradient
// loops and serving re-apply one structure with new coefficients, so plans and their
// compiled stages are kept per process (least recently used dropped beyond 64).
// The balanced planner's search (tens of plans) is remembered on disk beside the compiled
// stages: the winning cap and tie rule per structure, so a later process plans once.
static std::vector<Stage> plan_balanced_cached(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                                               int B, bool dual, const std::vector<uint64_t>& key) {
  const bool searches = !(pp.max_ops > 0 || dual || n < kBalanceMinQubits || pp.tile_bits >= (int)n);
  const std::string dir = searches ? jit_cache_dir() : std::string();
  std::string path;
  if (!dir.empty()) {
    std::string bytes((const char*)key.data(), key.size() * sizeof(uint64_t));
    path = dir + "/plan1-" + cache_key_hex(bytes) + ".txt";
    std::ifstream f(path);
    long long cap = -1, seed = -1, slack = -1;
    if (f >> cap >> seed >> slack && cap >= 0 && cap <= 1000000 && seed >= 0 && seed < (1ll << 32) && slack >= 0 &&
        slack <= 8) {
      PlanParams q = pp;
      q.max_ops = (int)cap;
      q.seed = (uint32_t)seed;
      q.tie_slack = (int)slack;
      if (const char* e = std::getenv("QR_JIT_CACHE_LOG"))
        if (e[0] == '1') std::fprintf(stderr, "[qr plan] plan choice from the disk cache\n");
      return plan_fused(n, ops, q, B, dual);
    }
  }
  if (!path.empty())
    if (const char* e = std::getenv("QR_JIT_CACHE_LOG"))
      if (e[0] == '1') std::fprintf(stderr, "[qr plan] searching plans\n");
  std::vector<PlanParams> chosen;
  std::vector<std::vector<Stage>> c = plan_fused_candidates(n, ops, pp, B, dual, 1, &chosen);
  if (!path.empty() && !chosen.empty())
    (void)cache_write(path, std::to_string(chosen[0].max_ops) + " " + std::to_string(chosen[0].seed) + " " +
                                std::to_string(chosen[0].tie_slack) + "\n");
  return std::move(c[0]);
}

static std::mutex g_plan_mu;
static std::map<std::vector<uint64_t>, std::shared_ptr<CachedPlan>> g_plans;
static uint64_t g_plan_clock = 0;

std::shared_ptr<CachedPlan> get_plan(const qr_state* st, const std::vector<FastOp>& ops, const PlanParams& pp,
                                     bool fused, bool dual, bool persistent) {
  return get_plan_for(st->n, st->prec, ops, pp, fused, dual, persistent);
}

std::shared_ptr<CachedPlan> get_plan_for(uint32_t n, int prec, const std::vector<FastOp>& ops, const PlanParams& pp,
                                         bool fused, bool dual, bool persistent) {
  std::vector<uint64_t> key{n, (uint64_t)prec, (uint64_t)pp.tile_bits, (uint64_t)pp.reg_bits,
                            (uint64_t)pp.low_bits, (uint64_t)pp.max_ops, (uint64_t)fused, (uint64_t)dual,
                            (uint64_t)persistent, (uint64_t)pp.lds_split, (uint64_t)pp.dyn_low};
  key.reserve(8 + 3 * ops.size());
  for (const FastOp& op : ops) plan_key_append(op, key);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
      it->second->stamp = ++g_plan_clock;
      return it->second;
    }
  }
  auto cp = std::make_shared<CachedPlan>();
  const int B = prec == 0 ? 16 : 8;
  if (fused) {
    cp->stages = plan_balanced_cached(n, ops, pp, B, dual, key);
  } else {
    for (size_t i = 0; i < ops.size(); ++i) {
      Stage s;
      s.tile = false;
      s.standalone = (int)i;
      s.bytes = 2.0 * (double)(1ull << n) * B * ops[i].f_touch;
      s.ngates = 1;
      cp->stages.push_back(std::move(s));
    }
  }
  std::lock_guard<std::mutex> lk(g_plan_mu);
  cp->stamp = ++g_plan_clock;
  g_plans[key] = cp;
  while (g_plans.size() > 64) {
    auto victim = g_plans.begin();
    for (auto it = g_plans.begin(); it != g_plans.end(); ++it)
      if (it->second->stamp < victim->second->stamp) victim = it;
    g_plans.erase(victim);
  }
  return cp;
}

// Compiles (once) the tile stages of a cached plan and resolves their kernels. Only this
// plan's lock is held while compiling, so other threads keep finding and using other plans;
// threads that use the plan take the same lock before reading its kernels.
void prepare_compiled(CachedPlan& cp, int prec, int RB, bool persistent, bool dual) {
  std::lock_guard<std::mutex> lk(cp.compile_mu);
  if (!cp.fns.empty() || cp.stages.empty()) return;
  cp.src = jit_stage_sources(cp.stages, prec, RB, persistent, dual);
  jit_prepare_without_spills(cp.src);
  std::vector<hipFunction_t> f(cp.stages.size(), nullptr);
  for (size_t i = 0; i < cp.stages.size(); ++i)
    if (cp.stages[i].tile) f[i] = jit_function(cp.src[i]);
  cp.fns = std::move(f);
}

// Candidate plans a tuning structure measures (the planner's lowest estimates).
constexpr int kTuneCandidates = 3;

namespace {

// Events around every launch of one measuring application, destroyed with it.
struct StageTimer {
  std::vector<hipEvent_t> ev;
  explicit StageTimer(size_t stages) : ev(2 * stages, nullptr) {
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  ~StageTimer() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  StageTimer(const StageTimer&) = delete;
  StageTimer& operator=(const StageTimer&) = delete;
};

// The tuning record of a structure, made on first use: candidate 0 is `cp` itself.
std::shared_ptr<TuneState> tune_state(qr_state* st, const std::shared_ptr<CachedPlan>& cp,
                                      const std::vector<FastOp>& ops, const PlanParams& pp) {
  std::lock_guard<std::mutex> lk(cp->tune_mu);
  if (cp->tune) return cp->tune;
  auto ts = std::make_shared<TuneState>();
  ts->cands.push_back(cp);
  std::vector<std::vector<Stage>> alt =
      plan_fused_candidates(st->n, ops, pp, (int)st->bytes_per_amp(), false, kTuneCandidates);
  for (size_t i = 1; i < alt.size(); ++i) {
    auto c = std::make_shared<CachedPlan>();
    c->stages = std::move(alt[i]);
    ts->cands.push_back(std::move(c));
  }
  ts->cand_ms.assign(ts->cands.size(), 0.0);
  ts->stage_ms.resize(ts->cands.size());
  cp->tune = ts;
  return ts;
}

}  // namespace

// Plans (fused tile sweeps, or one sweep per op), uploads and launches a fast-form program.
void run_fast(qr_state* st, const std::vector<FastOp>& ops, bool fused, qr_plan_stats* stats) {
  const PlanParams pp = plan_params(st);
  const bool persistent = st->opt.tile_variant == 2;
  std::shared_ptr<CachedPlan> cp = get_plan(st, ops, pp, fused, false, persistent);
  // ---- measured selection (qr_options.tune): pick this application's plan and tile orders
  std::shared_ptr<TuneState> ts;
  std::unique_lock<std::mutex> tune_lock;
  int step = -1;  // >= 0: this application measures
  const std::vector<uint32_t>* orders = nullptr;
  std::shared_ptr<CachedPlan> use = cp;
  if (fused && st->opt.tune && st->opt.jit != 0 && !persistent) {
    ts = tune_state(st, cp, ops, pp);
    tune_lock = std::unique_lock<std::mutex>(ts->mu);
    const int C = (int)ts->cands.size();
    if (ts->step <= C) {
      step = ts->step;
      use = ts->cands[step < C ? step : ts->chosen];
    } else {
      use = ts->cands[ts->chosen];
      orders = &ts->order;
      tune_lock.unlock();
    }
  }
  PackedPlan P;
  P.stages = use->stages;
  std::vector<Stage>& stages = P.stages;
  pack_plan(st, ops, P);
  // ---- runtime-specialised stages: always with jit = 1, from 28 qubits or from the second
  // application of a structure with jit = 2, and for every measured plan
  bool use_jit = false;
  if (fused && st->opt.jit != 0) {
    use_jit = st->opt.jit == 1 || st->n >= 28 || cp->runs > 0 || ts != nullptr;
    if (use_jit) prepare_compiled(*use, st->prec, pp.reg_bits, persistent, false);
  }
  cp->runs++;
  std::unique_ptr<StageTimer> timer;
  if (step >= 0) timer = std::make_unique<StageTimer>(stages.size());
  // tiles dealt to the XCDs in contiguous eighths by default (faster on most sweeps of the
  // 30-qubit brickwork); QR_TILE_ORDER=0 deals them in index order. Measuring applications
  // try the default, then the other order.
  uint32_t default_order = 1u;
  if (const char* e = std::getenv("QR_TILE_ORDER")) default_order = e[0] == '0' ? 0u : 1u;
  const uint32_t measure_order =
      step >= 0 && step == (int)ts->cands.size() ? 1u - default_order : default_order;
  // ---- launch --------------------------------------------------------------------------
  qr_plan_stats ps{};
  for (size_t si = 0; si < stages.size(); ++si) {
    const Stage& s = stages[si];
    ps.sweeps++;
    ps.bytes += s.bytes;
    if (timer) hip_check(hipEventRecord(timer->ev[2 * si], st->stream), "hipEventRecord");
    if (s.tile) {
      ps.tile_sweeps++;
      const TileArgs a = tile_args(st, P, si);
      ProfScope ps_(st, "tile_sweep", s.bytes);
      if (use_jit) {
        uint32_t order = step >= 0 ? measure_order : (orders ? (*orders)[si] : default_order);
        if (persistent || (a.ntiles & 7) != 0) order = 0;
        ps.xcd_order_sweeps += order;
        hip_check(jit_launch(use->fns[si], s.k, pp.reg_bits, persistent, st->amp, a.coefs, a.ntiles,
                             a.nontile_mask, st->stream, order),
                  "compiled tile kernel");
      } else {
        hip_check(launch_tile(st->prec, s.k, pp.reg_bits, st->opt.tile_variant == 2, st->amp, a, st->stream),
                  "tile kernel");
      }
    } else {
      launch_standalone(st, st->amp, ops[s.standalone], P, si);
    }
    if (timer) hip_check(hipEventRecord(timer->ev[2 * si + 1], st->stream), "hipEventRecord");
  }
  // ---- a measuring application records its times and advances the selection
  if (timer) {
    hip_check(hipEventSynchronize(timer->ev.back()), "hipEventSynchronize");
    std::vector<float> ms(stages.size(), 0.0f);
    double total = 0.0;
    for (size_t si = 0; si < stages.size(); ++si) {
      hip_check(hipEventElapsedTime(&ms[si], timer->ev[2 * si], timer->ev[2 * si + 1]), "hipEventElapsedTime");
      total += ms[si];
    }
    const int C = (int)ts->cands.size();
    if (step < C) {
      ts->cand_ms[step] = total;
      ts->stage_ms[step] = ms;
      if (step == C - 1)
        ts->chosen = (int)(std::min_element(ts->cand_ms.begin(), ts->cand_ms.end()) - ts->cand_ms.begin());
    } else {
      const std::vector<float>& base = ts->stage_ms[ts->chosen];
      ts->order.assign(stages.size(), default_order);
      for (size_t si = 0; si < stages.size(); ++si)
        if (stages[si].tile && ms[si] < base[si]) ts->order[si] = measure_order;
    }
    ts->step++;
    ps.tuning = 1;
  }
  ps.gates = 0;
  for (const Stage& s : stages) {
    ps.gates += s.ngates;
    if (s.tile) ps.max_stage_gates = std::max<uint64_t>(ps.max_stage_gates, (uint64_t)s.ngates);
    if (s.tile && s.ngates == 0) ps.restore_sweeps++;
  }
  if (stats) *stats = ps;
}

// Reference-order execution: one sweep per gate (Circuit::apply_gate arithmetic).
void run_exact(qr_state* st, const std::vector<ExactGate>& gates, qr_plan_stats* stats) {
  std::vector<uint8_t> blob;
  std::vector<size_t> mo(gates.size()), oo(gates.size());
  for (size_t g = 0; g < gates.size(); ++g) {
    while (blob.size() % 16) blob.push_back(0);
    mo[g] = blob.size();
    const ExactGate& e = gates[g];
    std::vector<double2> M(e.M.size());
    for (size_t i = 0; i < M.size(); ++i) M[i] = make_double2(e.M[i].real(), e.M[i].imag());
    blob.resize(blob.size() + M.size() * sizeof(double2));
    std::memcpy(blob.data() + mo[g], M.data(), M.size() * sizeof(double2));
    if (e.K > 4) {  // wide: offsets, order and flags tables
      const uint64_t dim = 1ull << e.K;
      while (blob.size() % 16) blob.push_back(0);
      oo[g] = blob.size();
      std::vector<uint64_t> off(dim);
      for (uint64_t j = 0; j < dim; ++j) {
        off[j] = 0;
        for (int q = 0; q < e.K; ++q)
          if ((j >> (e.K - 1 - q)) & 1) off[j] |= 1ull << e.bits[q];
      }
      std::vector<uint32_t> order(dim);
      std::iota(order.begin(), order.end(), 0);
      std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return off[a] < off[b]; });
      std::vector<uint8_t> flags(dim);
      for (uint64_t j = 0; j < dim; ++j) flags[j] = (e.some[j] ? 1 : 0) | (e.keep[j] ? 2 : 0);
      size_t at = blob.size();
      blob.resize(at + dim * 8 + dim * 4 + dim);
      std::memcpy(blob.data() + at, off.data(), dim * 8);
      std::memcpy(blob.data() + at + dim * 8, order.data(), dim * 4);
      std::memcpy(blob.data() + at + dim * 12, flags.data(), dim);
    }
  }
  uint8_t* dblob = upload(st, st->aux, blob);
  qr_plan_stats ps{};
  for (size_t g = 0; g < gates.size(); ++g) {
    const ExactGate& e = gates[g];
    const uint64_t dim = 1ull << e.K;
    std::vector<int> sb(e.bits);
    std::sort(sb.begin(), sb.end());
    ps.sweeps++;
    const double bytes = 2.0 * (double)st->dim * st->bytes_per_amp();
    ps.bytes += bytes;
    if (e.K <= 4) {
      ExactArgs a{};
      a.K = e.K;
      for (int q = 0; q < e.K; ++q) a.sorted_bits[q] = (uint8_t)sb[q];
      for (uint64_t j = 0; j < dim; ++j) {
        a.off[j] = 0;
        for (int q = 0; q < e.K; ++q)
          if ((j >> (e.K - 1 - q)) & 1) a.off[j] |= 1ull << e.bits[q];
      }
      std::vector<int> order(dim);
      std::iota(order.begin(), order.end(), 0);
      std::sort(order.begin(), order.end(), [&](int x, int y) { return a.off[x] < a.off[y]; });
      for (uint64_t j = 0; j < dim; ++j) a.order[j] = (uint8_t)order[j];
      a.some = 0;
      a.keep = 0;
      for (uint64_t j = 0; j < dim; ++j) {
        if (e.some[j]) a.some |= 1u << j;
        if (e.keep[j]) a.keep |= 1u << j;
      }
      ProfScope ps_(st, "exact_gate", bytes);
      hip_check(launch_exact(st->prec, st->amp, st->dim >> e.K, a, (const double2*)(dblob + mo[g]), st->stream),
                "exact kernel");
    } else {
      ExactWideArgs a{};
      a.K = e.K;
      for (int q = 0; q < e.K; ++q) a.sorted_bits[q] = (uint8_t)sb[q];
      const uint64_t ngroups = st->dim >> e.K;
      uint64_t tg = std::min<uint64_t>(ngroups, std::max<uint64_t>(256, (64ull << 20) / (dim * 16)));
      tg = std::max<uint64_t>(tg / 256, 1) * 256;
      if (ngroups < 256) tg = ngroups;
      st->tmp.reserve(tg * dim * st->bytes_per_amp());
      const uint8_t* base = dblob + oo[g];
      ProfScope ps_(st, "exact_wide", bytes);
      hip_check(launch_exact_wide(st->prec, st->amp, st->tmp.p, tg, ngroups, a, (const double2*)(dblob + mo[g]),
                                  (const uint64_t*)base, (const uint32_t*)(base + dim * 8), base + dim * 12,
                                  st->stream),
                "exact wide kernel");
    }
  }
  ps.gates = gates.size();
  if (stats) *stats = ps;
}

qr_state* new_state(uint32_t n, qr_precision p) {
  if (n == 0) throw Error{QR_ERR_ARG, "The initialised circuit must have at least one wire."};
  if (p != QR_C128 && p != QR_C64) throw Error{QR_ERR_ARG, "unknown precision"};
  if (n > 40) throw Error{QR_ERR_UNSUPPORTED, "at most 40 qubits"};
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error{QR_ERR_NO_DEVICE, "no GPU visible"};
  auto st = std::make_unique<qr_state>();
  st->n = n;
  st->prec = p == QR_C128 ? 0 : 1;
  st->dim = 1ull << n;
  hip_check(hipGetDevice(&st->device), "hipGetDevice");
  hip_check(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking), "hipStreamCreate");
  const size_t bytes = st->dim * st->bytes_per_amp();
  hipError_t e = hipMalloc(&st->amp, bytes);
  if (e != hipSuccess) {
    st->amp = nullptr;
    (void)hipGetLastError();
    throw Error{QR_ERR_OOM, "cannot allocate " + std::to_string(bytes) + " bytes for a " + std::to_string(n) +
                                "-qubit state: " + hipGetErrorString(e)};
  }
  hip_check(launch_init_basis(st->prec, st->amp, st->dim, 0, st->stream), "init kernel");
  return st.release();
}

void do_apply(qr_state* st, const qr_program* prog) {
  validate_program(st->n, prog);
  std::vector<ExactGate> gates = lower_exact(st->n, prog);
  if (!st->opt.fused) {
    run_exact(st, gates, &st->last);
    return;
  }
  std::vector<FastOp> ops = lower_fast(st->n, gates, false);
  run_fast(st, ops, true, &st->last);
}

double sum_partials_host(const double* v, size_t count) {
  double s = 0.0, c = 0.0;
  for (size_t i = 0; i < count; ++i) {  // Neumaier
    const double x = v[i];
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  return s + c;
}

double sum_partials_host(const std::vector<double>& v) { return sum_partials_host(v.data(), v.size()); }

double norm_sqr(qr_state* st) {
  const int nb = 1024;
  st->aux2.reserve(nb * sizeof(double));
  hip_check(launch_norm_partial(st->prec, st->amp, st->dim, (double*)st->aux2.p, nb, st->stream), "norm");
  std::vector<double> h(nb);
  hip_check(hipMemcpyAsync(h.data(), st->aux2.p, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
  return sum_partials_host(h);
}

// <psi|O|psi> in one pass over the state (terms grouped by X mask), partials summed on the host.
double expectation(qr_state* st, const std::vector<PauliMasks>& terms) {
  if (terms.empty()) return 0.0;
  std::vector<PauliGroup> groups;
  std::vector<PauliZTerm> flat;
  pauli_tables(terms, groups, flat);
  const int G = (int)groups.size();
  std::vector<PauliPass> passes(G + 1);
  std::vector<uint32_t> lxm(G);
  const int np = pauli_passes(st->n, st->prec, groups.data(), G, passes.data(), lxm.data());
  // device tables: groups, their tile masks (padded to 8 bytes), terms
  const size_t gb = (size_t)G * sizeof(PauliGroup), lb = ((size_t)G * 4 + 7) & ~(size_t)7;
  const size_t tb = flat.size() * sizeof(PauliZTerm);
  std::vector<uint8_t> blob(gb + lb + tb, 0);
  std::memcpy(blob.data(), groups.data(), gb);
  std::memcpy(blob.data() + gb, lxm.data(), (size_t)G * 4);
  if (tb) std::memcpy(blob.data() + gb + lb, flat.data(), tb);
  uint8_t* d = upload(st, st->aux, blob);
  const int nb = pauli_expect_partials(st->dim);
  st->aux2.reserve(nb * sizeof(double));
  hip_check(launch_pauli_expect_passes(st->prec, st->amp, st->dim, (const PauliGroup*)d, (const uint32_t*)(d + gb),
                                       (const PauliZTerm*)(d + gb + lb), passes.data(), np, (double*)st->aux2.p,
                                       st->stream),
            "expectation");
  std::vector<double> h(nb);
  hip_check(hipMemcpyAsync(h.data(), st->aux2.p, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
  return sum_partials_host(h);
}


void apply_observable(int prec, void* lam, const void* psi, uint64_t dim, const std::vector<PauliMasks>& terms,
                      DevBuf& scratch, hipStream_t s) {
  std::vector<PauliGroup> groups;
  std::vector<PauliZTerm> flat;
  pauli_tables(terms, groups, flat);
  const size_t gb = groups.size() * sizeof(PauliGroup), tb = flat.size() * sizeof(PauliZTerm);
  scratch.reserve(gb + tb + 16);
  std::vector<uint8_t> blob(gb + tb);
  if (gb) std::memcpy(blob.data(), groups.data(), gb);
  if (tb) std::memcpy(blob.data() + gb, flat.data(), tb);
  if (!blob.empty())
    hip_check(hipMemcpyAsync(scratch.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s), "H2D");
  hip_check(launch_pauli_sum(prec, lam, psi, dim, (const PauliGroup*)scratch.p, (int)groups.size(),
                             (const PauliZTerm*)((uint8_t*)scratch.p + gb), s),
            "pauli sum");
  // the tables must outlive the launch that reads them
  hip_check(hipStreamSynchronize(s), "sync");
}

// Tile shape of the adjoint sweep: psi and lam share a workgroup, so tiles are half the
// width of the forward sweep's (2^10 complex128 / 2^11 complex64 amplitudes by default).
PlanParams dual_plan_params(const qr_state* st) {
  PlanParams pp;
  int k = st->prec == 0 ? 10 : 11;
  if (st->opt.tile_bits) k = std::min<int>(tile_max_bits(st->prec), (int)st->opt.tile_bits);
  pp.tile_bits = std::min<int>(k, (int)st->n);
  pp.reg_bits = std::min(4, pp.tile_bits);
  pp.low_bits = st->opt.low_bits ? (int)st->opt.low_bits : (st->prec == 0 ? 3 : 4);
  pp.max_ops = stage_op_limit(st, pp.tile_bits);
  return pp;
}

// The reverse sweep as one fused program over (psi, lam): for every op from last to first,
// a gradient tap (parameter ops) then the op's adjoint on both states. Tile stages carry
// both states and write per-wavefront tap partials; standalone ops run on each state.
void backward_fused(qr_grad* g, double grad_out, double* grads_out) {
  qr_state* ps = g->psi.get();
  qr_state* ls = g->lam.get();
  hipStream_t s = ps->stream;
  const std::vector<FastOp> rev = adjoint_sweep_ops(g->ops);
  const PlanParams pp = dual_plan_params(ps);
  std::shared_ptr<CachedPlan> cp = get_plan(ps, rev, pp, true, true, false);
  prepare_compiled(*cp, ps->prec, pp.reg_bits, false, true);
  PackedPlan P;
  P.stages = cp->stages;
  pack_plan(ps, rev, P);
  const uint64_t nslots = jit_dual_slots((int)ps->n, pp.tile_bits, pp.reg_bits);
  const uint64_t np = g->nparams;
  ls->aux2.reserve((std::max<uint64_t>(np, 1) * nslots + np + 1) * sizeof(double));
  double* part = (double*)ls->aux2.p;
  double* sums = part + std::max<uint64_t>(np, 1) * nslots;
  for (size_t si = 0; si < P.stages.size(); ++si) {
    const Stage& st = P.stages[si];
    if (st.tile) {
      const TileArgs a = tile_args(ps, P, si);
      ProfScope ps_(ps, "adjoint_sweep", 2.0 * st.bytes);
      hip_check(jit_launch_dual(cp->fns[si], st.k, pp.reg_bits, ps->amp, ls->amp, a.coefs, a.ntiles, a.nontile_mask, part,
                                nslots, s),
                "compiled adjoint stage");
    } else {
      const FastOp& op = rev[st.standalone];
      if (op.kind == FK_GRAD) throw Error{QR_ERR_UNSUPPORTED, "internal error: gradient tap outside a tile"};
      launch_standalone(ps, ps->amp, op, P, si);
      launch_standalone(ps, ls->amp, op, P, si);
    }
  }
  hip_check(launch_sum_rows(part, nslots, (uint32_t)np, sums, s), "gradient sums");
  std::vector<double> h(np);
  if (np) hip_check(hipMemcpyAsync(h.data(), sums, np * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
  hip_check(hipStreamSynchronize(s), "sync");
  g->have_forward = false;
  for (const auto& op : g->ops) {
    if (op.param < 0) continue;
    // dE/dtheta = Im <lam|G|psi> for exp(-i theta G/2); -Im <lam|psi> for exp(i theta/2) I
    const double d = (op.gen == 4) ? -h[op.param] : h[op.param];
    grads_out[op.param] = grad_out * d;
  }
}

}  // namespace qr
