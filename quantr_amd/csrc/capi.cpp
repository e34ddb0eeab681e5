// capi.cpp — the C ABI of include/quantr_hip.h.
//
// Orchestration only: lowering (lower.cpp), sweep planning (planner.cpp) and the kernels
// (kernels.hip). Every entry point catches errors and reports them as a qr_status plus a
// thread-local message; nothing unwinds across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"

namespace qr {
const char* last_error_cstr();
}

using namespace qr;

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    throw Error{QR_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e)};
  throw Error{QR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class F>
qr_status guard(F&& f) {
  try {
    f();
    clear_error();
    return QR_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return QR_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return QR_ERR_HIP;
  } catch (...) {
    set_error("unknown error");
    return QR_ERR_HIP;
  }
}

// Device buffer that grows on demand.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  void reserve(size_t bytes) {
    if (bytes <= cap) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 1 << 16);
    hip_check(hipMalloc(&p, want), "hipMalloc (workspace)");
    cap = want;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};

}  // namespace

struct qr_state {
  uint32_t n = 0;
  int prec = 0;  // 0 = c128, 1 = c64
  uint64_t dim = 0;
  void* amp = nullptr;
  hipStream_t stream = nullptr;
  int device = 0;
  qr_options opt{1, 0, 0, 2, 0, 0, 2};
  DevBuf aux;   // plan / coefficients / small uploads
  DevBuf aux2;  // partial sums, sampler scratch
  DevBuf tmp;   // wide-gate staging (bounded, never state-sized)
  qr_plan_stats last{};
  bool profiling = false;
  std::vector<ProfRec> prof;
  size_t bytes_per_amp() const { return prec == 0 ? 16 : 8; }
  ~qr_state() {
    for (auto& r : prof) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (amp) (void)hipFree(amp);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct qr_grad {
  std::unique_ptr<qr_state> psi;
  std::unique_ptr<qr_state> lam;
  std::vector<FastOp> ops;  // forward program (fast form, with parameters)
  struct Term {
    double coeff;
    uint64_t xm, zm;
    int ny;
  };
  std::vector<Term> terms;
  uint64_t nparams = 0;
  bool have_forward = false;
};

namespace {

// ---- profiling helpers -----------------------------------------------------------------
struct ProfScope {
  qr_state* st;
  const char* name;
  double bytes;
  hipEvent_t a{}, b{};
  ProfScope(qr_state* s, const char* nm, double by) : st(s), name(nm), bytes(by) {
    if (!st->profiling) return;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    hip_check(hipEventRecord(a, st->stream), "hipEventRecord");
  }
  ~ProfScope() {
    if (!st->profiling) return;
    (void)hipEventRecord(b, st->stream);
    st->prof.push_back({name, a, b, bytes});
  }
};

template <class T>
T* upload(qr_state* st, DevBuf& buf, const std::vector<T>& v) {
  buf.reserve(v.size() * sizeof(T) + 256);
  if (!v.empty())
    hip_check(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st->stream),
              "hipMemcpyAsync H2D");
  hip_check(hipStreamSynchronize(st->stream), "hipStreamSynchronize");
  return (T*)buf.p;
}

// Default tile shape: 2^12 complex128 / 2^13 complex64 amplitudes (64 KB of LDS, two
// workgroups per CU), 16 amplitudes per thread.
int pick_tile_bits(const qr_state* st) {
  int k = st->prec == 0 ? 12 : 13;
  if (st->opt.tile_bits) k = std::min<int>(tile_max_bits(st->prec), (int)st->opt.tile_bits);
  return std::min<int>(k, (int)st->n);
}

PlanParams plan_params(const qr_state* st) {
  PlanParams pp;
  pp.tile_bits = pick_tile_bits(st);
  pp.reg_bits = tile_reg_bits(pp.tile_bits);
  if ((st->opt.reg_bits == 5 || st->opt.reg_bits == 3) && pp.tile_bits >= 10) pp.reg_bits = (int)st->opt.reg_bits;
  pp.low_bits = st->opt.low_bits ? (int)st->opt.low_bits : (st->prec == 0 ? 3 : 4);
  return pp;
}

// Device copy of a plan: every tile stage's ops, coefficients, slot maps and LDS patterns,
// and the standalone ops' tables, in one upload.
struct Loc {
  size_t ops = 0, maps = 0, coefs = 0, off = 0;
};
struct PackedPlan {
  std::vector<Stage> stages;
  std::vector<Loc> loc;
  uint8_t* dblob = nullptr;
};

void pack_plan(qr_state* st, const std::vector<FastOp>& ops, PackedPlan& P) {
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
  std::vector<Stage>& stages = P.stages;
  P.loc.assign(stages.size(), Loc{});
  for (size_t si = 0; si < stages.size(); ++si) {
    Stage& s = stages[si];
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
        if (h.src >= 0) {  // the current program's coefficients (plans are shared by structure)
          const std::vector<cd>& m = ops[h.src].m;
          for (int q = 0; q < 4; ++q)
            c[4 * i + q] = q < (int)m.size() ? make_double2(m[q].real(), m[q].imag()) : make_double2(0.0, 0.0);
        } else {
          for (int q = 0; q < 4; ++q) c[4 * i + q] = make_double2(h.coef[q].real(), h.coef[q].imag());
        }
      }
      P.loc[si].ops = append(d.data(), d.size() * sizeof(TileOpDev));
      align(16);
      P.loc[si].coefs = append(c.data(), c.size() * sizeof(double2));
      align(4);
      std::vector<uint32_t> m(s.maps.size());
      for (size_t i = 0; i < m.size(); ++i)
        m[i] = (uint32_t)s.tbits[s.maps[i]] | ((uint32_t)s.maps[i] << 8);
      P.loc[si].maps = append(m.data(), m.size() * sizeof(uint32_t));
      std::vector<uint32_t> z(s.swz.begin(), s.swz.end());
      if (z.empty()) z.push_back(0);
      P.loc[si].off = append(z.data(), z.size() * sizeof(uint32_t));
    } else {
      const FastOp& op = ops[s.standalone];
      if (op.kind == FK_DENSEK || op.kind == FK_DIAGK) {
        align(16);
        std::vector<double2> c(op.m.size());
        for (size_t i = 0; i < c.size(); ++i) c[i] = make_double2(op.m[i].real(), op.m[i].imag());
        P.loc[si].coefs = append(c.data(), c.size() * sizeof(double2));
        align(8);
        const int dim = 1 << op.k;
        std::vector<uint64_t> off(dim, 0);
        for (int j = 0; j < dim; ++j)
          for (int q = 0; q < op.k; ++q)
            if ((j >> (op.k - 1 - q)) & 1) off[j] |= 1ull << op.bits[q];
        P.loc[si].off = append(off.data(), off.size() * sizeof(uint64_t));
      }
    }
  }
  P.dblob = upload(st, st->aux, blob);
}

TileArgs tile_args(const qr_state* st, const PackedPlan& P, size_t si) {
  const Stage& s = P.stages[si];
  TileArgs a{};
  a.ops = (const TileOpDev*)(P.dblob + P.loc[si].ops);
  a.coefs = (const double2*)(P.dblob + P.loc[si].coefs);
  a.maps = (const uint32_t*)(P.dblob + P.loc[si].maps);
  a.swz = (const uint32_t*)(P.dblob + P.loc[si].off);
  a.nops = (uint32_t)s.ops.size();
  a.ntiles = 1ull << (st->n - s.k);
  uint64_t tm = 0;
  for (int b : s.tbits) tm |= 1ull << b;
  a.nontile_mask = (st->n >= 64 ? ~0ull : ((1ull << st->n) - 1)) & ~tm;
  return a;
}

// One standalone (unfused) op on `amp`, a state of st's size, on st's stream.
void launch_standalone(qr_state* st, void* amp, const FastOp& op, const PackedPlan& P, size_t si) {
  const Stage& s = P.stages[si];
  switch (op.kind) {
    case FK_DENSE1: {
      double2 u[4];
      for (int i = 0; i < 4; ++i) u[i] = make_double2(op.m[i].real(), op.m[i].imag());
      ProfScope ps_(st, "dense1", s.bytes);
      hip_check(launch_dense1(st->prec, amp, st->dim / 2, op.bits[0], op.cmask, u, st->stream), "dense1");
      break;
    }
    case FK_DIAG1: {
      ProfScope ps_(st, "diag1", s.bytes);
      hip_check(launch_diag1(st->prec, amp, st->dim, op.bits[0], op.cmask, make_double2(op.m[0].real(), op.m[0].imag()),
                             make_double2(op.m[1].real(), op.m[1].imag()), st->stream),
                "diag1");
      break;
    }
    case FK_DENSEK: {
      DenseKArgs a{};
      a.k = op.k;
      a.cmask = op.cmask;
      std::vector<int> sb(op.bits, op.bits + op.k);
      std::sort(sb.begin(), sb.end());
      for (int q = 0; q < op.k; ++q) a.sorted_bits[q] = (uint8_t)sb[q];
      const uint64_t dimk = 1ull << op.k;
      const uint64_t ngroups = st->dim >> op.k;
      uint64_t tg = std::min<uint64_t>(ngroups, std::max<uint64_t>(256, (64ull << 20) / (dimk * 16)));
      tg = std::max<uint64_t>(tg / 256, 1) * 256;
      if (ngroups < 256) tg = ngroups;
      st->tmp.reserve(tg * dimk * st->bytes_per_amp());
      ProfScope ps_(st, "densek", s.bytes);
      hip_check(launch_densek(st->prec, amp, st->tmp.p, tg, ngroups, a, (const double2*)(P.dblob + P.loc[si].coefs),
                              (const uint64_t*)(P.dblob + P.loc[si].off), st->stream),
                "densek");
      break;
    }
    case FK_DIAGK: {
      DiagKArgs a{};
      a.k = op.k;
      a.cmask = op.cmask;
      for (int q = 0; q < op.k; ++q) a.bits[q] = (uint8_t)op.bits[q];
      ProfScope ps_(st, "diagk", s.bytes);
      hip_check(launch_diagk(st->prec, amp, st->dim, a, (const double2*)(P.dblob + P.loc[si].coefs), st->stream),
                "diagk");
      break;
    }
    default:
      throw Error{QR_ERR_UNSUPPORTED, "internal error: op kind has no standalone kernel"};
  }
}

// ---- plan cache ---------------------------------------------------------------------------
// Plans depend on the program's structure only (plan_key_append); parameter sweeps, gradient
// loops and serving re-apply one structure with new coefficients, so plans and their
// compiled stages are kept per process (least recently used dropped beyond 64).
struct CachedPlan {
  std::vector<Stage> stages;
  std::vector<std::string> src;    // compiled-stage sources (filled on first compiled use)
  std::vector<hipFunction_t> fns;  // their kernels
  uint64_t runs = 0;
  uint64_t stamp = 0;
};
std::mutex g_plan_mu;
std::map<std::vector<uint64_t>, std::shared_ptr<CachedPlan>> g_plans;
uint64_t g_plan_clock = 0;

std::shared_ptr<CachedPlan> get_plan(const qr_state* st, const std::vector<FastOp>& ops, const PlanParams& pp,
                                     bool fused, bool dual, bool persistent) {
  std::vector<uint64_t> key{st->n, (uint64_t)st->prec, (uint64_t)pp.tile_bits, (uint64_t)pp.reg_bits,
                            (uint64_t)pp.low_bits, (uint64_t)fused, (uint64_t)dual, (uint64_t)persistent};
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
  const int B = (int)st->bytes_per_amp();
  if (fused) {
    cp->stages = plan_fused(st->n, ops, pp, B, dual);
  } else {
    for (size_t i = 0; i < ops.size(); ++i) {
      Stage s;
      s.tile = false;
      s.standalone = (int)i;
      s.bytes = 2.0 * (double)st->dim * B * ops[i].f_touch;
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

// Compiles (once) the tile stages of a cached plan and resolves their kernels.
void prepare_compiled(CachedPlan& cp, int prec, int RB, bool persistent, bool dual) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  if (!cp.fns.empty() || cp.stages.empty()) return;
  cp.src = jit_stage_sources(cp.stages, prec, RB, persistent, dual);
  jit_prepare(cp.src);
  std::vector<hipFunction_t> f(cp.stages.size(), nullptr);
  for (size_t i = 0; i < cp.stages.size(); ++i)
    if (cp.stages[i].tile) f[i] = jit_function(cp.src[i]);
  cp.fns = std::move(f);
}

// Plans (fused tile sweeps, or one sweep per op), uploads and launches a fast-form program.
void run_fast(qr_state* st, const std::vector<FastOp>& ops, bool fused, qr_plan_stats* stats) {
  const PlanParams pp = plan_params(st);
  const bool persistent = st->opt.tile_variant == 2;
  std::shared_ptr<CachedPlan> cp = get_plan(st, ops, pp, fused, false, persistent);
  PackedPlan P;
  P.stages = cp->stages;
  std::vector<Stage>& stages = P.stages;
  pack_plan(st, ops, P);
  // ---- runtime-specialised stages: always with jit = 1, from 28 qubits or from the second
  // application of a structure with jit = 2
  bool use_jit = false;
  if (fused && st->opt.jit != 0) {
    use_jit = st->opt.jit == 1 || st->n >= 28 || cp->runs > 0;
    if (use_jit) prepare_compiled(*cp, st->prec, pp.reg_bits, persistent, false);
  }
  cp->runs++;
  // ---- launch --------------------------------------------------------------------------
  qr_plan_stats ps{};
  for (size_t si = 0; si < stages.size(); ++si) {
    const Stage& s = stages[si];
    ps.sweeps++;
    ps.bytes += s.bytes;
    if (s.tile) {
      ps.tile_sweeps++;
      const TileArgs a = tile_args(st, P, si);
      ProfScope ps_(st, "tile_sweep", s.bytes);
      if (use_jit)
        hip_check(jit_launch(cp->fns[si], s.k, pp.reg_bits, persistent, st->amp, a.coefs, a.ntiles, a.nontile_mask,
                             st->stream),
                  "compiled tile kernel");
      else
        hip_check(launch_tile(st->prec, s.k, pp.reg_bits, st->opt.tile_variant == 2, st->amp, a, st->stream),
                  "tile kernel");
      continue;
    }
    launch_standalone(st, st->amp, ops[s.standalone], P, si);
  }
  ps.gates = 0;
  for (const Stage& s : stages) ps.gates += s.ngates;
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

double sum_partials_host(const std::vector<double>& v) {
  double s = 0.0, c = 0.0;
  for (double x : v) {  // Neumaier
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  return s + c;
}

double norm_sqr(qr_state* st) {
  const int nb = 1024;
  st->aux2.reserve(nb * sizeof(double));
  hip_check(launch_norm_partial(st->prec, st->amp, st->dim, (double*)st->aux2.p, nb, st->stream), "norm");
  std::vector<double> h(nb);
  hip_check(hipMemcpyAsync(h.data(), st->aux2.p, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
  return sum_partials_host(h);
}

std::vector<qr_grad::Term> parse_observable(uint32_t n, const qr_observable* o) {
  if (!o) throw Error{QR_ERR_ARG, "observable is NULL"};
  std::vector<qr_grad::Term> out;
  for (uint64_t t = 0; t < o->nterms; ++t) {
    const qr_pauli_term& pt = o->terms[t];
    if (pt.nfactors && (!o->wires || !o->paulis || pt.offset + pt.nfactors > o->nfactors))
      throw Error{QR_ERR_ARG, "observable term " + std::to_string(t) + " is out of range"};
    qr_grad::Term tm{pt.coeff, 0, 0, 0};
    uint64_t seen = 0;
    for (uint32_t f = 0; f < pt.nfactors; ++f) {
      const uint32_t w = o->wires[pt.offset + f];
      if (w >= n) throw Error{QR_ERR_ARG, "observable wire " + std::to_string(w) + " is out of range"};
      const uint64_t b = 1ull << (n - 1 - w);
      if (seen & b) throw Error{QR_ERR_ARG, "observable term repeats wire " + std::to_string(w)};
      seen |= b;
      switch (o->paulis[pt.offset + f]) {
        case 'X': case 1: tm.xm |= b; break;
        case 'Y': case 2: tm.xm |= b; tm.zm |= b; tm.ny++; break;
        case 'Z': case 3: tm.zm |= b; break;
        case 'I': case 0: break;
        default: throw Error{QR_ERR_ARG, "observable pauli must be X, Y or Z"};
      }
    }
    out.push_back(tm);
  }
  return out;
}

// <psi|O|psi>
double expectation(qr_state* st, const std::vector<qr_grad::Term>& terms) {
  const int nb = 512;
  std::vector<uint64_t> zm;
  std::vector<double> zc;
  std::vector<size_t> offdiag;
  for (size_t t = 0; t < terms.size(); ++t) {
    if (terms[t].xm == 0) {
      zm.push_back(terms[t].zm);
      zc.push_back(terms[t].coeff * ((terms[t].ny & 1) ? 0.0 : 1.0));
    } else {
      offdiag.push_back(t);
    }
  }
  const size_t nparts = (zm.empty() ? 0 : nb) + offdiag.size() * nb * 2;
  std::vector<uint8_t> blob(zm.size() * 8 + zc.size() * 8);
  if (!zm.empty()) {
    std::memcpy(blob.data(), zm.data(), zm.size() * 8);
    std::memcpy(blob.data() + zm.size() * 8, zc.data(), zc.size() * 8);
  }
  uint8_t* d = upload(st, st->aux, blob);
  st->aux2.reserve(std::max<size_t>(nparts, 1) * sizeof(double));
  double* parts = (double*)st->aux2.p;
  size_t at = 0;
  if (!zm.empty()) {
    hip_check(launch_zdiag_partial(st->prec, st->amp, st->dim, (const uint64_t*)d, (const double*)(d + zm.size() * 8),
                                   (int)zm.size(), parts, nb, st->stream),
              "zdiag");
    at += nb;
  }
  for (size_t i = 0; i < offdiag.size(); ++i) {
    const auto& tm = terms[offdiag[i]];
    hip_check(launch_pauli_partial(st->prec, st->amp, st->amp, st->dim, tm.xm, tm.zm, (double2*)(parts + at), nb,
                                   st->stream),
              "pauli");
    at += 2 * nb;
  }
  std::vector<double> h(nparts);
  if (nparts)
    hip_check(hipMemcpyAsync(h.data(), parts, nparts * sizeof(double), hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
  double total = 0.0;
  size_t rd = 0;
  if (!zm.empty()) {
    total += sum_partials_host(std::vector<double>(h.begin(), h.begin() + nb));
    rd = nb;
  }
  for (size_t i = 0; i < offdiag.size(); ++i) {
    const auto& tm = terms[offdiag[i]];
    std::vector<double> re(nb), im(nb);
    for (int b = 0; b < nb; ++b) {
      re[b] = h[rd + 2 * b];
      im[b] = h[rd + 2 * b + 1];
    }
    rd += 2 * nb;
    double sr = sum_partials_host(re), si = sum_partials_host(im);
    // multiply by i^ny, keep the real part
    for (int q = 0; q < (tm.ny & 3); ++q) {
      const double t = sr;
      sr = -si;
      si = t;
    }
    total += tm.coeff * sr;
  }
  return total;
}

}  // namespace

// ===========================================================================================
extern "C" {

uint32_t qr_abi_version(void) { return QR_ABI_VERSION; }
const char* qr_version(void) { return "quantr-hip 0.1.0 (gfx950)"; }
const char* qr_last_error(void) { return last_error_cstr(); }

qr_status qr_device_count(int* count) {
  return guard([&] {
    if (!count) throw Error{QR_ERR_ARG, "count is NULL"};
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
  });
}

qr_status qr_state_create(uint32_t num_qubits, qr_precision precision, qr_state** out) {
  return guard([&] {
    if (!out) throw Error{QR_ERR_ARG, "out is NULL"};
    *out = nullptr;
    *out = new_state(num_qubits, precision);
  });
}

qr_status qr_state_destroy(qr_state* state) {
  return guard([&] {
    if (!state) return;
    (void)hipStreamSynchronize(state->stream);
    delete state;
  });
}

qr_status qr_state_num_qubits(const qr_state* st, uint32_t* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = st->n;
  });
}

qr_status qr_state_precision(const qr_state* st, qr_precision* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = st->prec == 0 ? QR_C128 : QR_C64;
  });
}

qr_status qr_state_reset(qr_state* st) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    hip_check(launch_init_basis(st->prec, st->amp, st->dim, 0, st->stream), "init kernel");
  });
}

qr_status qr_state_set_amplitudes(qr_state* st, const double* re_im, uint64_t len) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    if (len != st->dim) {
      // Circuit::change_register (circuit.rs:463-468) compares product dimensions.
      uint32_t pd = 0;
      while (pd < 64 && (1ull << pd) < len) ++pd;
      const bool pow2 = len && ((len & (len - 1)) == 0);
      throw Error{QR_ERR_ARG, "The custom register has a product state dimension of " +
                                  (pow2 ? std::to_string(pd) : std::string("(not a power of two: ") + std::to_string(len) + " amplitudes)") +
                                  ", while the number of qubits is " + std::to_string(st->n) +
                                  ". These must equal each other."};
    }
    if (!re_im) throw Error{QR_ERR_ARG, "amplitudes pointer is NULL"};
    hip_check(hipStreamSynchronize(st->stream), "sync");
    if (st->prec == 0) {
      hip_check(hipMemcpy(st->amp, re_im, len * 16, hipMemcpyHostToDevice), "hipMemcpy H2D");
    } else {
      const uint64_t chunk = 1ull << 24;
      st->tmp.reserve(chunk * 16);
      for (uint64_t o = 0; o < len; o += chunk) {
        const uint64_t c = std::min(chunk, len - o);
        hip_check(hipMemcpy(st->tmp.p, re_im + 2 * o, c * 16, hipMemcpyHostToDevice), "hipMemcpy H2D");
        hip_check(launch_f64_to_c64((float2*)st->amp + o, (const double2*)st->tmp.p, c, st->stream), "convert");
        hip_check(hipStreamSynchronize(st->stream), "sync");
      }
    }
  });
}

qr_status qr_state_get_amplitudes_range(qr_state* st, uint64_t first, uint64_t count, double* out) {
  return guard([&] {
    if (!st || (!out && count)) throw Error{QR_ERR_ARG, "NULL argument"};
    if (first > st->dim || count > st->dim - first) throw Error{QR_ERR_ARG, "range out of bounds"};
    hip_check(hipStreamSynchronize(st->stream), "sync");
    if (st->prec == 0) {
      hip_check(hipMemcpy(out, (double2*)st->amp + first, count * 16, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    } else {
      const uint64_t chunk = 1ull << 24;
      st->tmp.reserve(chunk * 16);
      for (uint64_t o = 0; o < count; o += chunk) {
        const uint64_t c = std::min(chunk, count - o);
        hip_check(launch_c64_to_f64((double2*)st->tmp.p, (const float2*)st->amp + first + o, c, st->stream), "convert");
        hip_check(hipStreamSynchronize(st->stream), "sync");
        hip_check(hipMemcpy(out + 2 * o, st->tmp.p, c * 16, hipMemcpyDeviceToHost), "hipMemcpy D2H");
      }
    }
  });
}

qr_status qr_state_get_amplitudes(qr_state* st, double* out, uint64_t len) {
  if (st && len != st->dim) {
    set_error("output length " + std::to_string(len) + " does not match the state's " +
              std::to_string(st ? st->dim : 0) + " amplitudes");
    return QR_ERR_ARG;
  }
  return qr_state_get_amplitudes_range(st, 0, len, out);
}

qr_status qr_state_set_options(qr_state* st, const qr_options* o) {
  return guard([&] {
    if (!st || !o) throw Error{QR_ERR_ARG, "NULL argument"};
    if (o->tile_bits > 16) throw Error{QR_ERR_ARG, "tile_bits must be <= 16"};
    if (o->low_bits > 8) throw Error{QR_ERR_ARG, "low_bits must be <= 8"};
    if (o->tile_variant > 2) throw Error{QR_ERR_ARG, "tile_variant must be 0, 1 or 2"};
    if (o->reg_bits != 0 && (o->reg_bits < 3 || o->reg_bits > 5)) throw Error{QR_ERR_ARG, "reg_bits must be 0, 3, 4 or 5"};
    if (o->jit > 2) throw Error{QR_ERR_ARG, "jit must be 0, 1 or 2"};
    st->opt = *o;
  });
}

qr_status qr_state_get_options(const qr_state* st, qr_options* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = st->opt;
  });
}

qr_status qr_state_synchronize(qr_state* st) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    hip_check(hipStreamSynchronize(st->stream), "hipStreamSynchronize");
  });
}

qr_status qr_validate_program(uint32_t num_qubits, const qr_program* program) {
  return guard([&] { validate_program(num_qubits, program); });
}

qr_status qr_apply_program(qr_state* st, const qr_program* program) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    do_apply(st, program);
  });
}

qr_status qr_state_last_plan(const qr_state* st, qr_plan_stats* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = st->last;
  });
}

qr_status qr_sample(qr_state* st, const double* uniforms, uint64_t shots, uint64_t* idx_out) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    if (shots == 0) return;
    if (!uniforms || !idx_out) throw Error{QR_ERR_ARG, "NULL argument"};
    for (uint64_t s = 0; s < shots; ++s)
      if (std::isnan(uniforms[s])) throw Error{QR_ERR_ARG, "uniform draws must not be NaN"};
    std::vector<uint64_t> perm(shots);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](uint64_t a, uint64_t b) { return uniforms[a] < uniforms[b]; });
    std::vector<double> us(shots);
    for (uint64_t s = 0; s < shots; ++s) us[s] = uniforms[perm[s]];
    std::vector<uint64_t> res(shots, ~0ull);
    const bool exact = st->opt.exact_sample == 1 || (st->opt.exact_sample == 2 && st->n <= 26);
    if (exact) {
      std::vector<uint8_t> blob(shots * 8);
      std::memcpy(blob.data(), us.data(), shots * 8);
      uint8_t* d = upload(st, st->aux, blob);
      st->aux2.reserve(shots * 8);
      ProfScope ps_(st, "sample_fold", (double)st->dim * st->bytes_per_amp());
      hip_check(launch_sample_fold(st->prec, st->amp, st->dim, (const double*)d, shots, (uint64_t*)st->aux2.p,
                                   st->stream),
                "sample fold");
    } else {
      const int bb = std::min<int>(14, (int)st->n);
      const uint64_t nblk = st->dim >> bb;
      st->aux2.reserve(nblk * 8);
      hip_check(launch_block_mass(st->prec, st->amp, st->dim, bb, (double*)st->aux2.p, st->stream), "block mass");
      std::vector<double> mass(nblk);
      hip_check(hipMemcpyAsync(mass.data(), st->aux2.p, nblk * 8, hipMemcpyDeviceToHost, st->stream), "D2H");
      hip_check(hipStreamSynchronize(st->stream), "sync");
      std::vector<double> incl(nblk);
      double cum = 0.0;
      for (uint64_t b = 0; b < nblk; ++b) {
        cum += mass[b];
        incl[b] = cum;
      }
      std::vector<BlockJob> jobs;
      uint64_t s = 0;
      while (s < shots) {
        const uint64_t b = (uint64_t)(std::upper_bound(incl.begin(), incl.end(), us[s]) - incl.begin());
        if (b >= nblk) break;  // u >= total mass: None for this and all larger draws
        uint64_t e = s;
        while (e < shots && us[e] < incl[b]) ++e;
        jobs.push_back(BlockJob{b, b ? incl[b - 1] : 0.0, s, e});
        s = e;
      }
      std::vector<uint8_t> blob(shots * 8 + jobs.size() * sizeof(BlockJob));
      std::memcpy(blob.data(), us.data(), shots * 8);
      if (!jobs.empty()) std::memcpy(blob.data() + shots * 8, jobs.data(), jobs.size() * sizeof(BlockJob));
      uint8_t* d = upload(st, st->aux, blob);
      st->aux2.reserve(shots * 8);
      hip_check(hipMemsetAsync(st->aux2.p, 0xff, shots * 8, st->stream), "memset");
      hip_check(launch_sample_blocks(st->prec, st->amp, st->dim, bb, (const BlockJob*)(d + shots * 8), jobs.size(),
                                     (const double*)d, (uint64_t*)st->aux2.p, st->stream),
                "sample blocks");
    }
    hip_check(hipMemcpyAsync(res.data(), st->aux2.p, shots * 8, hipMemcpyDeviceToHost, st->stream), "D2H");
    hip_check(hipStreamSynchronize(st->stream), "sync");
    for (uint64_t s = 0; s < shots; ++s) idx_out[perm[s]] = res[s];
  });
}

qr_status qr_state_norm_sqr(qr_state* st, double* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = norm_sqr(st);
  });
}

qr_status qr_expectation(qr_state* st, const qr_observable* observable, double* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    auto terms = parse_observable(st->n, observable);
    *out = expectation(st, terms);
  });
}

// ---- adjoint differentiation ------------------------------------------------------------

qr_status qr_grad_create(uint32_t num_qubits, qr_precision precision, qr_grad** out) {
  return guard([&] {
    if (!out) throw Error{QR_ERR_ARG, "out is NULL"};
    *out = nullptr;
    auto g = std::make_unique<qr_grad>();
    g->psi.reset(new_state(num_qubits, precision));
    g->lam.reset(new_state(num_qubits, precision));
    *out = g.release();
  });
}

qr_status qr_grad_destroy(qr_grad* g) {
  return guard([&] { delete g; });
}

qr_status qr_grad_forward(qr_grad* g, const qr_program* program, const qr_observable* observable,
                          double* expval_out) {
  return guard([&] {
    if (!g) throw Error{QR_ERR_ARG, "grad is NULL"};
    qr_state* st = g->psi.get();
    validate_program(st->n, program);
    auto gates = lower_exact(st->n, program);
    for (const auto& e : gates)
      if (e.custom) {  // the reverse sweep needs U^dagger = U^-1
        const int dim = 1 << e.K;
        for (int r = 0; r < dim; ++r)
          for (int c = 0; c < dim; ++c) {
            cd s(0, 0);
            for (int k = 0; k < dim; ++k) s += std::conj(e.M[k * dim + r]) * e.M[k * dim + c];
            if (std::abs(s - cd(r == c ? 1.0 : 0.0, 0.0)) > 1e-10)
              throw Error{QR_ERR_UNSUPPORTED, "adjoint differentiation needs unitary custom gates"};
          }
      }
    g->ops = lower_fast(st->n, gates, true);
    g->terms = parse_observable(st->n, observable);
    g->nparams = 0;
    for (const auto& op : g->ops)
      if (op.param >= 0) g->nparams++;
    hip_check(launch_init_basis(st->prec, st->amp, st->dim, 0, st->stream), "init");
    run_fast(st, g->ops, st->opt.fused != 0, &st->last);
    const double e = expectation(st, g->terms);
    if (expval_out) *expval_out = e;
    g->have_forward = true;
  });
}

qr_status qr_grad_set_options(qr_grad* g, const qr_options* o) {
  return guard([&] {
    if (!g || !o) throw Error{QR_ERR_ARG, "NULL argument"};
    qr_status st = qr_state_set_options(g->psi.get(), o);
    if (st != QR_OK) throw Error{st, qr_last_error()};
    st = qr_state_set_options(g->lam.get(), o);
    if (st != QR_OK) throw Error{st, qr_last_error()};
  });
}

qr_status qr_grad_num_params(const qr_grad* g, uint64_t* out) {
  return guard([&] {
    if (!g || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = g->nparams;
  });
}

namespace {

// Tile shape of the adjoint sweep: psi and lam share a workgroup, so tiles are half the
// width of the forward sweep's (2^10 complex128 / 2^11 complex64 amplitudes by default).
PlanParams dual_plan_params(const qr_state* st) {
  PlanParams pp;
  int k = st->prec == 0 ? 10 : 11;
  if (st->opt.tile_bits) k = std::min<int>(tile_max_bits(st->prec), (int)st->opt.tile_bits);
  pp.tile_bits = std::min<int>(k, (int)st->n);
  pp.reg_bits = std::min(4, pp.tile_bits);
  pp.low_bits = st->opt.low_bits ? (int)st->opt.low_bits : (st->prec == 0 ? 3 : 4);
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

}  // namespace

qr_status qr_grad_backward(qr_grad* g, double grad_out, double* grads_out, uint64_t nparams) {
  return guard([&] {
    if (!g) throw Error{QR_ERR_ARG, "grad is NULL"};
    if (!g->have_forward) throw Error{QR_ERR_ARG, "backward called before forward"};
    if (nparams != g->nparams)
      throw Error{QR_ERR_ARG, "expected " + std::to_string(g->nparams) + " gradient slots, got " + std::to_string(nparams)};
    if (nparams && !grads_out) throw Error{QR_ERR_ARG, "grads_out is NULL"};
    qr_state* ps = g->psi.get();
    qr_state* ls = g->lam.get();
    // The two states share psi's stream so that every step is ordered.
    hip_check(hipStreamSynchronize(ls->stream), "sync");
    hipStream_t s = ps->stream;
    // lambda = O psi
    bool first = true;
    for (const auto& tm : g->terms) {
      double cr = tm.coeff, ci = 0.0;
      for (int q = 0; q < (tm.ny & 3); ++q) {
        const double t = cr;
        cr = -ci;
        ci = t;
      }
      hip_check(launch_pauli_apply(ps->prec, ls->amp, ps->amp, ps->dim, tm.xm, tm.zm, cr, ci, first ? 0 : 1, s),
                "pauli apply");
      first = false;
    }
    if (first) hip_check(hipMemsetAsync(ls->amp, 0, ls->dim * ls->bytes_per_amp(), s), "memset");
    if (ps->opt.fused && ps->opt.jit) {
      backward_fused(g, grad_out, grads_out);
      return;
    }
    const int nb = 256;
    ls->aux2.reserve(std::max<uint64_t>(g->nparams, 1) * nb * sizeof(double2));
    double2* parts = (double2*)ls->aux2.p;
    // reverse sweep, one launch per op and state
    for (size_t i = g->ops.size(); i-- > 0;) {
      const FastOp& op = g->ops[i];
      if (op.param >= 0) {
        hip_check(launch_gen_partial(ps->prec, ls->amp, ps->amp, ps->dim, op.bits[0], op.gen,
                                     parts + (size_t)op.param * nb, nb, s),
                  "gradient term");
      }
      const FastOp adj = adjoint(op);
      for (qr_state* x : {ps, ls}) {
        switch (adj.kind) {
          case FK_DENSE1: {
            double2 u[4];
            for (int q = 0; q < 4; ++q) u[q] = make_double2(adj.m[q].real(), adj.m[q].imag());
            hip_check(launch_dense1(x->prec, x->amp, x->dim / 2, adj.bits[0], adj.cmask, u, s), "dense1");
            break;
          }
          case FK_DIAG1:
            hip_check(launch_diag1(x->prec, x->amp, x->dim, adj.bits[0], adj.cmask,
                                   make_double2(adj.m[0].real(), adj.m[0].imag()),
                                   make_double2(adj.m[1].real(), adj.m[1].imag()), s),
                      "diag1");
            break;
          default: {
            std::vector<FastOp> one{adj};
            // reuse the standalone machinery on this state's own stream
            hip_check(hipStreamSynchronize(s), "sync");
            run_fast(x, one, false, nullptr);
            hip_check(hipStreamSynchronize(x->stream), "sync");
          }
        }
      }
    }
    std::vector<double2> h((size_t)g->nparams * nb);
    if (!h.empty())
      hip_check(hipMemcpyAsync(h.data(), parts, h.size() * sizeof(double2), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
    // psi and lambda now hold U^dagger psi_P and U^dagger O psi_P; forward state is gone.
    g->have_forward = false;
    std::vector<double> re(nb), im(nb);
    for (const auto& op : g->ops) {
      if (op.param < 0) continue;
      for (int b = 0; b < nb; ++b) {
        re[b] = h[(size_t)op.param * nb + b].x;
        im[b] = h[(size_t)op.param * nb + b].y;
      }
      const double sim = sum_partials_host(im);
      // dE/dtheta = Im <lam|G|psi> for exp(-i theta G/2); -Im <lam|psi> for exp(i theta/2) I
      const double d = (op.gen == 4) ? -sim : sim;
      grads_out[op.param] = grad_out * d;
    }
  });
}

qr_status qr_batch_grad(uint64_t ninstances, const uint32_t* num_qubits, const qr_program* programs,
                        const qr_observable* observables, qr_precision precision, double* expvals_out,
                        double* grads_out, const uint64_t* grad_offsets, double* latency_us_out) {
  return guard([&] {
    if (!ninstances) return;
    if (!num_qubits || !programs || !observables || !expvals_out)
      throw Error{QR_ERR_ARG, "NULL argument"};
    hipEvent_t t0;
    hip_check(hipEventCreate(&t0), "event");
    bool started = false;
    for (uint64_t i = 0; i < ninstances; ++i) {
      qr_grad* g = nullptr;
      qr_status s = qr_grad_create(num_qubits[i], precision, &g);
      if (s != QR_OK) throw Error{s, qr_last_error()};
      std::unique_ptr<qr_grad> hold(g);
      if (!started) {
        hip_check(hipEventRecord(t0, g->psi->stream), "event");
        started = true;
      }
      double e = 0.0;
      s = qr_grad_forward(g, &programs[i], &observables[i], &e);
      if (s != QR_OK) throw Error{s, qr_last_error()};
      expvals_out[i] = e;
      if (grads_out) {
        uint64_t np = g->nparams;
        s = qr_grad_backward(g, 1.0, grads_out + (grad_offsets ? grad_offsets[i] : 0), np);
        if (s != QR_OK) throw Error{s, qr_last_error()};
      }
      if (latency_us_out) {
        hipEvent_t t1;
        hip_check(hipEventCreate(&t1), "event");
        hip_check(hipEventRecord(t1, g->psi->stream), "event");
        hip_check(hipEventSynchronize(t1), "event");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, t0, t1), "event");
        latency_us_out[i] = ms * 1000.0;
        (void)hipEventDestroy(t1);
      }
    }
    (void)hipEventDestroy(t0);
  });
}

qr_status qr_profile_enable(qr_state* st, int enable) {
  return guard([&] {
    if (!st) throw Error{QR_ERR_ARG, "state is NULL"};
    hip_check(hipStreamSynchronize(st->stream), "sync");
    for (auto& r : st->prof) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    st->prof.clear();
    st->profiling = enable != 0;
  });
}

qr_status qr_profile_read(qr_state* st, qr_profile_entry* entries, uint64_t capacity, uint64_t* count_out) {
  return guard([&] {
    if (!st || !count_out) throw Error{QR_ERR_ARG, "NULL argument"};
    hip_check(hipStreamSynchronize(st->stream), "sync");
    std::vector<qr_profile_entry> agg;
    for (auto& r : st->prof) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
      auto it = std::find_if(agg.begin(), agg.end(), [&](const qr_profile_entry& e) { return r.name == e.name; });
      if (it == agg.end()) {
        qr_profile_entry e{};
        std::strncpy(e.name, r.name.c_str(), sizeof(e.name) - 1);
        agg.push_back(e);
        it = agg.end() - 1;
      }
      it->launches++;
      it->total_ms += ms;
      it->bytes += r.bytes;
    }
    *count_out = agg.size();
    if (entries)
      for (uint64_t i = 0; i < std::min<uint64_t>(capacity, agg.size()); ++i) entries[i] = agg[i];
  });
}

}  // extern "C"
