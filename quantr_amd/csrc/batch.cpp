// batch.cpp — many parameterised programs of uneven qubit counts, each with its expectation
// value and full adjoint gradient, executed as one captured HIP graph.
//
// Creation plans every instance (forward tile stages, then the dual-state reverse sweep),
// compiles the stages of every distinct structure, lays all states and plan blobs out in
// three device allocations, and captures the whole batch: instances are spread over a few
// streams (largest first, each to the least loaded stream; each stream then runs its own
// instances smallest first), every launch recorded once. A run
// uploads the programs' coefficients into the plan blob, launches the graph and reads back
// expectation values, gradients and one device timestamp per instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace qr;

namespace {

constexpr int kDotBlocks = 64;
constexpr size_t kAlign = 256;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Inst {
  uint32_t n = 0;
  uint64_t dim = 0;
  std::vector<FastOp> fwd, rev;
  std::vector<uint64_t> key;  // structure of the program the batch was captured for
  std::vector<PauliMasks> terms;
  PlanParams ppf, ppb;
  std::shared_ptr<CachedPlan> pf, pb;
  std::vector<Loc> locf, locb;
  size_t blob_f = 0, blob_b = 0;  // offsets of the two plan blobs in the batch blob
  size_t blob_f_size = 0, blob_b_size = 0;
  size_t obs_groups = 0, obs_terms = 0;  // byte offsets of the observable's tables in the blob
  int ngroups = 0;
  uint64_t nparams = 0, nslots = 0;
  std::vector<int> gens;      // generator of each parameter (4 = global phase)
  size_t psi = 0, lam = 0;    // byte offsets in the state pool
  size_t part = 0, dot = 0;   // double offsets in the work buffer
  size_t grad = 0;            // index of the first gradient in the output
  int stream = 0;
  double work = 0.0;          // estimated bytes moved (stream balancing)
  size_t tmp_bytes = 0;       // staging needed by standalone wide gates
};

}  // namespace

struct qr_batch {
  int prec = 0;
  std::vector<Inst> inst;
  qr::DevBuf states, blob, work, out;
  std::vector<qr::DevBuf> tmp;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  hipStream_t origin = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> host_blob;
  uint64_t total_params = 0;
  double clock_khz = 0.0;
  ~qr_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : streams) (void)hipStreamDestroy(s);
    if (origin) (void)hipStreamDestroy(origin);
  }
};

namespace qr {

namespace {

// Tile widths of one instance (0 = the default for its precision).
struct Widths {
  int fwd = 0, dual = 0;
};

PlanParams fwd_params(uint32_t n, int prec, int k = 0) {
  PlanParams pp;
  pp.tile_bits = std::min<int>(k ? k : (prec == 0 ? 12 : 13), (int)n);
  pp.reg_bits = tile_reg_bits(pp.tile_bits);
  pp.low_bits = prec == 0 ? 3 : 4;
  if (pp.tile_bits == (int)n) pp.max_ops = kWholeStateStageOps;
  return pp;
}

PlanParams dual_params(uint32_t n, int prec, int k = 0) {
  PlanParams pp;
  pp.tile_bits = std::min<int>(k ? k : (prec == 0 ? 10 : 11), (int)n);
  pp.reg_bits = std::min(4, pp.tile_bits);
  pp.low_bits = prec == 0 ? 3 : 4;
  if (pp.tile_bits == (int)n) pp.max_ops = kWholeStateStageOps;
  return pp;
}

void check_unitary_customs(const std::vector<ExactGate>& gates) {
  for (const auto& e : gates)
    if (e.custom) {
      const int dim = 1 << e.K;
      for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
          cd s(0, 0);
          for (int k = 0; k < dim; ++k) s += std::conj(e.M[k * dim + r]) * e.M[k * dim + c];
          if (std::abs(s - cd(r == c ? 1.0 : 0.0, 0.0)) > 1e-10)
            throw Error{QR_ERR_UNSUPPORTED, "adjoint differentiation needs unitary custom gates"};
        }
    }
}

// Lowers one instance's program (forward ops and the reverse-sweep ops).
void lower_instance(Inst& I, const qr_program* prog) {
  validate_program(I.n, prog);
  const std::vector<ExactGate> gates = lower_exact(I.n, prog);
  check_unitary_customs(gates);
  I.fwd = lower_fast(I.n, gates, true);
  I.rev = adjoint_sweep_ops(I.fwd);
}

std::vector<uint64_t> structure_key(const std::vector<FastOp>& ops) {
  std::vector<uint64_t> k;
  k.reserve(3 * ops.size());
  for (const auto& op : ops) plan_key_append(op, k);
  return k;
}

size_t densek_tmp_bytes(const FastOp& op, uint64_t dim, int prec) {
  if (op.kind != FK_DENSEK) return 0;
  const uint64_t dimk = 1ull << op.k, ngroups = dim >> op.k;
  uint64_t tg = std::min<uint64_t>(ngroups, std::max<uint64_t>(256, (64ull << 20) / (dimk * 16)));
  tg = std::max<uint64_t>(tg / 256, 1) * 256;
  if (ngroups < 256) tg = ngroups;
  return tg * dimk * (prec == 0 ? 16 : 8);
}

// Writes every instance's plan blobs and Pauli term tables at their offsets.
void fill_blob(qr_batch* b) {
  for (Inst& I : b->inst) {
    std::vector<Loc> lf, lb;
    const std::vector<uint8_t> f = pack_blob(I.fwd, I.pf->stages, lf);
    const std::vector<uint8_t> r = pack_blob(I.rev, I.pb->stages, lb);
    if (f.size() != I.blob_f_size || r.size() != I.blob_b_size)
      throw Error{QR_ERR_ARG, "program structure differs from the batch's"};
    std::memcpy(b->host_blob.data() + I.blob_f, f.data(), f.size());
    std::memcpy(b->host_blob.data() + I.blob_b, r.data(), r.size());
  }
}

void launch_instance(qr_batch* b, const Inst& I, hipStream_t s, uint64_t* stamps, size_t idx) {
  const int prec = b->prec;
  uint8_t* base = (uint8_t*)b->states.p;
  void* psi = base + I.psi;
  void* lam = base + I.lam;
  const uint8_t* dblob = (const uint8_t*)b->blob.p;
  double* work = (double*)b->work.p;
  double* outd = (double*)b->out.p;
  hip_check(launch_init_basis(prec, psi, I.dim, 0, s), "init");
  // forward
  for (size_t si = 0; si < I.pf->stages.size(); ++si) {
    const Stage& st = I.pf->stages[si];
    if (st.tile) {
      const TileArgs a = tile_args_at(I.n, st, dblob + I.blob_f, I.locf[si]);
      hip_check(jit_launch(I.pf->fns[si], st.k, I.ppf.reg_bits, false, psi, a.coefs, a.ntiles, a.nontile_mask, s),
                "compiled tile kernel");
    } else {
      launch_standalone_raw(prec, I.dim, psi, b->tmp[I.stream], I.fwd[st.standalone], dblob + I.blob_f, I.locf[si], s);
    }
  }
  // lam = O psi (one pass), E = Re <psi|lam>
  hip_check(launch_pauli_sum(prec, lam, psi, I.dim, (const PauliGroup*)(dblob + I.obs_groups), I.ngroups,
                             (const PauliZTerm*)(dblob + I.obs_terms), s),
            "pauli sum");
  hip_check(launch_dot_re_partial(prec, psi, lam, I.dim, work + I.dot, kDotBlocks, s), "dot");
  hip_check(launch_sum_rows(work + I.dot, kDotBlocks, 1, outd + idx, s), "expectation sum");
  // reverse sweep
  for (size_t si = 0; si < I.pb->stages.size(); ++si) {
    const Stage& st = I.pb->stages[si];
    if (st.tile) {
      const TileArgs a = tile_args_at(I.n, st, dblob + I.blob_b, I.locb[si]);
      hip_check(jit_launch_dual(I.pb->fns[si], st.k, I.ppb.reg_bits, psi, lam, a.coefs, a.ntiles, a.nontile_mask,
                                work + I.part, I.nslots, s),
                "compiled adjoint stage");
    } else {
      const FastOp& op = I.rev[st.standalone];
      launch_standalone_raw(prec, I.dim, psi, b->tmp[I.stream], op, dblob + I.blob_b, I.locb[si], s);
      launch_standalone_raw(prec, I.dim, lam, b->tmp[I.stream], op, dblob + I.blob_b, I.locb[si], s);
    }
  }
  const size_t m = b->inst.size();
  if (I.nparams)
    hip_check(launch_sum_rows(work + I.part, I.nslots, (uint32_t)I.nparams, outd + m + I.grad, s), "gradient sums");
  hip_check(launch_stamp(stamps + 1 + idx, s), "stamp");
}


qr_batch* batch_create_widths(uint64_t m, const uint32_t* nq, const qr_program* programs,
                              const qr_observable* observables, qr_precision precision,
                              const std::vector<Widths>& widths, int nstreams);
void autotune_widths(uint64_t m, const uint32_t* nq, const qr_program* programs, const qr_observable* observables,
                     qr_precision precision, std::vector<Widths>& widths, bool tune_fwd, bool tune_dual);

}  // namespace

qr_batch* batch_create(uint64_t m, const uint32_t* nq, const qr_program* programs, const qr_observable* observables,
                       qr_precision precision, const qr_batch_options* o) {
  if (!m) throw Error{QR_ERR_ARG, "a batch needs at least one instance"};
  if (!nq || !programs || !observables) throw Error{QR_ERR_ARG, "NULL argument"};
  const int maxk = tile_max_bits(precision == QR_C128 ? 0 : 1);
  if (o && ((int)o->fwd_tile_bits > maxk || (int)o->dual_tile_bits > maxk))
    throw Error{QR_ERR_ARG, "tile width above " + std::to_string(maxk)};
  std::vector<Widths> widths(m);
  for (auto& w : widths) {
    w.fwd = o ? (int)o->fwd_tile_bits : 0;
    w.dual = o ? (int)o->dual_tile_bits : 0;
  }
  // 4 streams (the process's hardware queues): the same throughput as 8 or 16 within 3%, and
  // each stream finishes its small instances early (512-instance batch: p50 41 ms against 142)
  const int nstreams = (o && o->streams) ? (int)std::min<uint32_t>(o->streams, 32) : 4;
  if (o && o->autotune && (!o->fwd_tile_bits || !o->dual_tile_bits))
    autotune_widths(m, nq, programs, observables, precision, widths, !o->fwd_tile_bits, !o->dual_tile_bits);
  return batch_create_widths(m, nq, programs, observables, precision, widths, nstreams);
}

namespace {

qr_batch* batch_create_widths(uint64_t m, const uint32_t* nq, const qr_program* programs,
                              const qr_observable* observables, qr_precision precision,
                              const std::vector<Widths>& widths, int nstreams) {
  if (!m) throw Error{QR_ERR_ARG, "a batch needs at least one instance"};
  if (!nq || !programs || !observables) throw Error{QR_ERR_ARG, "NULL argument"};
  if (precision != QR_C128 && precision != QR_C64) throw Error{QR_ERR_ARG, "unknown precision"};
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error{QR_ERR_NO_DEVICE, "no GPU visible"};
  auto b = std::make_unique<qr_batch>();
  b->prec = precision == QR_C128 ? 0 : 1;
  const size_t B = b->prec == 0 ? 16 : 8;
  b->inst.resize(m);
  // ---- lower and plan -----------------------------------------------------------------------
  for (uint64_t i = 0; i < m; ++i) {
    Inst& I = b->inst[i];
    I.n = nq[i];
    if (I.n == 0 || I.n > 34) throw Error{QR_ERR_ARG, "instance " + std::to_string(i) + ": qubit count out of range"};
    I.dim = 1ull << I.n;
    lower_instance(I, &programs[i]);
    I.key = structure_key(I.fwd);
    I.terms = parse_observable(I.n, &observables[i]);
    I.ppf = fwd_params(I.n, b->prec, widths[i].fwd);
    I.ppb = dual_params(I.n, b->prec, widths[i].dual);
    I.pf = get_plan_for(I.n, b->prec, I.fwd, I.ppf, true, false, false);
    I.pb = get_plan_for(I.n, b->prec, I.rev, I.ppb, true, true, false);
    for (const auto& op : I.fwd)
      if (op.param >= 0) {
        I.nparams++;
        I.gens.push_back(op.gen);
      }
    I.nslots = jit_dual_slots((int)I.n, I.ppb.tile_bits, I.ppb.reg_bits);
    I.work = (double)I.dim * B * (2.0 * I.pf->stages.size() + 4.0 * I.pb->stages.size());
    for (const auto& op : I.fwd) I.tmp_bytes = std::max(I.tmp_bytes, densek_tmp_bytes(op, I.dim, b->prec));
    for (const auto& op : I.rev) I.tmp_bytes = std::max(I.tmp_bytes, densek_tmp_bytes(op, I.dim, b->prec));
  }
  // ---- compile every distinct stage once (in parallel), then resolve the kernels ------------
  {
    std::vector<std::string> all;
    for (Inst& I : b->inst) {
      if (I.pf->fns.empty()) {
        auto s = jit_stage_sources(I.pf->stages, b->prec, I.ppf.reg_bits, false, false);
        all.insert(all.end(), s.begin(), s.end());
      }
      if (I.pb->fns.empty()) {
        auto s = jit_stage_sources(I.pb->stages, b->prec, I.ppb.reg_bits, false, true);
        all.insert(all.end(), s.begin(), s.end());
      }
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    jit_prepare(all);
    for (Inst& I : b->inst) {
      prepare_compiled(*I.pf, b->prec, I.ppf.reg_bits, false, false);
      prepare_compiled(*I.pb, b->prec, I.ppb.reg_bits, false, true);
    }
  }
  // ---- layout ---------------------------------------------------------------------------------
  size_t blob_size = 0, state_bytes = 0, work_doubles = 0;
  for (Inst& I : b->inst) {
    std::vector<Loc> tmp;
    I.blob_f = blob_size = align_up(blob_size, kAlign);
    I.blob_f_size = pack_blob(I.fwd, I.pf->stages, I.locf).size();
    blob_size += I.blob_f_size;
    I.blob_b = blob_size = align_up(blob_size, kAlign);
    I.blob_b_size = pack_blob(I.rev, I.pb->stages, I.locb).size();
    blob_size += I.blob_b_size;
    // the observable's tables: terms grouped by X mask
    {
      std::vector<PauliGroup> groups;
      std::vector<PauliZTerm> flat;
      pauli_tables(I.terms, groups, flat);
      I.ngroups = (int)groups.size();
      I.obs_groups = blob_size = align_up(blob_size, 16);
      blob_size += groups.size() * sizeof(PauliGroup);
      I.obs_terms = blob_size = align_up(blob_size, 16);
      blob_size += flat.size() * sizeof(PauliZTerm);
    }
    I.psi = state_bytes = align_up(state_bytes, kAlign);
    state_bytes += I.dim * B;
    I.lam = state_bytes = align_up(state_bytes, kAlign);
    state_bytes += I.dim * B;
    I.part = work_doubles;
    work_doubles += std::max<uint64_t>(I.nparams, 1) * I.nslots;
    I.dot = work_doubles;
    work_doubles += kDotBlocks;
    I.grad = b->total_params;
    b->total_params += I.nparams;
  }
  b->host_blob.assign(blob_size, 0);
  for (Inst& I : b->inst) {  // observable tables (fixed for the batch's lifetime)
    std::vector<PauliGroup> groups;
    std::vector<PauliZTerm> flat;
    pauli_tables(I.terms, groups, flat);
    if (!groups.empty())
      std::memcpy(b->host_blob.data() + I.obs_groups, groups.data(), groups.size() * sizeof(PauliGroup));
    if (!flat.empty()) std::memcpy(b->host_blob.data() + I.obs_terms, flat.data(), flat.size() * sizeof(PauliZTerm));
  }
  fill_blob(b.get());
  const size_t out_doubles = m + b->total_params;
  b->states.reserve(std::max<size_t>(state_bytes, 1));
  b->blob.reserve(std::max<size_t>(blob_size, 1));
  b->work.reserve(std::max<size_t>(work_doubles, 1) * sizeof(double));
  b->out.reserve(out_doubles * sizeof(double) + (m + 1) * sizeof(uint64_t) + kAlign);
  // ---- streams: largest instances first, each to the least loaded stream ---------------------
  const int S = (int)std::min<uint64_t>((uint64_t)nstreams, m);
  hip_check(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking), "hipStreamCreate");
  b->streams.resize(S);
  b->tmp.resize(S);
  for (int s = 0; s < S; ++s) hip_check(hipStreamCreateWithFlags(&b->streams[s], hipStreamNonBlocking), "hipStreamCreate");
  std::vector<size_t> order(m);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->inst[x].work > b->inst[y].work; });
  std::vector<double> load(S, 0.0);
  std::vector<std::vector<size_t>> per_stream(S);
  for (size_t i : order) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->inst[i].stream = s;
    load[s] += b->inst[i].work;
    per_stream[s].push_back(i);
  }
  // inside a stream, smallest first: the stream's total is unchanged, and most instances
  // complete early instead of waiting behind the large ones
  for (auto& list : per_stream)
    std::stable_sort(list.begin(), list.end(), [&](size_t x, size_t y) { return b->inst[x].work < b->inst[y].work; });
  for (int s = 0; s < S; ++s) {
    size_t need = 0;
    for (size_t i : per_stream[s]) need = std::max(need, b->inst[i].tmp_bytes);
    if (need) b->tmp[s].reserve(need);
  }
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  int khz = 0;
  hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
  b->clock_khz = khz > 0 ? khz : 100000.0;
  // ---- capture ----------------------------------------------------------------------------------
  uint64_t* stamps = (uint64_t*)((uint8_t*)b->out.p + align_up(out_doubles * sizeof(double), kAlign));
  b->events.resize(S + 1);
  for (auto& e : b->events) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    hip_check(launch_stamp(stamps, b->origin), "stamp");
    hip_check(hipEventRecord(b->events[S], b->origin), "hipEventRecord");
    for (int s = 0; s < S; ++s) {
      hip_check(hipStreamWaitEvent(b->streams[s], b->events[S], 0), "hipStreamWaitEvent");
      for (size_t i : per_stream[s]) launch_instance(b.get(), b->inst[i], b->streams[s], stamps, i);
      hip_check(hipEventRecord(b->events[s], b->streams[s]), "hipEventRecord");
      hip_check(hipStreamWaitEvent(b->origin, b->events[s], 0), "hipStreamWaitEvent");
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->origin, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(b->origin, &b->graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  return b.release();
}

}  // namespace

void batch_run(qr_batch* b, const qr_program* programs, double* expvals, double* grads, const uint64_t* grad_offsets,
               double* latency_us) {
  const size_t m = b->inst.size();
  if (programs) {
    // lower everything first; the batch keeps its previous programs if any instance fails
    std::vector<Inst> fresh(m);
    for (size_t i = 0; i < m; ++i) {
      fresh[i].n = b->inst[i].n;
      lower_instance(fresh[i], &programs[i]);
      if (structure_key(fresh[i].fwd) != b->inst[i].key)
        throw Error{QR_ERR_ARG, "instance " + std::to_string(i) + ": program structure differs from the batch's"};
    }
    for (size_t i = 0; i < m; ++i) {
      b->inst[i].fwd = std::move(fresh[i].fwd);
      b->inst[i].rev = std::move(fresh[i].rev);
    }
    fill_blob(b);
  }
  hip_check(hipMemcpyAsync(b->blob.p, b->host_blob.data(), b->host_blob.size(), hipMemcpyHostToDevice, b->origin),
            "plan upload");
  hip_check(hipGraphLaunch(b->exec, b->origin), "hipGraphLaunch");
  const size_t out_doubles = m + b->total_params;
  std::vector<double> h(out_doubles);
  std::vector<uint64_t> st(m + 1);
  hip_check(hipMemcpyAsync(h.data(), b->out.p, out_doubles * sizeof(double), hipMemcpyDeviceToHost, b->origin), "D2H");
  hip_check(hipMemcpyAsync(st.data(), (uint8_t*)b->out.p + align_up(out_doubles * sizeof(double), kAlign),
                           (m + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, b->origin),
            "D2H");
  hip_check(hipStreamSynchronize(b->origin), "sync");
  for (size_t i = 0; i < m; ++i) {
    const Inst& I = b->inst[i];
    if (expvals) expvals[i] = h[i];
    if (grads) {
      double* g = grads + (grad_offsets ? grad_offsets[i] : I.grad);
      for (uint64_t p = 0; p < I.nparams; ++p) g[p] = I.gens[p] == 4 ? -h[m + I.grad + p] : h[m + I.grad + p];
    }
    if (latency_us) latency_us[i] = (double)(st[1 + i] - st[0]) * 1000.0 / b->clock_khz;
  }
}

void batch_param_counts(const qr_batch* b, uint64_t* out) {
  for (size_t i = 0; i < b->inst.size(); ++i) out[i] = b->inst[i].nparams;
}

void batch_destroy(qr_batch* b) { delete b; }

namespace {

std::mutex g_tune_mu;
std::map<std::pair<uint32_t, int>, Widths> g_tuned;  // (qubits, precision) -> tuned tile widths

// For every qubit count >= 16: the forward-sweep tile widths around the default (reverse
// width at its default), then the reverse-sweep widths around the default with the fastest
// forward width, timed on the instance with the longest program (one-instance batches, best
// of three runs). The stages of every candidate are compiled together first, in one parallel
// compilation. Widths the caller fixed are kept. Results are cached per process.
void autotune_widths(uint64_t m, const uint32_t* nq, const qr_program* programs, const qr_observable* observables,
                     qr_precision precision, std::vector<Widths>& widths, bool tune_fwd, bool tune_dual) {
  const int prec = precision == QR_C128 ? 0 : 1;
  const int def_f = prec == 0 ? 12 : 13, def_d = prec == 0 ? 10 : 11;
  std::map<uint32_t, uint64_t> rep;
  for (uint64_t i = 0; i < m; ++i)
    if (nq[i] >= 16 && (!rep.count(nq[i]) || programs[i].nops > programs[rep[nq[i]]].nops)) rep[nq[i]] = i;
  auto cands = [&](int def, uint32_t n, int fixed) {
    std::vector<int> v;
    if (fixed) return std::vector<int>{fixed};
    for (int k : {def - 1, def, def + 1})
      if (k >= 6 && k <= tile_max_bits(prec) && k <= (int)n) v.push_back(k);
    return v;
  };
  // ---- compile every candidate's stages at once
  std::vector<uint32_t> todo;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& kv : rep)
      if (!g_tuned.count({kv.first, prec})) todo.push_back(kv.first);
  }
  {
    std::vector<std::string> all;
    for (uint32_t n : todo) {
      const uint64_t i = rep[n];
      Inst I;
      I.n = n;
      I.dim = 1ull << n;
      lower_instance(I, &programs[i]);
      for (int k : cands(def_f, n, tune_fwd ? 0 : widths[i].fwd)) {
        const PlanParams pp = fwd_params(n, prec, k);
        auto cp = get_plan_for(n, prec, I.fwd, pp, true, false, false);
        if (cp->fns.empty()) {
          auto src = jit_stage_sources(cp->stages, prec, pp.reg_bits, false, false);
          all.insert(all.end(), src.begin(), src.end());
        }
      }
      for (int k : cands(def_d, n, tune_dual ? 0 : widths[i].dual)) {
        const PlanParams pp = dual_params(n, prec, k);
        auto cp = get_plan_for(n, prec, I.rev, pp, true, true, false);
        if (cp->fns.empty()) {
          auto src = jit_stage_sources(cp->stages, prec, pp.reg_bits, false, true);
          all.insert(all.end(), src.begin(), src.end());
        }
      }
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    jit_prepare(all);
  }
  // ---- time the candidates
  auto time_widths = [&](uint64_t i, Widths w) {
    std::vector<Widths> ws(1, w);
    std::unique_ptr<qr_batch, void (*)(qr_batch*)> b(
        batch_create_widths(1, &nq[i], &programs[i], &observables[i], precision, ws, 1), batch_destroy);
    std::vector<double> g(std::max<uint64_t>(b->total_params, 1));
    double e = 0.0;
    batch_run(b.get(), nullptr, &e, g.data(), nullptr, nullptr);  // warm-up
    double t = 1e300;
    for (int r = 0; r < 3; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      batch_run(b.get(), nullptr, &e, g.data(), nullptr, nullptr);
      t = std::min(t, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    return t;
  };
  for (uint32_t n : todo) {
    const uint64_t i = rep[n];
    const int dual0 = tune_dual ? def_d : widths[i].dual;
    Widths best{cands(def_f, n, tune_fwd ? 0 : widths[i].fwd)[0], dual0};
    double best_t = 1e300;
    for (int k : cands(def_f, n, tune_fwd ? 0 : widths[i].fwd)) {
      const double t = time_widths(i, Widths{k, dual0});
      if (t < best_t) {
        best_t = t;
        best = Widths{k, dual0};
      }
    }
    if (tune_dual) {
      for (int k : cands(def_d, n, 0)) {
        if (k == dual0) continue;  // timed above
        const double t = time_widths(i, Widths{best.fwd, k});
        if (t < best_t) {
          best_t = t;
          best.dual = k;
        }
      }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{n, prec}] = best;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto& kv : rep) {
    const Widths w = g_tuned[{kv.first, prec}];
    for (uint64_t j = 0; j < m; ++j)
      if (nq[j] == kv.first) {
        if (tune_fwd) widths[j].fwd = w.fwd;
        if (tune_dual) widths[j].dual = w.dual;
      }
  }
}

}  // namespace

}  // namespace qr
