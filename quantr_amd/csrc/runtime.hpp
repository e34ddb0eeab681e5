// runtime.hpp — device states, plan cache and launch helpers shared by the C ABI
// (capi.cpp) and the batched executor (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"

namespace qr {

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    throw Error{QR_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e)};
  throw Error{QR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
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


}  // namespace qr

struct qr_state {
  uint32_t n = 0;
  int prec = 0;  // 0 = c128, 1 = c64
  uint64_t dim = 0;
  void* amp = nullptr;
  hipStream_t stream = nullptr;
  int device = 0;
  qr_options opt{1, 0, 0, 2, 0, 0, 2, 0, 0};
  qr::DevBuf aux;   // plan / coefficients / small uploads
  qr::DevBuf aux2;  // partial sums, sampler scratch
  qr::DevBuf tmp;   // wide-gate staging (bounded, never state-sized)
  qr_plan_stats last{};
  bool profiling = false;
  std::vector<qr::ProfRec> prof;
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
  std::vector<qr::FastOp> ops;  // forward program (fast form, with parameters)
  using Term = qr::PauliMasks;
  std::vector<Term> terms;
  uint64_t nparams = 0;
  bool have_forward = false;
};

namespace qr {

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

// ---- plans ----------------------------------------------------------------------------------
PlanParams plan_params(const qr_state* st);       // forward sweeps
// The same for a state of n qubits, precision prec (0 = complex128) and options opt.
PlanParams plan_params(uint32_t n, int prec, const qr_options& opt);
PlanParams dual_plan_params(const qr_state* st);  // adjoint sweeps (psi and lam per tile)

// Offsets of one stage's data in a packed plan blob.
struct Loc {
  size_t ops = 0, maps = 0, coefs = 0, off = 0;
};

// A plan's stages plus the device copy of its blob.
struct PackedPlan {
  std::vector<Stage> stages;
  std::vector<Loc> loc;
  uint8_t* dblob = nullptr;
};

struct TuneState;

// A plan shared by every program of one structure, with its compiled stages.
struct CachedPlan {
  std::vector<Stage> stages;
  std::mutex compile_mu;           // held while this plan's stages compile (not the cache lock)
  std::vector<std::string> src;    // compiled-stage sources (filled on first compiled use)
  std::vector<hipFunction_t> fns;  // their kernels
  std::atomic<uint64_t> runs{0};  // applications (states on several threads may share a plan)
  uint64_t stamp = 0;
  std::mutex tune_mu;               // guards `tune`
  std::shared_ptr<TuneState> tune;  // measured selection (qr_options.tune), made on first use
};

// Measured selection for one structure. Applications with qr_options.tune measure, in turn:
// each candidate plan with the default tile order (step 0..C-1), then the fastest candidate
// with the other order (step C; the orders are index order and the XCDs' contiguous eighths).
// Each tile sweep then keeps the faster of its two orders. `mu` serialises the measuring
// applications of the structure.
struct TuneState {
  std::mutex mu;
  std::vector<std::shared_ptr<CachedPlan>> cands;  // [0] is the plan the estimate picks
  std::vector<double> cand_ms;                     // per candidate, sum of its sweeps' times
  std::vector<std::vector<float>> stage_ms;        // per candidate, per stage (index order)
  int step = 0;                                    // next measurement; C + 1 = done
  int chosen = 0;
  std::vector<uint32_t> order;  // per stage of the chosen plan: 0 index order, 1 per-XCD eighths
};

std::shared_ptr<CachedPlan> get_plan(const qr_state* st, const std::vector<FastOp>& ops, const PlanParams& pp,
                                     bool fused, bool dual, bool persistent);
std::shared_ptr<CachedPlan> get_plan_for(uint32_t n, int prec, const std::vector<FastOp>& ops, const PlanParams& pp,
                                         bool fused, bool dual, bool persistent);
void prepare_compiled(CachedPlan& cp, int prec, int RB, bool persistent, bool dual);
// Host image of a plan's blob (stage ops, coefficients from `ops`, maps, LDS patterns and the
// standalone tables); loc receives each stage's offsets.
std::vector<uint8_t> pack_blob(const std::vector<FastOp>& ops, const std::vector<Stage>& stages, std::vector<Loc>& loc);
void pack_plan(qr_state* st, const std::vector<FastOp>& ops, PackedPlan& P);
TileArgs tile_args_at(uint32_t n, const Stage& s, const uint8_t* dblob, const Loc& loc);
// One standalone op on a state of 2^n amplitudes (tmp: staging for wide dense gates).
void launch_standalone_raw(int prec, uint64_t dim, void* amp, DevBuf& tmp, const FastOp& op, const uint8_t* dblob,
                           const Loc& loc, hipStream_t s);

// ---- execution ------------------------------------------------------------------------------
qr_state* new_state(uint32_t n, qr_precision p);
void run_fast(qr_state* st, const std::vector<FastOp>& ops, bool fused, qr_plan_stats* stats);
void run_exact(qr_state* st, const std::vector<ExactGate>& gates, qr_plan_stats* stats);
void do_apply(qr_state* st, const qr_program* prog);
double sum_partials_host(const std::vector<double>& v);
// Compensated (Neumaier) sum of count partials, in index order.
double sum_partials_host(const double* v, size_t count);
// measure_all's sampling (sampler.cpp): draws `us` in ascending order; res[s] = the first index
// whose left-fold cumulative probability exceeds us[s], or UINT64_MAX (None).
void sample_exact_parallel(qr_state* st, const std::vector<double>& us, std::vector<uint64_t>& res);
// measure_all_without_cache's re-simulation, batched (shots.cpp): programs[j] simulated from
// |0...0>, then its draws[j] uniforms (consecutive in u) placed by the left fold into out.
void sample_programs(uint32_t n, int prec, const qr_program* programs, uint64_t count, const uint64_t* draws,
                     const double* u, uint64_t* out);
double norm_sqr(qr_state* st);
double expectation(qr_state* st, const std::vector<PauliMasks>& terms);
// The same value through fixed device and host arrays: no allocation (noalloc.cpp).
double expectation_noalloc(qr_state* st, const qr_observable* o);
// lam = O psi for an observable, on stream s (one pass over the state).
void apply_observable(int prec, void* lam, const void* psi, uint64_t dim, const std::vector<PauliMasks>& terms,
                      DevBuf& scratch, hipStream_t s);
void backward_fused(qr_grad* g, double grad_out, double* grads_out);

// ---- batches (batch.cpp) ------------------------------------------------------------------------
qr_batch* batch_create(uint64_t m, const uint32_t* nq, const qr_program* programs, const qr_observable* observables,
                       qr_precision precision, const qr_batch_options* options);
void batch_run(qr_batch* b, const qr_program* programs, double* expvals, double* grads, const uint64_t* grad_offsets,
               double* latency_us);
void batch_param_counts(const qr_batch* b, uint64_t* out);
void batch_destroy(qr_batch* b);

}  // namespace qr
