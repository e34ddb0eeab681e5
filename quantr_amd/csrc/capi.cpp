// capi.cpp — the C ABI of include/quantr_hip.h.
//
// Entry points only: execution lives in runtime.cpp and batch.cpp, lowering in lower.cpp,
// sweep planning in planner.cpp and the kernels in kernels.hip. Every entry point catches
// errors and reports them as a qr_status plus a thread-local message; nothing unwinds across
// the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace qr {
const char* last_error_cstr();
}

using namespace qr;

namespace {

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
    if (o->stage_ops > 65536) throw Error{QR_ERR_ARG, "stage_ops must be <= 65536"};
    if (o->tune > 1) throw Error{QR_ERR_ARG, "tune must be 0 or 1"};
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

qr_status qr_sample_programs(uint32_t num_qubits, qr_precision precision, const qr_program* programs,
                             uint64_t count, const uint64_t* draws_per_program, const double* uniforms,
                             uint64_t* idx_out) {
  return guard([&] {
    if (precision != QR_C128 && precision != QR_C64) throw Error{QR_ERR_ARG, "unknown precision"};
    sample_programs(num_qubits, precision == QR_C128 ? 0 : 1, programs, count, draws_per_program, uniforms, idx_out);
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
    if (st->opt.exact_sample == 1) {
      // checker mode: one lane folds the whole state in index order
      std::vector<uint8_t> blob(shots * 8);
      std::memcpy(blob.data(), us.data(), shots * 8);
      uint8_t* d = upload(st, st->aux, blob);
      st->aux2.reserve(shots * 8);
      ProfScope ps_(st, "sample_fold", (double)st->dim * st->bytes_per_amp());
      hip_check(launch_sample_fold(st->prec, st->amp, st->dim, (const double*)d, shots, (uint64_t*)st->aux2.p,
                                   st->stream),
                "sample fold");
      hip_check(hipMemcpyAsync(res.data(), st->aux2.p, shots * 8, hipMemcpyDeviceToHost, st->stream), "D2H");
      hip_check(hipStreamSynchronize(st->stream), "sync");
    } else {
      sample_exact_parallel(st, us, res);
    }
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

qr_status qr_expectation_noalloc(qr_state* st, const qr_observable* observable, double* out) {
  return guard([&] {
    if (!st || !out) throw Error{QR_ERR_ARG, "NULL argument"};
    *out = expectation_noalloc(st, observable);
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
    apply_observable(ps->prec, ls->amp, ps->amp, ps->dim, g->terms, ls->aux, s);
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

qr_status qr_batch_create(uint64_t ninstances, const uint32_t* num_qubits, const qr_program* programs,
                          const qr_observable* observables, qr_precision precision, const qr_batch_options* options,
                          qr_batch** out) {
  return guard([&] {
    if (!out) throw Error{QR_ERR_ARG, "out is NULL"};
    *out = nullptr;
    *out = batch_create(ninstances, num_qubits, programs, observables, precision, options);
  });
}

qr_status qr_batch_destroy(qr_batch* batch) {
  return guard([&] { batch_destroy(batch); });
}

qr_status qr_batch_param_counts(const qr_batch* batch, uint64_t* counts_out) {
  return guard([&] {
    if (!batch || !counts_out) throw Error{QR_ERR_ARG, "NULL argument"};
    batch_param_counts(batch, counts_out);
  });
}

qr_status qr_batch_run(qr_batch* batch, const qr_program* programs, double* expvals_out, double* grads_out,
                       const uint64_t* grad_offsets, double* latency_us_out) {
  return guard([&] {
    if (!batch) throw Error{QR_ERR_ARG, "batch is NULL"};
    batch_run(batch, programs, expvals_out, grads_out, grad_offsets, latency_us_out);
  });
}

qr_status qr_batch_grad(uint64_t ninstances, const uint32_t* num_qubits, const qr_program* programs,
                        const qr_observable* observables, qr_precision precision, double* expvals_out,
                        double* grads_out, const uint64_t* grad_offsets, double* latency_us_out) {
  return guard([&] {
    if (!ninstances) return;
    if (!expvals_out) throw Error{QR_ERR_ARG, "NULL argument"};
    std::unique_ptr<qr_batch, void (*)(qr_batch*)> b(batch_create(ninstances, num_qubits, programs, observables, precision,
                                                                  nullptr),
                                                     [](qr_batch* x) { batch_destroy(x); });
    batch_run(b.get(), nullptr, expvals_out, grads_out, grad_offsets, latency_us_out);
  });
}

qr_status qr_measure_stream(qr_state* st, uint32_t reps, double* ms_out) {
  return guard([&] {
    if (!st || !ms_out) throw Error{QR_ERR_ARG, "NULL argument"};
    if (reps == 0) throw Error{QR_ERR_ARG, "reps must be at least 1"};
    hip_check(launch_stream_scale(st->prec, st->amp, st->dim, 1.0, st->stream), "stream kernel");  // warm-up
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    hipError_t e = hipEventRecord(a, st->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
      ProfScope ps_(st, "stream_scale", 2.0 * (double)st->dim * st->bytes_per_amp());
      e = launch_stream_scale(st->prec, st->amp, st->dim, 1.0, st->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(b, st->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    hip_check(e, "stream timing");
    *ms_out = (double)ms / reps;
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
