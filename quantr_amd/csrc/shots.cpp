// shots.cpp — measure_all_without_cache's per-shot re-simulation, batched
// (SimulatedCircuit::measure_all_without_cache, simulated_circuit.rs:81-114).
//
// The reference re-simulates the whole circuit on |0...0> for every shot, because custom gate
// functions may answer differently each time (non-unitary channels such as post-selection,
// examples/post_select.rs:39-49), then measures that state once. Here the caller tabulates
// each shot's program on the host (the custom functions are called there, in shot order) and
// hands all of them over at once: programs of one structure share one plan and one set of
// compiled stages, their states sit back to back, and each stage runs as ONE launch over all
// of them (grid y = program, coefficients at a fixed stride). Each program's draws are then
// placed by the reference's left fold (one lane per state, binary64, no contraction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <map>
#include <numeric>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace qr {

namespace {

constexpr uint64_t kChunkBytes = 4ull << 30;  // states held at once
constexpr uint32_t kMaxGridY = 65535;

struct Prepared {
  std::vector<FastOp> ops;
  std::vector<uint64_t> key;
};

// fn(i) for i in [0, count) on the host's cores; the error of the lowest failing i is rethrown.
template <class F>
void parallel_for(uint64_t count, F fn) {
  const unsigned nt = (unsigned)std::min<uint64_t>(count, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<std::exception_ptr> err(count);
  std::atomic<uint64_t> next{0};
  auto work = [&] {
    for (uint64_t i = next++; i < count; i = next++) {
      try {
        fn(i);
      } catch (...) {
        err[i] = std::current_exception();
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace

void sample_programs(uint32_t n, int prec, const qr_program* programs, uint64_t count, const uint64_t* draws,
                     const double* u, uint64_t* out) {
  if (n == 0) throw Error{QR_ERR_ARG, "The initialised circuit must have at least one wire."};
  if (n > 34) throw Error{QR_ERR_UNSUPPORTED, "at most 34 qubits"};
  if (count == 0) return;
  if (!programs || !draws) throw Error{QR_ERR_ARG, "NULL argument"};
  uint64_t total = 0;
  std::vector<uint64_t> first(count + 1, 0);
  for (uint64_t j = 0; j < count; ++j) {
    if (draws[j] > (1ull << 40) - total) throw Error{QR_ERR_ARG, "more than 2^40 draws in all"};
    first[j] = total;
    total += draws[j];
  }
  first[count] = total;
  if (total && (!u || !out)) throw Error{QR_ERR_ARG, "NULL argument"};
  for (uint64_t d = 0; d < total; ++d)
    if (u[d] != u[d]) throw Error{QR_ERR_ARG, "uniform draws must not be NaN"};
  // ---- lower every program; group them by structure (first appearance order)
  std::vector<Prepared> pr(count);
  std::map<std::vector<uint64_t>, std::vector<uint64_t>> groups;
  std::vector<std::vector<uint64_t>> order;
  parallel_for(count, [&](uint64_t j) {
    validate_program(n, &programs[j]);
    pr[j].ops = lower_fast(n, lower_exact(n, &programs[j]), false);
    for (const FastOp& op : pr[j].ops) plan_key_append(op, pr[j].key);
  });
  for (uint64_t j = 0; j < count; ++j) {
    auto it = groups.find(pr[j].key);
    if (it == groups.end()) {
      groups[pr[j].key] = {j};
      order.push_back(pr[j].key);
    } else {
      it->second.push_back(j);
    }
  }
  int count_dev = 0;
  if (hipGetDeviceCount(&count_dev) != hipSuccess || count_dev == 0) throw Error{QR_ERR_NO_DEVICE, "no GPU visible"};
  // planning parameters of a state of this size and precision with default options
  const qr_options defaults{1, 0, 0, 2, 0, 0, 2, 0, 0};
  const PlanParams pp = plan_params(n, prec, defaults);
  const uint64_t dim = 1ull << n;
  const size_t B = prec == 0 ? 16 : 8;
  hipStream_t s = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf states, blobs, tmp, dev_u, dev_off, dev_out;
  for (const auto& key : order) {
    const std::vector<uint64_t>& members = groups[key];
    std::shared_ptr<CachedPlan> cp = get_plan_for(n, prec, pr[members[0]].ops, pp, true, false, false);
    prepare_compiled(*cp, prec, pp.reg_bits, false, false);
    const std::vector<Stage>& stages = cp->stages;
    std::vector<Loc> loc;
    const size_t blob_size = pack_blob(pr[members[0]].ops, stages, loc).size();
    const size_t blob_stride = (blob_size + 255) & ~(size_t)255;
    const uint64_t per_chunk =
        std::max<uint64_t>(1, std::min<uint64_t>(kMaxGridY, kChunkBytes / (dim * B)));
    for (size_t c0 = 0; c0 < members.size(); c0 += per_chunk) {
      const uint64_t cnt = std::min<uint64_t>(per_chunk, members.size() - c0);
      // coefficient blobs of the chunk's programs, back to back (same layout: same plan)
      std::vector<uint8_t> host(cnt * blob_stride, 0);
      parallel_for(cnt, [&](uint64_t y) {
        std::vector<Loc> ly;
        std::vector<uint8_t> b = pack_blob(pr[members[c0 + y]].ops, stages, ly);
        if (b.size() != blob_size) throw Error{QR_ERR_UNSUPPORTED, "internal error: blob layouts differ"};
        std::memcpy(host.data() + y * blob_stride, b.data(), b.size());
      });
      blobs.reserve(host.size());
      hip_check(hipMemcpyAsync(blobs.p, host.data(), host.size(), hipMemcpyHostToDevice, s), "H2D");
      states.reserve(cnt * dim * B);
      hip_check(launch_init_instances(prec, states.p, dim, cnt, s), "init states");
      const uint8_t* db = (const uint8_t*)blobs.p;
      for (size_t si = 0; si < stages.size(); ++si) {
        const Stage& st = stages[si];
        if (st.tile) {
          const TileArgs a = tile_args_at(n, st, db, loc[si]);
          const uint32_t ord = (a.ntiles & 7) == 0 ? 1u : 0u;  // XCD-contiguous tiles, as run_fast
          hip_check(jit_launch(cp->fns[si], st.k, pp.reg_bits, false, states.p, a.coefs, a.ntiles, a.nontile_mask, s,
                               ord, (uint32_t)cnt, dim, blob_stride / sizeof(double2)),
                    "compiled tile kernel (programs)");
        } else {
          for (uint64_t y = 0; y < cnt; ++y)
            launch_standalone_raw(prec, dim, (uint8_t*)states.p + y * dim * B, tmp,
                                  pr[members[c0 + y]].ops[st.standalone], db + y * blob_stride, loc[si], s);
        }
      }
      // the chunk's draws, each program's in ascending order
      std::vector<uint64_t> off(cnt + 1, 0), perm;
      std::vector<double> uu;
      for (uint64_t y = 0; y < cnt; ++y) {
        const uint64_t j = members[c0 + y];
        std::vector<uint64_t> idx(draws[j]);
        std::iota(idx.begin(), idx.end(), first[j]);
        std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return u[a] < u[b]; });
        for (uint64_t d : idx) {
          perm.push_back(d);
          uu.push_back(u[d]);
        }
        off[y + 1] = uu.size();
      }
      if (!uu.empty()) {
        dev_u.reserve(uu.size() * sizeof(double));
        dev_off.reserve(off.size() * sizeof(uint64_t));
        dev_out.reserve(uu.size() * sizeof(uint64_t));
        hip_check(hipMemcpyAsync(dev_u.p, uu.data(), uu.size() * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
        hip_check(hipMemcpyAsync(dev_off.p, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s),
                  "H2D");
        hip_check(launch_sample_fold_instances(prec, states.p, dim, cnt, (const double*)dev_u.p,
                                               (const uint64_t*)dev_off.p, (uint64_t*)dev_out.p, s),
                  "sample fold (programs)");
        std::vector<uint64_t> got(uu.size());
        hip_check(hipMemcpyAsync(got.data(), dev_out.p, got.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s),
                  "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        for (size_t q = 0; q < got.size(); ++q) out[perm[q]] = got[q];
      } else {
        hip_check(hipStreamSynchronize(s), "sync");
      }
    }
  }
}

}  // namespace qr
