// sampler.cpp — measure_all's sampling step: SuperPosition::measure (super_positions.rs:332-342)
// for a batch of injected draws, reproducing the reference's sequential left fold bit for bit.
//
// The reference draws u and returns the first index i with u < S_i, where
// S_i = fl(S_{i-1} + (re_i * re_i + im_i * im_i)) is folded in index order (no fused
// operations), or None when u >= S_last. The fold is rebuilt exactly in parallel:
//  1. block masses (2^14 amplitudes per block, any order) give approximate block starts and so
//     the binade e in which each block's fold most likely runs;
//  2. per block, exact integer sums of rint(p_i / 2^(e-52)) for e and its two neighbours;
//  3. on the host, the exact start of every block, in order: while a block's fold stays in the
//     binade of its start, fl(S + p) = S + rint(p / U) * U exactly (U = 2^(e-52)), so its end is
//     its start plus the integer sum. Blocks that cross into the next binade, hold a halfway
//     case, or start at 0 are folded sequentially on the host from their exact start;
//  4. each draw goes to the block whose exact fold range holds it; blocks of the integer kind
//     resolve their draws on the device from an exact in-block prefix sum, the others on the
//     host by the sequential fold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace qr {

namespace {

constexpr double kTwo53 = 9007199254740992.0;

// Fetches block b's amplitudes and folds them in index order from `s`, exactly as the
// reference does (binary64, no contraction). With draws, assigns each u[d] (sorted, d in
// [first, last)) the first index whose fold value exceeds it. Returns the fold's end value.
double host_fold(qr_state* st, uint64_t b, int bb, double s, const double* u, uint64_t first, uint64_t last,
                 uint64_t* out, std::vector<double>& buf) {
#pragma clang fp contract(off)
  const uint64_t len = 1ull << bb, b0 = b << bb;
  buf.resize(2 * len);
  if (st->prec == 0) {
    hip_check(hipMemcpy(buf.data(), (const double2*)st->amp + b0, len * 16, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  } else {
    std::vector<float> f(2 * len);
    hip_check(hipMemcpy(f.data(), (const float2*)st->amp + b0, len * 8, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (uint64_t i = 0; i < 2 * len; ++i) buf[i] = (double)f[i];
  }
  uint64_t d = first;
  for (uint64_t i = 0; i < len; ++i) {
    const double re = buf[2 * i], im = buf[2 * i + 1];
    const double p = re * re + im * im;
    s = s + p;
    while (d < last && u[d] < s) out[d++] = b0 + i;
  }
  return s;
}

struct BlockFold {
  double start = 0.0;
  double end = 0.0;
  int e = 0;         // binade of the integer fold
  bool exact = false;  // integer fold on the device (else sequential on the host)
};

}  // namespace

// us: draws in ascending order (no NaN); res: shots slots, UINT64_MAX = None.
void sample_exact_parallel(qr_state* st, const std::vector<double>& us, std::vector<uint64_t>& res) {
  const uint64_t shots = us.size();
  res.assign(shots, ~0ull);
  if (shots == 0) return;
  const int bb = std::min<int>(14, (int)st->n);
  const uint64_t nblk = st->dim >> bb;
  std::vector<double> mass(nblk);
  std::vector<double> sums(3 * nblk);
  std::vector<uint32_t> flags(nblk);
  std::vector<int32_t> ebase(nblk, -2000);
  hip_check(hipStreamSynchronize(st->stream), "sync");
  if (bb >= 12) {
    // 1. block masses
    st->aux2.reserve(nblk * 8);
    {
      ProfScope ps_(st, "block_mass", (double)st->dim * st->bytes_per_amp());
      hip_check(launch_block_mass(st->prec, st->amp, st->dim, bb, (double*)st->aux2.p, st->stream), "block mass");
    }
    hip_check(hipMemcpyAsync(mass.data(), st->aux2.p, nblk * 8, hipMemcpyDeviceToHost, st->stream), "D2H");
    hip_check(hipStreamSynchronize(st->stream), "sync");
    // 2. candidate binades from the approximate (compensated) prefix, then exact unit sums
    double acc = 0.0, comp = 0.0;
    for (uint64_t b = 0; b < nblk; ++b) {
      const double a = acc + comp;
      if (a > 0.0 && std::isfinite(a)) ebase[b] = std::ilogb(a) - 1;
      const double t = acc + mass[b];
      comp += std::fabs(acc) >= std::fabs(mass[b]) ? (acc - t) + mass[b] : (mass[b] - t) + acc;
      acc = t;
    }
    const size_t eb_bytes = nblk * sizeof(int32_t);
    std::vector<uint8_t> blob(eb_bytes);
    std::memcpy(blob.data(), ebase.data(), eb_bytes);
    const uint8_t* d_ebase = upload(st, st->aux, blob);
    st->aux2.reserve(nblk * (3 * 8 + 4));
    double* d_sums = (double*)st->aux2.p;
    uint32_t* d_flags = (uint32_t*)(d_sums + 3 * nblk);
    {
      ProfScope ps_(st, "block_exact_sums", (double)st->dim * st->bytes_per_amp());
      hip_check(launch_block_exact_sums(st->prec, st->amp, st->dim, bb, (const int32_t*)d_ebase, d_sums, d_flags,
                                        st->stream),
                "block exact sums");
    }
    hip_check(hipMemcpyAsync(sums.data(), d_sums, nblk * 24, hipMemcpyDeviceToHost, st->stream), "D2H");
    hip_check(hipMemcpyAsync(flags.data(), d_flags, nblk * 4, hipMemcpyDeviceToHost, st->stream), "D2H");
    hip_check(hipStreamSynchronize(st->stream), "sync");
  } else {
    for (uint64_t b = 0; b < nblk; ++b) mass[b] = 1.0;  // one small block: folded on the host
  }
  // 3. exact block starts, in order
  std::vector<BlockFold> bf(nblk);
  std::vector<double> ends(nblk);
  std::vector<double> buf;
  double s = 0.0;
  for (uint64_t b = 0; b < nblk; ++b) {
    BlockFold& f = bf[b];
    f.start = s;
    if (mass[b] == 0.0) {  // every p is +0: the fold does not move
      f.end = ends[b] = s;
      continue;
    }
    if (s > 0.0 && std::isfinite(s) && bb >= 12) {
      const int e = std::ilogb(s);
      const int c = e - ebase[b];
      if (c >= 0 && c <= 2 && !((flags[b] >> c) & 1u)) {
        const double unit = std::ldexp(1.0, e - 52);
        const double m = s / unit;  // integer in [2^52, 2^53), exact
        const double tot = sums[3 * b + c];
        if (std::isfinite(tot) && m + tot < kTwo53) {
          f.exact = true;
          f.e = e;
          s = (m + tot) * unit;
          f.end = ends[b] = s;
          continue;
        }
      }
    }
    s = host_fold(st, b, bb, s, nullptr, 0, 0, nullptr, buf);
    f.end = ends[b] = s;
  }
  // 4. draws to blocks: block b holds the draws with start_b <= u < end_b (the first block
  // whose end exceeds u); draws at or above the total are None
  std::vector<ExactJob> jobs;
  uint64_t d = 0;
  while (d < shots) {
    const uint64_t b = (uint64_t)(std::upper_bound(ends.begin(), ends.end(), us[d]) - ends.begin());
    if (b >= nblk) break;
    uint64_t e = d;
    while (e < shots && us[e] < ends[b]) ++e;
    const BlockFold& f = bf[b];
    if (f.exact) {
      const double unit = std::ldexp(1.0, f.e - 52);
      jobs.push_back(ExactJob{b, f.start / unit, unit, std::ldexp(1.0, 52 - f.e), d, e});
    } else {
      host_fold(st, b, bb, f.start, us.data(), d, e, res.data(), buf);
    }
    d = e;
  }
  if (jobs.empty()) return;
  const size_t ub = shots * 8, jb = jobs.size() * sizeof(ExactJob);
  std::vector<uint8_t> blob(ub + jb);
  std::memcpy(blob.data(), us.data(), ub);
  std::memcpy(blob.data() + ub, jobs.data(), jb);
  const uint8_t* dblob = upload(st, st->aux, blob);
  st->aux2.reserve(shots * 8);
  hip_check(hipMemcpyAsync(st->aux2.p, res.data(), shots * 8, hipMemcpyHostToDevice, st->stream), "H2D");
  {
    ProfScope ps_(st, "sample_exact_blocks", (double)jobs.size() * (double)(1ull << bb) * st->bytes_per_amp());
    hip_check(launch_sample_exact_blocks(st->prec, st->amp, st->dim, bb, (const ExactJob*)(dblob + ub), jobs.size(),
                                         (const double*)dblob, (uint64_t*)st->aux2.p, st->stream),
              "sample exact blocks");
  }
  hip_check(hipMemcpyAsync(res.data(), st->aux2.p, shots * 8, hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
}

}  // namespace qr
