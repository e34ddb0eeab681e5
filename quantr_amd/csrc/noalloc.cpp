// noalloc.cpp — qr_expectation_noalloc: <psi|O|psi> with no device or host allocation.
//
// Callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc) get the expectation value through the same passes and kernels as
// qr_expectation, with its device working set (the observable's X-mask groups, their tile
// masks, its Z-mask terms and the per-workgroup partial sums) in a fixed device array instead
// of workspaces that grow by hipMalloc. Host staging lives in fixed static arrays as well.
// Results are bitwise equal to qr_expectation's: same tables in the same order, same passes,
// same kernels, same partial count and the same host summation.
//
// The device array is one per GPU (each GPU holds its own copy until the process ends). Every
// call on a GPU shares it and it keeps the last call's contents; nothing in it needs to start
// at zero (the kernel writes every partial it later reads). Calls that overlapped on one GPU
// would race on it, so one host mutex serialises them (it also guards the host staging), and a
// call returns only after its kernel and copies have finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace qr {
namespace {

constexpr size_t kScratchBytes = 1u << 20;   // the device array: 1 MiB
constexpr uint64_t kMaxTerms = 16384;         // observable terms at most
constexpr int kMaxParts = 1024;               // partial sums (one per workgroup)
constexpr int kMaxDevices = 64;

// groups (16 B), their tile masks (4 B) and terms (24 B) of at most kMaxTerms terms, then the
// partials
static_assert(kMaxTerms * (sizeof(PauliGroup) + sizeof(uint32_t) + sizeof(PauliZTerm)) + 8 +
                      kMaxParts * sizeof(double) + 256 <=
                  kScratchBytes,
              "scratch layout");

}  // namespace
}  // namespace qr

static __device__ __align__(256) unsigned char scratch[qr::kScratchBytes];

namespace qr {
namespace {

std::mutex g_mu;  // serialises calls: host staging and per-device set-up
void* g_scratch_dev[kMaxDevices] = {};

// host staging (static: no allocation per call)
alignas(256) unsigned char g_stage[kScratchBytes];
struct KeyIdx {
  uint64_t xm;
  uint64_t idx;
};
KeyIdx g_keys[kMaxTerms];
PauliZTerm g_terms_in[kMaxTerms];
PauliPass g_passes[kMaxTerms + 1];
uint32_t g_lxm[kMaxTerms];
double g_parts[kMaxParts];

// The device address of this GPU's copy of `scratch`, looked up once per device with that
// device current. Caller holds g_mu.
void* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error{QR_ERR_UNSUPPORTED, "device index out of range"};
  if (!g_scratch_dev[device]) {
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device) hip_check(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    hip_check(e, "hipGetSymbolAddress (scratch)");
    if (!p) throw Error{QR_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array"};
    g_scratch_dev[device] = p;
  }
  return g_scratch_dev[device];
}

}  // namespace

double expectation_noalloc(qr_state* st, const qr_observable* o) {
  if (!o) throw Error{QR_ERR_ARG, "observable is NULL"};
  if (o->nterms && !o->terms) throw Error{QR_ERR_ARG, "observable terms are NULL"};
  if (o->nterms > kMaxTerms)
    throw Error{QR_ERR_ARG, "the observable has " + std::to_string(o->nterms) +
                                " terms; qr_expectation_noalloc takes at most " + std::to_string(kMaxTerms)};
  std::lock_guard<std::mutex> lk(g_mu);
  unsigned char* dev = (unsigned char*)scratch_on(st->device);
  const uint64_t T = o->nterms;
  if (T == 0) return 0.0;
  // terms as device terms, then grouped by X mask: ascending masks, input order inside a
  // group (the order qr_expectation uses)
  for (uint64_t t = 0; t < T; ++t) {
    const PauliMasks tm = parse_term(st->n, o, t);
    g_terms_in[t] = z_term(tm);
    g_keys[t] = KeyIdx{tm.xm, t};
  }
  std::sort(g_keys, g_keys + T,
            [](const KeyIdx& a, const KeyIdx& b) { return a.xm != b.xm ? a.xm < b.xm : a.idx < b.idx; });
  PauliGroup* groups = (PauliGroup*)g_stage;
  uint64_t G = 0;
  for (uint64_t t = 0; t < T; ++t)
    if (t == 0 || g_keys[t].xm != g_keys[t - 1].xm) groups[G++] = PauliGroup{g_keys[t].xm, (uint32_t)t, 0u};
  for (uint64_t g = 0; g < G; ++g) {
    const uint64_t end = g + 1 < G ? groups[g + 1].first : T;
    groups[g].count = (uint32_t)(end - groups[g].first);
  }
  // the same passes as qr_expectation (groups reordered in place)
  const int np = pauli_passes(st->n, st->prec, groups, (int)G, g_passes, g_lxm);
  const size_t gb = (size_t)G * sizeof(PauliGroup), lb = ((size_t)G * 4 + 7) & ~(size_t)7;
  std::memcpy(g_stage + gb, g_lxm, (size_t)G * 4);
  PauliZTerm* flat = (PauliZTerm*)(g_stage + gb + lb);
  for (uint64_t t = 0; t < T; ++t) flat[t] = g_terms_in[g_keys[t].idx];
  const size_t tb = (size_t)T * sizeof(PauliZTerm);
  const size_t part_off = (gb + lb + tb + 255) & ~(size_t)255;
  const int nb = pauli_expect_partials(st->dim);
  static_assert(kMaxParts >= 1024, "pauli_expect_partials returns at most 1024");
  hip_check(hipMemcpyAsync(dev, g_stage, gb + lb + tb, hipMemcpyHostToDevice, st->stream), "hipMemcpyAsync H2D");
  hip_check(launch_pauli_expect_passes(st->prec, st->amp, st->dim, (const PauliGroup*)dev,
                                       (const uint32_t*)(dev + gb), (const PauliZTerm*)(dev + gb + lb), g_passes,
                                       np, (double*)(dev + part_off), st->stream),
            "expectation");
  hip_check(hipMemcpyAsync(g_parts, dev + part_off, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream), "D2H");
  hip_check(hipStreamSynchronize(st->stream), "sync");
  return sum_partials_host(g_parts, (size_t)nb);
}

}  // namespace qr
