// jit.cpp — runtime-specialised tile sweeps.
//
// The interpreted tile sweep (k_tile in kernels.hip) decodes every op at run time; its
// dispatch switch costs the compiler register copies and branches on every op. For circuits
// that are applied repeatedly (benchmarks, gradient loops, serving) each stage's op list is
// instead turned into straight-line code: the same register-level gate templates
// (tile_ops.hpp, compiled from its text with hipRTC), with register slots, forms, control
// masks, slot maps and LDS patterns as compile-time constants. Coefficients stay run-time
// data, so a compiled stage serves every circuit of the same structure (e.g. all
// parameter-shifted variants). Modules are cached per process, keyed by their source text.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <fcntl.h>
#include <cstring>
#include <fstream>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "jit.hpp"
#include "kernels.hpp"

namespace qr {

extern const char* const kTileOpsSource;  // text of tile_ops.hpp, embedded at build time
}  // namespace qr

extern char** environ;

namespace qr {

namespace {

struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::map<std::string, JitKernel> g_cache;

std::string hex64(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
  return b;
}

// Source of one stage kernel. `prec` 0 = complex128, 1 = complex64. With `persistent`, each
// workgroup walks tiles blockIdx.x, +gridDim.x, ... and loads the next tile into a second
// register set while the current one is processed. With `dual` (the adjoint sweep) the
// kernel carries two states, psi and lam, through the same ops and evaluates gradient taps
// (TC_GRAD) on them; each tap stores per-wavefront partials into its parameter's row.
std::string stage_source(const Stage& s, int prec, int RB, bool persistent, bool dual) {
  const int K = s.k, NT = K - RB, NR = 1 << RB, TPB = 1 << NT;
  uint64_t tile_mask = 0;  // physical bits of the tile
  for (int b : s.pload) tile_mask |= 1ull << b;
  const char* R = prec == 0 ? "double" : "float";
  const char* C = prec == 0 ? "double2" : "float2";
  const int amp_bytes = prec == 0 ? 16 : 8;
  // both states through LDS at once when that fits in 64 KB, else one after the other
  const bool lds2 = dual && (2 * (1 << K) * amp_bytes <= 65536);
  // remaps in two halves (forward stages only): the LDS buffer holds half the tile
  bool split = false;
  if (!dual)
    for (const TileOpHost& op : s.ops) split = split || (op.code == TC_REMAP && (op.flags & 1));
  std::vector<std::string> sets = dual ? std::vector<std::string>{"r", "l"} : std::vector<std::string>{"r"};
  std::vector<std::string> arrays = dual ? std::vector<std::string>{"psi", "lam"} : std::vector<std::string>{"psi"};
  auto gbit = [&](int mi, int slot) { return s.phys_bit(mi, s.maps[(size_t)mi * K + slot]); };
  auto thread_part = [&](int mi) {
    std::ostringstream os;
    os << "base";
    for (int sl = 0; sl < NT; ++sl) os << " | ((uint64_t)((tid >> " << sl << ") & 1u) << " << gbit(mi, sl) << ")";
    return os.str();
  };
  auto reg_off = [&](int mi, int m) {
    uint64_t x = 0;
    for (int j = 0; j < RB; ++j)
      if ((m >> j) & 1) x |= 1ull << gbit(mi, NT + j);
    return x;
  };
  // Amplitudes move between memory and registers one component at a time: a whole-struct
  // copy can be lowered to a memcpy through a private stack slot (scratch memory) in stages
  // without gates, while the component copies still combine into one 16/8-byte access.
  // states larger than twice the 256 MB memory-side cache are streamed once per sweep and
  // never found in a cache by the next one: their amplitudes are loaded and stored
  // non-temporal (30-qubit brickwork 169.6 -> 163.7 ms); smaller states keep cached accesses,
  // which later sweeps reuse. QR_NT=0 or 1 forces either.
  bool nt = !dual && s.bytes >= 2.0 * (512.0 * 1024 * 1024);
  if (const char* e = std::getenv("QR_NT")) nt = e[0] == '1';
  auto load_amp = [&](const std::string& dst, const std::string& arr, const char* idx, uint64_t off) {
    std::ostringstream e;
    if (nt) {
      e << "qr::dev::load_nt(" << dst << ", &" << arr << "[" << idx << " | " << hex64(off) << "]);";
      return e.str();
    }
    e << "{ const " << C << "* s_ = &" << arr << "[" << idx << " | " << hex64(off) << "]; " << dst << ".x = s_->x; "
      << dst << ".y = s_->y; }";
    return e.str();
  };
  auto store_amp = [&](const std::string& src, const std::string& arr, const char* idx, uint64_t off) {
    std::ostringstream e;
    if (nt) {
      e << "qr::dev::store_nt(&" << arr << "[" << idx << " | " << hex64(off) << "], " << src << ");";
      return e.str();
    }
    e << "{ " << C << "* d_ = &" << arr << "[" << idx << " | " << hex64(off) << "]; d_->x = " << src << ".x; d_->y = "
      << src << ".y; }";
    return e.str();
  };
  std::ostringstream o;
  // occupancy target of the stage kernels: QR_STAGE_WAVES_PER_EU (1-8) asks the compiler to
  // fit that many wavefronts per SIMD (fewer registers per thread); unset leaves it free
  int wpe = persistent ? 2 : 0;
  // a 256-thread stage exchanging through half the LDS fits three workgroups per CU when it
  // uses at most 168 registers per lane: ask the compiler for three wavefronts per SIMD. For
  // states of 512 MB and more ask for four (128 registers) first: stages that fit run four
  // workgroups per CU, the others step down to three when compiled without spills (30-qubit
  // brickwork 157.3 -> 156.3 ms); smaller states skip the extra compilation.
  if (split && TPB == 256) wpe = s.bytes >= 2.0 * (512.0 * 1024 * 1024) ? 4 : 3;
  if (const char* e = std::getenv("QR_STAGE_WAVES_PER_EU")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 8) wpe = v;
  }
  o << "extern \"C\" __global__ __launch_bounds__(" << TPB << ")";
  if (wpe) o << " __attribute__((amdgpu_waves_per_eu(" << wpe << ", " << wpe << ")))";
  o << " void "
    << (dual ? "qr_dual_stage(" : "qr_stage(") << C
    << "* __restrict__ psi, ";
  if (dual) o << C << "* __restrict__ lam, ";
  o << "const double2* __restrict__ cf, uint64_t nontile, uint64_t ntiles";
  if (dual) o << ", double* __restrict__ part, uint64_t nslots";
  else o << ", uint32_t order, uint64_t inst_stride, uint64_t cf_stride";
  o << ") {\n";
  // several states of one structure in one launch: instance blockIdx.y has its amplitudes and
  // coefficients at fixed strides (zero for a single state)
  if (!dual) o << "  psi += (uint64_t)blockIdx.y * inst_stride;\n  cf += (uint64_t)blockIdx.y * cf_stride;\n";
  o << "  using R = " << R << ";\n";
  o << "  __shared__ " << C << " lds[" << (split ? (1 << (K - 1)) : (lds2 ? 2 : 1) * (1 << K)) << "];\n";
  o << "  const uint32_t tid = threadIdx.x;\n";
  o << "  uint64_t tile = blockIdx.x;\n";
  o << "  if (tile >= ntiles) return;\n";
  // order 1 deals tiles so that workgroup w takes tile (w mod 8) * ntiles/8 + w / 8: with
  // workgroups handed to the 8 XCDs in turn, each XCD walks one contiguous eighth of the tiles
  if (!dual && !persistent)
    o << "  if (order == 1u && (ntiles & 7ull) == 0) tile = (tile & 7ull) * (ntiles >> 3) + (tile >> 3);\n";
  o << "  uint64_t base = qr::dev::deposit_bits(tile, nontile);\n";
  for (const auto& v : sets) o << "  qr::dev::Regs<R, " << RB << "> " << v << ";\n";
  int mi = 0;
  const std::string thread0 = thread_part(0).substr(4);  // " | ..." part after "base"
  o << "  uint64_t g0 = " << thread_part(mi) << ";\n";
  for (size_t v = 0; v < sets.size(); ++v)
    for (int m = 0; m < NR; ++m)
      o << "  " << load_amp(sets[v] + ".v[" + std::to_string(m) + "]", arrays[v], "g0", reg_off(mi, m)) << "\n";
  if (persistent) {
    o << "  qr::dev::Regs<R, " << RB << "> nx;\n";
    o << "  while (true) {\n";
    o << "  const uint64_t ntile = tile + gridDim.x;\n";
    o << "  const bool more = ntile < ntiles;\n";
    o << "  uint64_t nbase = 0;\n";
    o << "  if (more) {\n    nbase = qr::dev::deposit_bits(ntile, nontile);\n";
    o << "    const uint64_t ng = nbase" << thread0 << ";\n";
    for (int m = 0; m < NR; ++m)
      o << "    " << load_amp("nx.v[" + std::to_string(m) + "]", "psi", "ng", reg_off(0, m)) << "\n";
    o << "  }\n";
  }
  auto coef = [&](size_t i) {
    std::ostringstream c;
    c << "qr::dev::Coef4{{cf[" << 4 * i << "], cf[" << 4 * i + 1 << "], cf[" << 4 * i + 2 << "], cf["
      << 4 * i + 3 << "]}}";
    return c.str();
  };
  for (size_t i = 0; i < s.ops.size(); ++i) {
    const TileOpHost& op = s.ops[i];
    const std::string ns = std::to_string(op.ns) + "u";
    const char* has_ns = op.ns ? "true" : "false";
    switch (op.code) {
      case TC_U1: {
        const uint64_t outside = op.lane_cmask & ~tile_mask, inside = op.lane_cmask & tile_mask;
        if (inside == 0) {
          // controls outside the tile are the same for the whole workgroup: a scalar branch
          if (outside) o << "  if ((base & " << hex64(outside) << ") == " << hex64(outside) << ")";
          o << "  { const qr::dev::Coef4 u = " << coef(i) << ";";
          for (const auto& v : sets)
            o << " qr::dev::u1_slot<R, " << RB << ", " << op.a << ", " << op.b << ", " << has_ns << ">(" << v
              << ", u, " << ns << ");";
          o << " }\n";
        } else if (op.b == TF_X) {  // controlled X: a per-lane exchange, no arithmetic
          o << "  { const bool ok = (g0 & " << hex64(op.lane_cmask) << ") == " << hex64(op.lane_cmask) << ";";
          for (const auto& v : sets)
            o << " qr::dev::u1_xswap<R, " << RB << ", " << op.a << ", " << has_ns << ">(" << v << ", ok, " << ns << ");";
          o << " }\n";
        } else {
          o << "  { const bool ok = (g0 & " << hex64(op.lane_cmask) << ") == " << hex64(op.lane_cmask) << ";\n"
            << "    qr::dev::Coef4 u = " << coef(i) << ";\n"
            << "    u.c[0] = ok ? u.c[0] : make_double2(1.0, 0.0); u.c[1] = ok ? u.c[1] : make_double2(0.0, 0.0);\n"
            << "    u.c[2] = ok ? u.c[2] : make_double2(0.0, 0.0); u.c[3] = ok ? u.c[3] : make_double2(1.0, 0.0);\n";
          for (const auto& v : sets)
            o << "    qr::dev::u1_slot<R, " << RB << ", " << op.a << ", 0, " << has_ns << ">(" << v << ", u, " << ns
              << ");\n";
          o << "  }\n";
        }
        break;
      }
      case TC_D1: {
        const bool only1 = op.flags & 1;
        const int form = (int)(op.flags >> 4);
        const uint64_t outside = op.lane_cmask & ~tile_mask, inside = op.lane_cmask & tile_mask;
        if (op.b != 0xFFu && inside == 0) {
          if (outside) o << "  if ((base & " << hex64(outside) << ") == " << hex64(outside) << ")";
          o << "  { const qr::dev::Coef4 d = " << coef(i) << ";";
          for (const auto& v : sets)
            o << " qr::dev::d1_reg<R, " << RB << ", " << op.b << ", " << (only1 ? "true" : "false") << ", "
              << (only1 ? form : 0) << ", " << has_ns << ">(" << v << ", d, " << ns << ");";
          o << " }\n";
        } else if (op.b != 0xFFu) {
          o << "  { const bool ok = (g0 & " << hex64(op.lane_cmask) << ") == " << hex64(op.lane_cmask) << ";\n"
            << "    qr::dev::Coef4 d = " << coef(i) << ";\n"
            << "    d.c[0] = ok ? d.c[0] : make_double2(1.0, 0.0); d.c[1] = ok ? d.c[1] : make_double2(1.0, 0.0);\n";
          for (const auto& v : sets)
            o << "    qr::dev::d1_reg<R, " << RB << ", " << op.b << ", " << (only1 ? "true" : "false") << ", 0, "
              << has_ns << ">(" << v << ", d, " << ns << ");\n";
          o << "  }\n";
        } else {
          o << "  { const bool bit = (g0 >> " << op.a << ") & 1;\n"
            << "    double2 f = bit ? cf[" << 4 * i + 1 << "] : cf[" << 4 * i << "];\n";
          if (op.lane_cmask)
            o << "    const bool ok = (g0 & " << hex64(op.lane_cmask) << ") == " << hex64(op.lane_cmask) << ";\n"
              << "    f = ok ? f : make_double2(1.0, 0.0);\n";
          for (const auto& v : sets)
            o << "    qr::dev::scale_all<R, " << RB << ", 0, " << has_ns << ">(" << v << ", (R)f.x, (R)f.y, " << ns
              << ");\n";
          o << "  }\n";
        }
        break;
      }
      case TC_GRAD: {
        if (!dual) break;
        const int gen = (int)op.flags;
        o << "  { double acc;\n";
        if (gen == 1 || gen == 2) {
          o << "    acc = qr::dev::tap_pair<R, " << RB << ", " << op.a << ", " << gen << ">(l, r);\n";
        } else if (gen == 3 && op.b != 0xFFu) {
          o << "    acc = qr::dev::tap_diag<R, " << RB << ", " << op.b << ">(l, r);\n";
        } else if (gen == 3) {
          o << "    acc = qr::dev::tap_diag<R, " << RB << ", -1>(l, r);\n"
            << "    if ((g0 >> " << op.a << ") & 1) acc = -acc;\n";
        } else {
          o << "    acc = qr::dev::tap_diag<R, " << RB << ", -1>(l, r);\n";
        }
        // the parameter's row comes from the op's coefficient slot, so that stages differing
        // only in parameter numbering share one kernel
        o << "    qr::dev::tap_store<" << TPB << ">(acc, part + (uint64_t)cf[" << 4 * i << "].x * nslots, tile, tid); }\n";
        break;
      }
      case TC_REMAP: {
        const uint16_t* P = &s.swz[(size_t)op.swz * 2 * K];
        auto lds_thread = [&](const uint16_t* Q) {
          std::ostringstream os;
          os << "0u";
          for (int sl = 0; sl < NT; ++sl) os << " ^ ((0u - ((tid >> " << sl << ") & 1u)) & " << Q[sl] << "u)";
          return os.str();
        };
        auto lds_reg = [&](const uint16_t* Q, int m) {
          uint32_t x = 0;
          for (int j = 0; j < RB; ++j)
            if ((m >> j) & 1) x ^= Q[NT + j];
          return x;
        };
        // Byte offset of register m's slot: (thread part) ^ (register part). The register part's
        // bits that no thread pattern can set (H) add, and fold into the access's immediate
        // offset; its other bits (E, from the swizzle) are XORed into one base per distinct E,
        // declared here, so an exchange takes a few XORs rather than one per access.
        auto lds_bases = [&](const char* name, const uint16_t* Q, uint32_t mask) {
          uint32_t T = 0;
          for (int sl = 0; sl < NT; ++sl) T |= Q[sl] & mask;
          std::vector<std::string> ex(NR);
          std::vector<uint32_t> seen;
          for (int m = 0; m < NR; ++m) {
            const uint32_t v = lds_reg(Q, m) & mask, E = v & T, H = v & ~T;
            std::string b = name;
            if (E) {
              b += "_" + std::to_string(E);
              if (std::find(seen.begin(), seen.end(), E) == seen.end()) {
                seen.push_back(E);
                o << "    const uint32_t " << b << " = " << name << " ^ " << E * amp_bytes << "u;\n";
              }
            }
            ex[m] = "(" + b + " + " + std::to_string(H * amp_bytes) + "u)";
          }
          return ex;
        };
        if (split && (op.flags & 1)) {
          // two rounds through half the LDS, split by tile position op.b
          const int h = (int)op.b, nmi = (int)op.a;
          int osl = -1, nsl = -1;
          for (int sl = 0; sl < K; ++sl) {
            if (s.maps[(size_t)mi * K + sl] == h) osl = sl;
            if (s.maps[(size_t)nmi * K + sl] == h) nsl = sl;
          }
          const uint32_t hm = (1u << (K - 1)) - 1;
          o << "  { const uint32_t w = ((" << lds_thread(P) << ") & " << hm << "u) * " << amp_bytes << "u;\n";
          o << "    const uint32_t q = ((" << lds_thread(P + K) << ") & " << hm << "u) * " << amp_bytes << "u;\n";
          const std::vector<std::string> wa = lds_bases("w", P, hm), qa = lds_bases("q", P + K, hm);
          auto in_round = [&](int slot, int m, int rho) {  // register m of this half (register slot)
            return ((m >> (slot - NT)) & 1) == rho;
          };
          std::vector<int> tmp;
          for (int m = 0; m < NR; ++m)
            if ((nsl < NT || in_round(nsl, m, 0)) && split_read_needs_tmp(osl, nsl, NT, m)) tmp.push_back(m);
          for (int m : tmp) o << "    " << C << " t" << m << ";\n";
          for (int rho = 0; rho < 2; ++rho) {
            o << "    __syncthreads();\n";
            if (osl < NT) o << "    if (((tid >> " << osl << ") & 1u) == " << rho << "u) {\n";
            for (int m = 0; m < NR; ++m)
              if (osl < NT || in_round(osl, m, rho))
                o << "    qr::dev::lds_at(lds, " << wa[m] << ") = r.v[" << m << "];\n";
            if (osl < NT) o << "    }\n";
            o << "    __syncthreads();\n";
            if (nsl < NT) o << "    if (((tid >> " << nsl << ") & 1u) == " << rho << "u) {\n";
            for (int m = 0; m < NR; ++m) {
              if (!(nsl < NT || in_round(nsl, m, rho))) continue;
              const bool t = rho == 0 && std::find(tmp.begin(), tmp.end(), m) != tmp.end();
              o << "    " << (t ? "t" + std::to_string(m) : "r.v[" + std::to_string(m) + "]")
                << " = qr::dev::lds_at(lds, " << qa[m] << ");\n";
            }
            if (nsl < NT) o << "    }\n";
          }
          if (!tmp.empty()) {
            if (nsl < NT) o << "    if (((tid >> " << nsl << ") & 1u) == 0u) {\n";
            for (int m : tmp) o << "    r.v[" << m << "] = t" << m << ";\n";
            if (nsl < NT) o << "    }\n";
          }
          mi = nmi;
          o << "    g0 = " << thread_part(mi) << "; }\n";
          break;
        }
        o << "  { const uint32_t w = (" << lds_thread(P) << ") * " << amp_bytes << "u;\n";
        o << "    const uint32_t q = (" << lds_thread(P + K) << ") * " << amp_bytes << "u;\n";
        const std::vector<std::string> wa = lds_bases("w", P, 0xFFFFFFFFu), qa = lds_bases("q", P + K, 0xFFFFFFFFu);
        if (lds2 || !dual) {
          o << "    __syncthreads();\n";
          for (size_t v = 0; v < sets.size(); ++v)
            for (int m = 0; m < NR; ++m)
              o << "    qr::dev::lds_at(lds, " << v * (1 << K) * amp_bytes << "u + " << wa[m] << ") = " << sets[v]
                << ".v[" << m << "];\n";
          o << "    __syncthreads();\n";
          for (size_t v = 0; v < sets.size(); ++v)
            for (int m = 0; m < NR; ++m)
              o << "    " << sets[v] << ".v[" << m << "] = qr::dev::lds_at(lds, " << v * (1 << K) * amp_bytes << "u + "
                << qa[m] << ");\n";
        } else {
          for (const auto& v : sets) {
            o << "    __syncthreads();\n";
            for (int m = 0; m < NR; ++m)
              o << "    qr::dev::lds_at(lds, " << wa[m] << ") = " << v << ".v[" << m << "];\n";
            o << "    __syncthreads();\n";
            for (int m = 0; m < NR; ++m)
              o << "    " << v << ".v[" << m << "] = qr::dev::lds_at(lds, " << qa[m] << ");\n";
          }
        }
        mi = (int)op.a;
        o << "    g0 = " << thread_part(mi) << "; }\n";
        break;
      }
      case TC_SWAPSLOTS:
        mi = (int)op.a;
        o << "  g0 = " << thread_part(mi) << ";\n";
        break;
      case TC_SCALE:
        o << "  { const double2 f = cf[" << 4 * i << "];";
        for (const auto& v : sets)
          o << " qr::dev::scale_all<R, " << RB << ", 0, false>(" << v << ", (R)f.x, (R)f.y, 0u);";
        o << " }\n";
        break;
      default:
        break;
    }
  }
  for (size_t v = 0; v < sets.size(); ++v)
    for (int m = 0; m < NR; ++m)
      o << "  " << store_amp(sets[v] + ".v[" + std::to_string(m) + "]", arrays[v], "g0", reg_off(mi, m)) << "\n";
  if (persistent) {
    o << "  if (!more) break;\n";
    o << "  tile = ntile;\n  base = nbase;\n";
    o << "  g0 = " << thread_part(0) << ";\n";
    for (int m = 0; m < NR; ++m) o << "  r.v[" << m << "].x = nx.v[" << m << "].x; r.v[" << m << "].y = nx.v[" << m << "].y;\n";
    o << "  }\n";
  }
  o << "}\n";
  return o.str();
}

std::string full_source(const std::string& kernel) {
  std::string src = "typedef unsigned int uint32_t;\ntypedef unsigned long long uint64_t;\n";
  src += kTileOpsSource;
  src += "\n";
  src += kernel;
  return src;
}

// Compiles one program; returns the code object or an error message.
bool compile_one(const std::string& src, std::vector<char>& code, std::string& err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "qr_stage.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 2000);
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  code.resize(sz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return true;
}

}  // namespace

std::vector<std::string> jit_stage_sources(const std::vector<Stage>& stages, int prec, int RB, bool persistent,
                                           bool dual) {
  std::vector<std::string> out(stages.size());
  for (size_t i = 0; i < stages.size(); ++i)
    if (stages[i].tile) out[i] = stage_source(stages[i], prec, RB, persistent && !dual, dual);
  return out;
}

namespace {

// ---- parallel compilation in child processes ----------------------------------------------
// hipRTC compiles one program at a time per process, so a plan of tens of stages would compile
// serially. Stages are instead split into groups, each written as one source with uniquely
// named kernels and compiled by the HIP compiler in its own child process (all groups at
// once); every group becomes one module. hipRTC in this process is the fallback.

std::string compiler_path() {
  const char* rp = std::getenv("ROCM_PATH");
  std::string p = std::string(rp && *rp ? rp : "/opt/rocm") + "/lib/llvm/bin/clang++";
  return access(p.c_str(), X_OK) == 0 ? p : std::string();
}

// Entry point of a stage source: qr_stage (forward) or qr_dual_stage (adjoint sweep).
std::string entry_name(const std::string& kernel) {
  return kernel.find("void qr_dual_stage(") != std::string::npos ? "qr_dual_stage" : "qr_stage";
}

// The source with its entry point renamed to <entry>_<j> (unique within a group; the prefix
// keeps profiles grouped by kind).
std::string renamed(const std::string& kernel, size_t j) {
  std::string k = kernel;
  const std::string e = entry_name(kernel);
  const size_t at = k.find("void " + e + "(");
  if (at != std::string::npos) k.replace(at, 6 + e.size(), "void " + e + "_" + std::to_string(j) + "(");
  return k;
}

bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

struct Group {
  std::vector<size_t> items;  // indices into the todo list
  std::string src, obj, log;
  pid_t pid = -1;
  bool ok = false;
};

// Compiles todo[i] for every i of every group; returns the groups that produced a code object.
void compile_in_children(const std::vector<std::string>& todo, std::vector<Group>& groups, const std::string& cc) {
  // a private directory under $TMPDIR, else /tmp
  auto make_dir = [](const std::string& parent, std::string& out) {
    std::string t = parent + "/qr_stages_XXXXXX";
    std::vector<char> buf(t.begin(), t.end());
    buf.push_back('\0');
    if (!mkdtemp(buf.data())) return false;
    out = buf.data();
    return true;
  };
  const char* td = std::getenv("TMPDIR");
  std::string d;
  if (!(td && *td && make_dir(td, d)) && !make_dir("/tmp", d)) return;
  for (size_t g = 0; g < groups.size(); ++g) {
    Group& G = groups[g];
    G.src = d + "/g" + std::to_string(g) + ".hip";
    G.obj = d + "/g" + std::to_string(g) + ".co";
    G.log = d + "/g" + std::to_string(g) + ".log";
    std::ofstream f(G.src);
    f << "#include <hip/hip_runtime.h>\n" << kTileOpsSource << "\n";
    for (size_t j = 0; j < G.items.size(); ++j) f << renamed(todo[G.items[j]], j);
    f.close();
    if (!f) continue;
    std::vector<std::string> args = {cc, "-x", "hip", "--offload-arch=gfx950", "--cuda-device-only",
                                     "--no-gpu-bundle-output", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                     "-Xclang", "-packed-fp32-ops", "-Wno-pragma-once-outside-header", "-w",
                                     "-o", G.obj, G.src};
    std::vector<char*> argv;
    for (auto& a : args) argv.push_back(&a[0]);
    argv.push_back(nullptr);
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_addopen(&fa, 1, G.log.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    posix_spawn_file_actions_adddup2(&fa, 1, 2);
    if (posix_spawn(&G.pid, cc.c_str(), &fa, nullptr, argv.data(), environ) != 0) G.pid = -1;
    posix_spawn_file_actions_destroy(&fa);
  }
  for (auto& G : groups) {
    if (G.pid <= 0) continue;
    int st = 0;
    while (waitpid(G.pid, &st, 0) < 0 && errno == EINTR) {
    }
    G.ok = WIFEXITED(st) && WEXITSTATUS(st) == 0;
  }
  for (auto& G : groups) {
    (void)unlink(G.src.c_str());
    (void)unlink(G.log.c_str());
  }
}

// ---- code objects on disk ---------------------------------------------------------------
// Compiled stages outlive the process: each group's code object is stored once in a cache
// directory, with one small index file per kernel (named by a hash of the kernel source and of
// what the compiler made of it: the tile-ops text, the options, the compiler binary's size and
// time) giving the object file and the kernel's name in it. Writes go to a temporary name
// and are renamed into place, so concurrent processes never read a partial file. Any cache
// error just means compiling. QR_JIT_CACHE=0 disables the cache, QR_JIT_CACHE=<dir> moves it;
// the default is $XDG_CACHE_HOME/quantr_amd or ~/.cache/quantr_amd.

struct Hash128 {
  uint64_t a = 0xcbf29ce484222325ull, b = 0x84222325cbf29ce4ull;
  void add(const std::string& s) {
    for (unsigned char ch : s) {
      a = (a ^ ch) * 0x100000001b3ull;
      b = (b ^ ch) * 0x9e3779b97f4a7c15ull;
      b ^= b >> 29;
    }
    a ^= s.size();
    b += s.size() * 0xbf58476d1ce4e5b9ull;
  }
  std::string hex() const {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%016llx%016llx", (unsigned long long)a, (unsigned long long)b);
    return buf;
  }
};

const char* const kCompileOptions = "-x hip --offload-arch=gfx950 --cuda-device-only -O3 -std=c++17 "
                                    "-target-feature -packed-fp32-ops";

std::string cache_dir() {
  static std::string dir = [] {
    const char* e = std::getenv("QR_JIT_CACHE");
    std::string base;
    if (e && *e) {
      if (std::string(e) == "0") return std::string();
      base = e;
    } else {
      const char* x = std::getenv("XDG_CACHE_HOME");
      const char* h = std::getenv("HOME");
      if (x && *x) base = std::string(x) + "/quantr_amd";
      else if (h && *h) base = std::string(h) + "/.cache/quantr_amd";
      else return std::string();
    }
    // one directory per toolchain identity
    Hash128 id;
    id.add(kTileOpsSource);
    id.add(kCompileOptions);
    const std::string cc = compiler_path();
    struct stat sb{};
    if (!cc.empty() && stat(cc.c_str(), &sb) == 0)
      id.add(cc + ":" + std::to_string((long long)sb.st_size) + ":" + std::to_string((long long)sb.st_mtime));
    std::string d = base + "/jit-" + id.hex().substr(0, 16);
    // create the path one level at a time
    for (size_t at = 1; at != std::string::npos;) {
      at = d.find('/', at + 1);
      const std::string part = d.substr(0, at);
      if (mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return std::string();
    }
    return access(d.c_str(), W_OK) == 0 ? d : std::string();
  }();
  return dir;
}

std::string kernel_key(const std::string& kernel) {
  Hash128 h;
  h.add(kernel);
  return h.hex();
}

bool write_file_atomic(const std::string& path, const char* data, size_t size) {
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return false;
    f.write(data, (std::streamsize)size);
    if (!f) {
      (void)unlink(tmp.c_str());
      return false;
    }
  }
  if (rename(tmp.c_str(), path.c_str()) != 0) {
    (void)unlink(tmp.c_str());
    return false;
  }
  return true;
}

std::map<std::string, hipModule_t> g_disk_modules;  // object file -> module (under g_mu)

// The kernel for `kernel` from the disk cache, if there. Caller holds no lock.
std::string content_hash(const std::vector<char>& code) {
  Hash128 h;
  h.add(std::string(code.begin(), code.end()));
  return h.hex();
}

// An index file names the code object, the kernel's symbol in it and the object's content
// hash; an object whose bytes do not hash to that (a damaged or foreign file) is never loaded.
bool load_cached(const std::string& dir, const std::string& kernel, JitKernel& out) {
  std::ifstream f(dir + "/" + kernel_key(kernel) + ".k");
  std::string obj, name, sum;
  if (!f || !std::getline(f, obj) || !std::getline(f, name) || !std::getline(f, sum) || obj.empty() ||
      name.empty() || sum.size() != 32)
    return false;
  if (obj.find('/') != std::string::npos) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  hipModule_t mod = nullptr;
  auto it = g_disk_modules.find(obj);
  if (it != g_disk_modules.end()) {
    mod = it->second;
  } else {
    std::vector<char> code;
    if (!read_file(dir + "/" + obj, code) || content_hash(code) != sum) return false;
    if (hipModuleLoadData(&mod, code.data()) != hipSuccess) return false;
    g_disk_modules[obj] = mod;
  }
  hipFunction_t fn = nullptr;
  if (hipModuleGetFunction(&fn, mod, name.c_str()) != hipSuccess) return false;
  out = JitKernel{mod, fn};
  return true;
}

// Stores a compiled group and indexes each of its kernels.
void store_group(const std::string& dir, const std::vector<std::string>& todo, const Group& G,
                 const std::vector<char>& code) {
  Hash128 gh;
  for (size_t j = 0; j < G.items.size(); ++j) gh.add(todo[G.items[j]]);
  const std::string obj = "g" + gh.hex() + ".co";
  // written even where the file exists, so that a damaged copy is replaced
  if (!write_file_atomic(dir + "/" + obj, code.data(), code.size())) return;
  const std::string sum = content_hash(code);
  for (size_t j = 0; j < G.items.size(); ++j) {
    const std::string& k = todo[G.items[j]];
    const std::string idx = obj + "\n" + entry_name(k) + "_" + std::to_string(j) + "\n" + sum + "\n";
    (void)write_file_atomic(dir + "/" + kernel_key(k) + ".k", idx.data(), idx.size());
  }
}

}  // namespace

void jit_prepare(const std::vector<std::string>& sources) {
  std::vector<std::string> todo;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& s : sources)
      if (!s.empty() && !g_cache.count(s) && std::find(todo.begin(), todo.end(), s) == todo.end()) todo.push_back(s);
  }
  if (todo.empty()) return;
  // kernels already on disk
  const std::string dir = cache_dir();
  size_t from_disk = 0;
  if (!dir.empty()) {
    std::vector<std::string> left;
    for (const auto& s : todo) {
      JitKernel k;
      if (load_cached(dir, s, k)) {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_cache.count(s)) g_cache[s] = k;
        ++from_disk;
      } else {
        left.push_back(s);
      }
    }
    todo.swap(left);
  }
  if (const char* e = std::getenv("QR_JIT_CACHE_LOG"))
    if (e[0] == '1')
      std::fprintf(stderr, "[qr jit] %zu kernels from the disk cache, %zu to compile\n", from_disk, todo.size());
  if (todo.empty()) return;
  std::vector<char> done(todo.size(), 0);
  std::vector<JitKernel> made(todo.size());
  const std::string cc = todo.size() > 1 ? compiler_path() : std::string();
  if (!cc.empty()) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t ng = std::min<size_t>(hw, todo.size());
    std::vector<Group> groups(ng);
    for (size_t i = 0; i < todo.size(); ++i) groups[i * ng / todo.size()].items.push_back(i);
    compile_in_children(todo, groups, cc);
    for (auto& G : groups) {
      std::vector<char> code;
      if (G.ok && read_file(G.obj, code)) {
        if (!dir.empty()) store_group(dir, todo, G, code);
        hipModule_t mod = nullptr;
        if (hipModuleLoadData(&mod, code.data()) == hipSuccess) {
          for (size_t j = 0; j < G.items.size(); ++j) {
            hipFunction_t fn = nullptr;
            const std::string nm = entry_name(todo[G.items[j]]) + "_" + std::to_string(j);
            if (hipModuleGetFunction(&fn, mod, nm.c_str()) == hipSuccess) {
              made[G.items[j]] = JitKernel{mod, fn};
              done[G.items[j]] = 1;
            }
          }
        }
      }
      (void)unlink(G.obj.c_str());
    }
    if (!groups.empty()) {
      const std::string dir = groups[0].obj.substr(0, groups[0].obj.rfind('/'));
      if (!dir.empty()) (void)rmdir(dir.c_str());
    }
  }
  // fallback: hipRTC in this process
  for (size_t i = 0; i < todo.size(); ++i) {
    if (done[i]) continue;
    std::vector<char> code;
    std::string err;
    if (!compile_one(full_source(todo[i]), code, err))
      throw Error{QR_ERR_HIP, "runtime compilation of a tile stage failed: " + err};
    if (!dir.empty()) {
      const std::string key = kernel_key(todo[i]), obj = "s" + key + ".co";
      if (write_file_atomic(dir + "/" + obj, code.data(), code.size())) {
        const std::string idx = obj + "\n" + entry_name(todo[i]) + "\n" + content_hash(code) + "\n";
        (void)write_file_atomic(dir + "/" + key + ".k", idx.data(), idx.size());
      }
    }
    JitKernel k;
    hipError_t e = hipModuleLoadData(&k.mod, code.data());
    if (e != hipSuccess) throw Error{QR_ERR_HIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e)};
    e = hipModuleGetFunction(&k.fn, k.mod, entry_name(todo[i]).c_str());
    if (e != hipSuccess) throw Error{QR_ERR_HIP, std::string("hipModuleGetFunction: ") + hipGetErrorString(e)};
    made[i] = k;
    done[i] = 1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (size_t i = 0; i < todo.size(); ++i)
    if (!g_cache.count(todo[i])) g_cache[todo[i]] = made[i];
}

// The occupancy hint asks the compiler to fit more wavefronts per SIMD; for a stage whose
// code then needs more registers than that allows, the compiler spills to scratch memory,
// which costs HBM traffic. Such stages are compiled again with a hint of three wavefronts
// (from a higher one), and then without a hint.
void jit_prepare_without_spills(std::vector<std::string>& sources) {
  jit_prepare(sources);
  // a stage that spills under a hint above three wavefronts per SIMD tries three, then none
  for (int round = 0; round < 2; ++round) {
    std::vector<size_t> redo;
    std::vector<std::string> next;
    for (size_t i = 0; i < sources.size(); ++i) {
      if (sources[i].empty()) continue;
      const std::string tag = " __attribute__((amdgpu_waves_per_eu(";
      const size_t at = sources[i].find(tag);
      if (at == std::string::npos) continue;
      int scratch = 0;
      if (hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, jit_function(sources[i])) != hipSuccess ||
          scratch <= 0)
        continue;
      const size_t end = sources[i].find(")))", at);
      if (end == std::string::npos) continue;
      const int w = std::atoi(sources[i].c_str() + at + tag.size());
      std::string p = sources[i];
      if (w > 3)
        p.replace(at, end + 3 - at, " __attribute__((amdgpu_waves_per_eu(3, 3)))");
      else
        p.erase(at, end + 3 - at);
      redo.push_back(i);
      next.push_back(std::move(p));
    }
    if (redo.empty()) return;
    jit_prepare(next);
    for (size_t j = 0; j < redo.size(); ++j) sources[redo[j]] = std::move(next[j]);
  }
}

std::string jit_cache_dir() { return cache_dir(); }

std::string cache_key_hex(const std::string& bytes) {
  Hash128 h;
  h.add(bytes);
  return h.hex();
}

bool cache_write(const std::string& path, const std::string& text) {
  return write_file_atomic(path, text.data(), text.size());
}

hipFunction_t jit_function(const std::string& source) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(source);
  if (it == g_cache.end()) throw Error{QR_ERR_HIP, "internal error: tile stage used before it was compiled"};
  return it->second.fn;
}

hipError_t jit_launch(hipFunction_t fn, int K, int RB, bool persistent, void* psi, const double2* coefs,
                      uint64_t ntiles, uint64_t nontile_mask, hipStream_t s, uint32_t order, uint32_t instances,
                      uint64_t inst_stride, uint64_t cf_stride) {
  if (!fn) return hipErrorInvalidValue;
  if (ntiles == 0 || ntiles > 0x7fffffffull) return hipErrorInvalidValue;
  if (instances == 0 || instances > 65535 || (instances > 1 && persistent)) return hipErrorInvalidValue;
  unsigned grid = (unsigned)ntiles;
  const unsigned threads = 1u << (K - RB);
  if (persistent) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess || per_cu <= 0)
      per_cu = 1;
    if (cus <= 0) cus = 256;
    grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)cus * per_cu);
  }
  if (persistent) order = 0;
  void* args[] = {&psi, (void*)&coefs, &nontile_mask, &ntiles, &order, &inst_stride, &cf_stride};
  return hipModuleLaunchKernel(fn, grid, instances, 1, threads, 1, 1, 0, s, args, nullptr);
}

hipError_t jit_launch_dual(hipFunction_t fn, int K, int RB, void* psi, void* lam, const double2* coefs,
                           uint64_t ntiles, uint64_t nontile_mask, double* part, uint64_t nslots, hipStream_t s) {
  if (!fn) return hipErrorInvalidValue;
  if (ntiles == 0 || ntiles > 0x7fffffffull) return hipErrorInvalidValue;
  const unsigned threads = 1u << (K - RB);
  void* args[] = {&psi, &lam, (void*)&coefs, &nontile_mask, &ntiles, &part, &nslots};
  return hipModuleLaunchKernel(fn, (unsigned)ntiles, 1, 1, threads, 1, 1, 0, s, args, nullptr);
}

// Partials written per tile by a dual stage: one per wavefront (or one per tile when a tile
// has fewer than 64 threads).
uint64_t jit_dual_slots(int n, int K, int RB) {
  const uint64_t tpb = 1ull << (K - RB);
  return (1ull << (n - K)) * (tpb < 64 ? 1 : tpb / 64);
}

size_t jit_cache_size() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_cache.size();
}

}  // namespace qr
