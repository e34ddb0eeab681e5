// jit.hpp — runtime-specialised tile sweeps (see jit.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "engine.hpp"

namespace qr {

// Kernel source of every tile stage (empty string for standalone stages).
// With `dual`, stages carry psi and lam together and evaluate gradient taps (adjoint sweep).
std::vector<std::string> jit_stage_sources(const std::vector<Stage>& stages, int prec, int RB, bool persistent,
                                           bool dual = false);
// Compiles (in parallel) and loads every source not yet in the process-wide cache. Throws Error.
void jit_prepare(const std::vector<std::string>& sources);
// jit_prepare, then stages that spill registers under their occupancy hint are compiled again
// without it; `sources` then names the kernels to use.
void jit_prepare_without_spills(std::vector<std::string>& sources);
// The kernel of a prepared source. Throws Error if it was not prepared.
hipFunction_t jit_function(const std::string& source);
// Launches a prepared stage: one workgroup per tile. `order` 0 takes tiles in index order;
// 1 gives each XCD a contiguous eighth of them (non-persistent stages, tile counts divisible
// by 8; otherwise index order). Results do not depend on it. With `instances` > 1
// (non-persistent stages), instance y's state starts inst_stride amplitudes and its
// coefficients cf_stride double2 after instance 0's (grid y = instance, at most 65535).
hipError_t jit_launch(hipFunction_t fn, int K, int RB, bool persistent, void* psi, const double2* coefs,
                      uint64_t ntiles, uint64_t nontile_mask, hipStream_t s, uint32_t order = 0,
                      uint32_t instances = 1, uint64_t inst_stride = 0, uint64_t cf_stride = 0);
// Launches a prepared dual stage; gradient taps write part[param * nslots + slot].
hipError_t jit_launch_dual(hipFunction_t fn, int K, int RB, void* psi, void* lam, const double2* coefs,
                           uint64_t ntiles, uint64_t nontile_mask, double* part, uint64_t nslots, hipStream_t s);
uint64_t jit_dual_slots(int n, int K, int RB);
size_t jit_cache_size();
// The on-disk cache directory ("" when disabled or unusable), a 128-bit hex hash of some bytes,
// and an atomic file write (temporary name, then rename) for entries kept there.
std::string jit_cache_dir();
std::string cache_key_hex(const std::string& bytes);
bool cache_write(const std::string& path, const std::string& text);

}  // namespace qr
