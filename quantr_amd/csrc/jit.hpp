// jit.hpp — runtime-specialised tile sweeps (see jit.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "engine.hpp"

namespace qr {

// Kernel source of every tile stage (empty string for standalone stages).
std::vector<std::string> jit_stage_sources(const std::vector<Stage>& stages, int prec, int RB, bool persistent);
// Compiles (in parallel) and loads every source not yet in the process-wide cache. Throws Error.
void jit_prepare(const std::vector<std::string>& sources);
// Launches a prepared stage: one workgroup per tile.
hipError_t jit_launch(const std::string& source, int K, int RB, bool persistent, void* psi, const double2* coefs,
                      uint64_t ntiles, uint64_t nontile_mask, hipStream_t s);
size_t jit_cache_size();

}  // namespace qr
