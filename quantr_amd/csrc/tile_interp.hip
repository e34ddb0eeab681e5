// tile_interp.hip — the interpreted fused tile sweep (k_tile): decodes a stage's op list at
// run time. Kept apart from kernels.hip because its many (precision, tile width, register
// width, prefetch) instantiations dominate compile time. Compiled stages (jit.cpp) are the
// default for large or repeated circuits; this kernel serves the first application of a
// structure below 28 qubits and every plan when run-time compilation is off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"
#include "tile_ops.hpp"

namespace qr {
namespace dev {
// =====================================================================================
// Fused tile sweep (interpreter). See planner.cpp for the program it runs.
// =====================================================================================

template <int K, int RB>
struct TileCtx {
  static constexpr int NT = K - RB;
  uint64_t g0;       // global index of register 0 of this thread
  uint64_t gb[RB];   // global bit of each register slot
  // Branch-free: every map word is loaded (uniformly) before any per-lane use.
  __device__ __forceinline__ void setup(const uint32_t* __restrict__ mp, uint64_t base, uint32_t tid) {
    uint32_t w[K];
#pragma unroll
    for (int s = 0; s < K; ++s) w[s] = mp[s];
    g0 = base;
#pragma unroll
    for (int s = 0; s < NT; ++s) g0 |= (uint64_t)((tid >> s) & 1u) << (w[s] & 0xFF);
#pragma unroll
    for (int j = 0; j < RB; ++j) gb[j] = 1ull << (w[NT + j] & 0xFF);
  }
  __device__ __forceinline__ uint64_t gidx(int m) const {
    uint64_t x = g0;
#pragma unroll
    for (int j = 0; j < RB; ++j)
      if ((m >> j) & 1) x |= gb[j];
    return x;
  }
};

// LDS offsets of a thread's registers from per-slot patterns P (XOR-linear in the slots).
template <int K, int RB>
__device__ __forceinline__ void lds_offsets(const uint32_t* __restrict__ P, uint32_t tid, uint32_t& t0,
                                            uint32_t (&rb)[RB]) {
  constexpr int NT = K - RB;
  uint32_t w[K];
#pragma unroll
  for (int s = 0; s < K; ++s) w[s] = P[s];
  t0 = 0;
#pragma unroll
  for (int s = 0; s < NT; ++s) t0 ^= w[s] & (0u - ((tid >> s) & 1u));
#pragma unroll
  for (int j = 0; j < RB; ++j) rb[j] = w[NT + j];
}

template <int RB>
__device__ __forceinline__ uint32_t lds_at(uint32_t t0, const uint32_t (&rb)[RB], int m) {
  uint32_t x = t0;
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if ((m >> j) & 1) x ^= rb[j];
  return x;
}

// An op and its four coefficients, loaded together (scalar loads).
struct OpPack {
  TileOpDev op;
  Coef4 cf;
};

__device__ __forceinline__ void load_op(const TileArgs& a, uint32_t o, OpPack& p) {
  p.op = a.ops[o];
#pragma unroll
  for (int q = 0; q < 4; ++q) p.cf.c[q] = a.coefs[4 * (size_t)o + q];
}

// Deposits the bits of the tile number into the non-tile index bits (uniform scalar work).
__device__ __forceinline__ uint64_t tile_base(uint64_t tile, const TileArgs& a) {
  uint64_t base = 0, m = a.nontile_mask;
  while (m) {
    const uint64_t low = m & (0 - m);
    if (tile & 1) base |= low;
    tile >>= 1;
    m ^= low;
  }
  return base;
}

// Persistent tile sweep: each workgroup walks tiles blockIdx.x, +gridDim.x, ...; the next
// tile's amplitudes are loaded into registers while the current tile is processed.
template <typename R, int K, int RB, bool PF>
__global__ __launch_bounds__(1 << (K - RB)) void k_tile(typename CT<R>::T* __restrict__ psi, TileArgs a) {
  using C = typename CT<R>::T;
  constexpr int NT = K - RB;
  constexpr int NR = 1 << RB;
  __shared__ C lds[1 << K];
  const uint32_t tid = threadIdx.x;
  uint64_t tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  TileCtx<K, RB> c0;  // load map, thread part only
  c0.setup(a.maps, 0, tid);
  uint64_t base = tile_base(tile, a);
  Regs<R, RB> cur, nxt;
#pragma unroll
  for (int m = 0; m < NR; ++m) cur.v[m] = psi[base | c0.gidx(m)];
  while (true) {
    const uint64_t ntile = tile + gridDim.x;
    const bool more = ntile < a.ntiles;
    uint64_t nbase = 0;
    if (PF && more) {
      nbase = tile_base(ntile, a);
#pragma unroll
      for (int m = 0; m < NR; ++m) nxt.v[m] = psi[nbase | c0.gidx(m)];
    }
    TileCtx<K, RB> ctx;
    ctx.g0 = base | c0.g0;
#pragma unroll
    for (int j = 0; j < RB; ++j) ctx.gb[j] = c0.gb[j];
    for (uint32_t o = 0; o < a.nops; ++o) {
      OpPack pk;
      load_op(a, o, pk);
      const TileOpDev& op = pk.op;
      const uint32_t code = op.w0 & 0xFF, A = (op.w0 >> 8) & 0xFF, Bv = (op.w0 >> 16) & 0xFF, fl = op.w0 >> 24;
      const bool lane_ok = (ctx.g0 & op.lane_cmask) == op.lane_cmask;  // used only if lane_cmask != 0
      const Coef4 cf = pk.cf;
      switch (code) {
        // Only uniform branches below: controls held by threads become per-lane selects of
        // the coefficients (a divergent branch would make the compiler copy every register).
        case TC_U1:
          if (op.lane_cmask == 0) {
            if (op.ns) u1_dispatch<R, RB, true>(cur, A, cf, Bv, op.ns);
            else u1_dispatch<R, RB, false>(cur, A, cf, Bv, 0);
          } else {  // per-lane coefficients: the gate or the identity
            Coef4 u;
            u.c[0] = lane_ok ? cf.c[0] : make_double2(1.0, 0.0);
            u.c[1] = lane_ok ? cf.c[1] : make_double2(0.0, 0.0);
            u.c[2] = lane_ok ? cf.c[2] : make_double2(0.0, 0.0);
            u.c[3] = lane_ok ? cf.c[3] : make_double2(1.0, 0.0);
            if (op.ns) u1_dispatch<R, RB, true>(cur, A, u, 0, op.ns);
            else u1_dispatch<R, RB, false>(cur, A, u, 0, 0);
          }
          break;
        case TC_D1: {
          if (Bv != 0xFF && op.lane_cmask == 0) {
            if (op.ns) d1_reg_dispatch<R, RB, true>(cur, Bv, cf, fl, op.ns);
            else d1_reg_dispatch<R, RB, false>(cur, Bv, cf, fl, 0);
          } else if (Bv != 0xFF) {
            Coef4 d;
            d.c[0] = lane_ok ? cf.c[0] : make_double2(1.0, 0.0);
            d.c[1] = lane_ok ? cf.c[1] : make_double2(1.0, 0.0);
            d.c[2] = d.c[3] = make_double2(0.0, 0.0);
            const uint32_t f2 = fl & 1;  // general factors, |1>-only skipping kept
            if (op.ns) d1_reg_dispatch<R, RB, true>(cur, Bv, d, f2, op.ns);
            else d1_reg_dispatch<R, RB, false>(cur, Bv, d, f2, 0);
          } else {
            const bool bit = (ctx.g0 >> A) & 1;
            double2 f = bit ? cf.c[1] : cf.c[0];
            if (op.lane_cmask) f = lane_ok ? f : make_double2(1.0, 0.0);
            if (op.ns) scale_all<R, RB, 0, true>(cur, (R)f.x, (R)f.y, op.ns);
            else scale_all<R, RB, 0, false>(cur, (R)f.x, (R)f.y, 0);
          }
          break;
        }
        case TC_REMAP: {
          const uint32_t* P = a.swz + (size_t)(2 * K) * op.swz;
          const uint32_t* nm = a.maps + (size_t)K * op.map;
          uint32_t t0, rb[RB];
          lds_offsets<K, RB>(P, tid, t0, rb);
          __syncthreads();
#pragma unroll
          for (int m = 0; m < NR; ++m) lds[lds_at<RB>(t0, rb, m)] = cur.v[m];
          __syncthreads();
          lds_offsets<K, RB>(P + K, tid, t0, rb);
#pragma unroll
          for (int m = 0; m < NR; ++m) cur.v[m] = lds[lds_at<RB>(t0, rb, m)];
          ctx.setup(nm, base, tid);
          break;
        }
        case TC_SWAPSLOTS:
          ctx.setup(a.maps + (size_t)K * op.map, base, tid);
          break;
        case TC_SCALE:
          scale_all<R, RB, 0, false>(cur, (R)cf.c[0].x, (R)cf.c[0].y, 0);
          break;
        default:
          break;
      }
    }
#pragma unroll
    for (int m = 0; m < NR; ++m) psi[ctx.gidx(m)] = cur.v[m];
    if (!PF || !more) break;
    tile = ntile;
    if (PF) {
#pragma unroll
      for (int m = 0; m < NR; ++m) cur.v[m] = nxt.v[m];
      base = nbase;
    } else {
      base = tile_base(tile, a);
#pragma unroll
      for (int m = 0; m < NR; ++m) cur.v[m] = psi[base | c0.gidx(m)];
    }
  }
}

}  // namespace dev
namespace {
// Workgroups that fit on the device at once for a tile kernel variant (persistent grid).
template <typename R, int K, int RB, bool PF>
unsigned tile_grid(uint64_t ntiles) {
  static unsigned cached = 0;
  if (!cached) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&dev::k_tile<R, K, RB, PF>),
                                                       1 << (K - RB), 0);
    if (cus <= 0) cus = 256;
    if (per_cu <= 0) per_cu = 1;
    cached = (unsigned)(cus * per_cu);
  }
  if (!PF) return (unsigned)ntiles;  // one tile per workgroup
  return (unsigned)std::min<uint64_t>(ntiles, cached);
}
}  // namespace

namespace {
template <typename R, int K, int RB>
hipError_t launch_tile_t(bool prefetch, void* psi, const TileArgs& a, hipStream_t s) {
  using C = typename dev::CT<R>::T;
  constexpr int threads = 1 << (K - RB);
  if (prefetch)
    dev::k_tile<R, K, RB, true><<<tile_grid<R, K, RB, true>(a.ntiles), threads, 0, s>>>((C*)psi, a);
  else
    dev::k_tile<R, K, RB, false><<<tile_grid<R, K, RB, false>(a.ntiles), threads, 0, s>>>((C*)psi, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_tile_r(int K, int RB, bool prefetch, void* psi, const TileArgs& a, hipStream_t s) {
  if (RB == 3 && K >= 10) {
    switch (K) {
      case 10: return launch_tile_t<R, 10, 3>(prefetch, psi, a, s);
      case 11: return launch_tile_t<R, 11, 3>(prefetch, psi, a, s);
      case 12: return launch_tile_t<R, 12, 3>(prefetch, psi, a, s);
      case 13: return launch_tile_t<R, 13, 3>(prefetch, psi, a, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (RB == 5) {
    switch (K) {
      case 10: return launch_tile_t<R, 10, 5>(prefetch, psi, a, s);
      case 11: return launch_tile_t<R, 11, 5>(prefetch, psi, a, s);
      case 12: return launch_tile_t<R, 12, 5>(prefetch, psi, a, s);
      case 13: return launch_tile_t<R, 13, 5>(prefetch, psi, a, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (RB != tile_reg_bits(K)) return hipErrorInvalidValue;
  switch (K) {
    case 1: return launch_tile_t<R, 1, 1>(prefetch, psi, a, s);
    case 2: return launch_tile_t<R, 2, 2>(prefetch, psi, a, s);
    case 3: return launch_tile_t<R, 3, 3>(prefetch, psi, a, s);
    case 4: return launch_tile_t<R, 4, 4>(prefetch, psi, a, s);
    case 5: return launch_tile_t<R, 5, 4>(prefetch, psi, a, s);
    case 6: return launch_tile_t<R, 6, 4>(prefetch, psi, a, s);
    case 7: return launch_tile_t<R, 7, 4>(prefetch, psi, a, s);
    case 8: return launch_tile_t<R, 8, 4>(prefetch, psi, a, s);
    case 9: return launch_tile_t<R, 9, 4>(prefetch, psi, a, s);
    case 10: return launch_tile_t<R, 10, 4>(prefetch, psi, a, s);
    case 11: return launch_tile_t<R, 11, 4>(prefetch, psi, a, s);
    case 12: return launch_tile_t<R, 12, 4>(prefetch, psi, a, s);
    case 13: return launch_tile_t<R, 13, 4>(prefetch, psi, a, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

hipError_t launch_tile(int prec, int K, int RB, bool prefetch, void* psi, const TileArgs& a, hipStream_t s) {
  if (a.ntiles == 0) return hipErrorInvalidValue;
  return prec == 0 ? launch_tile_r<double>(K, RB, prefetch, psi, a, s)
                   : launch_tile_r<float>(K, RB, prefetch, psi, a, s);
}
}  // namespace qr
