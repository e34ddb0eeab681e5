// kernels.hip — gfx950 kernels of the state-vector engine (wave64, CDNA4).
//
// State layout: 2^n amplitudes, interleaved (re, im), index i = sum_w bit_w * 2^(n-1-w)
// (wire 0 = most significant bit, product_states.rs:191-215). complex128 amplitudes are
// loaded as one 16-byte dwordx4, complex64 as one 8-byte dwordx2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"
#include "tile_ops.hpp"

namespace qr {
namespace dev {

// Inserts a zero bit at each position of `sorted_bits` (ascending) into x.
__device__ __forceinline__ uint64_t insert_zero_bits(uint64_t x, const uint8_t* sorted_bits, int nb) {
  for (int q = 0; q < nb; ++q) {
    const uint64_t b = sorted_bits[q];
    const uint64_t lo = x & ((1ull << b) - 1);
    x = ((x >> b) << (b + 1)) | lo;
  }
  return x;
}

// =====================================================================================
// |0...0> (SuperPosition::new_unchecked, super_positions_unchecked.rs:39-46)
// =====================================================================================
template <typename R>
__global__ void k_init_basis(typename CT<R>::T* __restrict__ psi, uint64_t dim, uint64_t one_at, int nt) {
  using C = typename CT<R>::T;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    C v;
    v.x = (i == one_at) ? R(1) : R(0);
    v.y = R(0);
    if (nt) store_nt(&psi[i], v);
    else psi[i] = v;
  }
}

// =====================================================================================
// Exact per-gate sweep: new[r] = sum over inputs j in ascending global index order of
// M[r][j] * a[j], the first product taken as is and the rest added, complex products in
// the textbook form with no fused multiply-add (Circuit::apply_gate, simulation.rs:64-180).
// Inputs whose custom image is None keep their amplitude (simulation.rs:120-133).
// One thread per group of 2^K amplitudes that differ only in the acting bits.
// =====================================================================================
template <typename R, int KR>
__global__ void k_exact(typename CT<R>::T* __restrict__ psi, uint64_t ngroups, ExactArgs a,
                        const double2* __restrict__ M) {
#pragma clang fp contract(off)
  using C = typename CT<R>::T;
  const int K = a.K;
  const int dim = 1 << K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zero_bits(g, a.sorted_bits, K);
    if (KR > 0) {
      C in[1 << (KR > 0 ? KR : 1)];
#pragma unroll
      for (int j = 0; j < (1 << KR); ++j)
        if (j < dim) in[j] = psi[base + a.off[j]];
      C out[1 << (KR > 0 ? KR : 1)];
#pragma unroll
      for (int r = 0; r < (1 << KR); ++r) {
        if (r >= dim) continue;
        if ((a.keep >> r) & 1) {
          out[r] = in[r];
          continue;
        }
        R acc_re = R(0), acc_im = R(0);
        bool first = true;
#pragma unroll
        for (int oi = 0; oi < (1 << KR); ++oi) {
          if (oi >= dim) continue;
          const int j = a.order[oi];
          if (!((a.some >> j) & 1)) continue;
          const double2 m = M[r * dim + j];
          const R mr = (R)m.x, mi = (R)m.y;
          C v;
#pragma unroll
          for (int jj = 0; jj < (1 << KR); ++jj)
            if (jj == j) v = in[jj];
          const R pr = mr * v.x - mi * v.y;
          const R pi = mr * v.y + mi * v.x;
          if (first) {
            acc_re = pr;
            acc_im = pi;
            first = false;
          } else {
            acc_re = acc_re + pr;
            acc_im = acc_im + pi;
          }
        }
        out[r].x = acc_re;
        out[r].y = acc_im;
      }
#pragma unroll
      for (int j = 0; j < (1 << KR); ++j)
        if (j < dim) psi[base + a.off[j]] = out[j];
    }
  }
}

// Generic exact sweep for wide custom gates (K > 4): inputs re-read from memory.
template <typename R>
__global__ void k_exact_wide(typename CT<R>::T* __restrict__ psi, typename CT<R>::T* __restrict__ tmp,
                             uint64_t ngroups, ExactWideArgs a, const double2* __restrict__ M,
                             const uint64_t* __restrict__ off, const uint32_t* __restrict__ order,
                             const uint8_t* __restrict__ flags) {
#pragma clang fp contract(off)
  using C = typename CT<R>::T;
  const int dim = 1 << a.K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // flags[j] bit0 = some, bit1 = keep
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zero_bits(g, a.sorted_bits, a.K);
    C* out = tmp + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * (uint64_t)dim;
    for (int r = 0; r < dim; ++r) {
      if (flags[r] & 2) {
        out[r] = psi[base + off[r]];
        continue;
      }
      R acc_re = R(0), acc_im = R(0);
      bool first = true;
      for (int oi = 0; oi < dim; ++oi) {
        const int j = (int)order[oi];
        if (!(flags[j] & 1)) continue;
        const double2 m = M[(uint64_t)r * dim + j];
        const C v = psi[base + off[j]];
        const R mr = (R)m.x, mi = (R)m.y;
        const R pr = mr * v.x - mi * v.y;
        const R pi = mr * v.y + mi * v.x;
        if (first) {
          acc_re = pr;
          acc_im = pi;
          first = false;
        } else {
          acc_re = acc_re + pr;
          acc_im = acc_im + pi;
        }
      }
      C o;
      o.x = acc_re;
      o.y = acc_im;
      out[r] = o;
    }
    for (int r = 0; r < dim; ++r) psi[base + off[r]] = out[r];
  }
}

// =====================================================================================
// Fast standalone sweeps (fused multiply-add allowed)
// =====================================================================================

template <typename R>
__device__ __forceinline__ void cmul_acc(R& xr, R& xi, R ur, R ui, R vr, R vi) {
  xr = fma(ur, vr, fma(-ui, vi, xr));
  xi = fma(ur, vi, fma(ui, vr, xi));
}

// 2x2 on target bit t for every pair whose control bits are all 1.
template <typename R>
__global__ void k_dense1(typename CT<R>::T* __restrict__ psi, uint64_t npairs, int t, uint64_t cmask,
                         double2 u00, double2 u01, double2 u10, double2 u11) {
  using C = typename CT<R>::T;
  const uint64_t tb = 1ull << t;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const R a_r = u00.x, a_i = u00.y, b_r = u01.x, b_i = u01.y, c_r = u10.x, c_i = u10.y, d_r = u11.x,
          d_i = u11.y;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
    const uint64_t i0 = ((p >> t) << (t + 1)) | (p & (tb - 1));
    if ((i0 & cmask) != cmask) continue;
    const uint64_t i1 = i0 | tb;
    const C x = psi[i0], y = psi[i1];
    R o0r = 0, o0i = 0, o1r = 0, o1i = 0;
    cmul_acc(o0r, o0i, a_r, a_i, (R)x.x, (R)x.y);
    cmul_acc(o0r, o0i, b_r, b_i, (R)y.x, (R)y.y);
    cmul_acc(o1r, o1i, c_r, c_i, (R)x.x, (R)x.y);
    cmul_acc(o1r, o1i, d_r, d_i, (R)y.x, (R)y.y);
    C r0, r1;
    r0.x = o0r;
    r0.y = o0i;
    r1.x = o1r;
    r1.y = o1i;
    psi[i0] = r0;
    psi[i1] = r1;
  }
}

// diag(d0, d1) on bit t where the control bits are all 1.
template <typename R>
__global__ void k_diag1(typename CT<R>::T* __restrict__ psi, uint64_t dim, int t, uint64_t cmask,
                        double2 d0, double2 d1) {
  using C = typename CT<R>::T;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    if ((i & cmask) != cmask) continue;
    const double2 d = ((i >> t) & 1) ? d1 : d0;
    const C x = psi[i];
    C y;
    y.x = (R)d.x * x.x - (R)d.y * x.y;
    y.y = (R)d.x * x.y + (R)d.y * x.x;
    psi[i] = y;
  }
}

// Dense 2^k x 2^k on k target bits (+ controls); one thread per group.
template <typename R>
__global__ void k_densek(typename CT<R>::T* __restrict__ psi, typename CT<R>::T* __restrict__ tmp,
                         uint64_t ngroups, DenseKArgs a, const double2* __restrict__ M,
                         const uint64_t* __restrict__ off) {
  using C = typename CT<R>::T;
  const int dim = 1 << a.k;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zero_bits(g, a.sorted_bits, a.k);
    if ((base & a.cmask) != a.cmask) continue;
    C* out = tmp + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * (uint64_t)dim;
    for (int r = 0; r < dim; ++r) {
      R ar = 0, ai = 0;
      for (int j = 0; j < dim; ++j) {
        const double2 m = M[(uint64_t)r * dim + j];
        const C v = psi[base + off[j]];
        cmul_acc(ar, ai, (R)m.x, (R)m.y, (R)v.x, (R)v.y);
      }
      C o;
      o.x = ar;
      o.y = ai;
      out[r] = o;
    }
    for (int r = 0; r < dim; ++r) psi[base + off[r]] = out[r];
  }
}

// Dense 2^K x 2^K on K <= 4 target bits (+ controls) with the group in registers: one thread
// per group of 2^K amplitudes, which it reads once, transforms and writes back in place.
// off[j] is the offset of local basis state j (MSB-first local index).
template <typename R, int K>
__global__ __launch_bounds__(256) void k_densek_reg(typename CT<R>::T* __restrict__ psi, uint64_t ngroups, DenseKArgs a,
                                                    const double2* __restrict__ M, const uint64_t* __restrict__ off) {
  using C = typename CT<R>::T;
  constexpr int D = 1 << K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zero_bits(g, a.sorted_bits, K);
    if ((base & a.cmask) != a.cmask) continue;
    C v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = psi[base + off[j]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      R ar = 0, ai = 0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const double2 m = M[r * D + j];
        cmul_acc(ar, ai, (R)m.x, (R)m.y, (R)v[j].x, (R)v[j].y);
      }
      C o;
      o.x = ar;
      o.y = ai;
      psi[base + off[r]] = o;
    }
  }
}

// 2^k diagonal phases over k bits (+ controls).
template <typename R>
__global__ void k_diagk(typename CT<R>::T* __restrict__ psi, uint64_t dim, DiagKArgs a,
                        const double2* __restrict__ D) {
  using C = typename CT<R>::T;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    if ((i & a.cmask) != a.cmask) continue;
    int l = 0;
    for (int q = 0; q < a.k; ++q) l = (l << 1) | (int)((i >> a.bits[q]) & 1);
    const double2 d = D[l];
    const C x = psi[i];
    C y;
    y.x = (R)d.x * x.x - (R)d.y * x.y;
    y.y = (R)d.x * x.y + (R)d.y * x.x;
    psi[i] = y;
  }
}

// =====================================================================================
// Sampling (SuperPosition::measure, super_positions.rs:332-342) with injected draws.
// =====================================================================================

// Exact left fold: one workgroup streams the probabilities through LDS; thread 0 adds them
// in index order (re*re + im*im, no fused operations) and assigns the sorted draws.
template <typename R>
__global__ __launch_bounds__(1024) void k_sample_fold(const typename CT<R>::T* __restrict__ psi,
                                                      uint64_t dim, const double* __restrict__ u,
                                                      uint64_t shots, uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int CH = 4096;
  __shared__ double p[CH];
  __shared__ int done;
  double cum = 0.0;
  uint64_t d = 0;
  double ud = shots ? u[0] : 0.0;
  if (threadIdx.x == 0) done = (shots == 0);
  __syncthreads();
  for (uint64_t base = 0; base < dim && !done; base += CH) {
    for (int q = threadIdx.x; q < CH; q += blockDim.x) {
      const uint64_t i = base + q;
      double v = 0.0;
      if (i < dim) {
        const double re = (double)psi[i].x, im = (double)psi[i].y;
        v = re * re + im * im;
      }
      p[q] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t cnt = (dim - base < (uint64_t)CH) ? (dim - base) : (uint64_t)CH;
      for (uint64_t q = 0; q < cnt; ++q) {
        cum = cum + p[q];
        while (d < shots && ud < cum) {
          out[d] = base + q;
          ++d;
          ud = (d < shots) ? u[d] : 0.0;
        }
        if (d >= shots) break;
      }
      done = (d >= shots);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (; d < shots; ++d) out[d] = ~0ull;
}

// Parallel path, step 1: probability mass of each block of 2^bb amplitudes (f64).
template <typename R>
__global__ __launch_bounds__(256) void k_block_mass(const typename CT<R>::T* __restrict__ psi,
                                                    uint64_t dim, int bb, double* __restrict__ mass, int nt) {
  __shared__ double red[256];
  const uint64_t b0 = (uint64_t)blockIdx.x << bb;
  const uint64_t len = 1ull << bb;
  double s = 0.0;
  for (uint64_t q = threadIdx.x; q < len; q += blockDim.x) {
    const uint64_t i = b0 + q;
    if (i < dim) {
      typename CT<R>::T v;
      if (nt) load_nt(v, &psi[i]);
      else v = psi[i];
      const double re = (double)v.x, im = (double)v.y;
      s += re * re + im * im;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) mass[blockIdx.x] = red[0];
}

// ---- Block-parallel exact fold ------------------------------------------------------------
// The reference's fold S_i = fl(S_{i-1} + p_i) is sequential, but inside one binade of S it is
// integer arithmetic: while 2^e <= S < 2^(e+1), every S is a multiple of U = 2^(e-52) and
// fl(S + p) = S + rint(p / U) * U whenever p / U is not exactly halfway between integers
// (round-to-nearest-even would then look at the parity of S). Sums of those integers are exact
// in binary64 while below 2^53, in any order. A block whose whole fold stays in one binade is
// therefore folded exactly by an integer prefix sum; blocks that cross a binade (a few per
// power of two of the cumulative mass) or hold a halfway case are folded sequentially instead.

// rint(p / U) for one amplitude, with p / U = p * scale exact (scale a power of two).
__device__ __forceinline__ double units_of(double p, double scale, bool& tie) {
  const double x = p * scale;
  const double r = rint(x);
  tie = fabs(x - r) == 0.5;
  return r;
}

// Per block b of 2^bb amplitudes: for the three candidate binades e = ebase[b] + c (c = 0, 1, 2)
// the exact sum of rint(p_i * 2^(52-e)) (sums[3b + c]; exact while below 2^53, and at least
// 2^53 exactly when the true sum is) and flags[b] bit c when some p_i * 2^(52-e) is a halfway
// case or e is outside [-960, 52] (then the scaling could overflow or the binade is unusable).
template <typename R>
__global__ __launch_bounds__(256) void k_block_exact_sums(const typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                          int bb, const int32_t* __restrict__ ebase,
                                                          double* __restrict__ sums, uint32_t* __restrict__ flags,
                                                          int nt) {
#pragma clang fp contract(off)
  __shared__ double red[3][256];
  __shared__ uint32_t fred[256];
  const uint64_t b0 = (uint64_t)blockIdx.x << bb;
  const uint64_t len = 1ull << bb;
  const int e0 = ebase[blockIdx.x];
  double sc[3];
  uint32_t fl = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int e = e0 + c;
    const bool ok = e >= -960 && e <= 52;
    sc[c] = ok ? ldexp(1.0, 52 - e) : 0.0;
    if (!ok) fl |= 1u << c;
  }
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  for (uint64_t q = threadIdx.x; q < len; q += blockDim.x) {
    const uint64_t i = b0 + q;
    if (i >= dim) break;
    typename CT<R>::T v;
    if (nt) load_nt(v, &psi[i]);
    else v = psi[i];
    const double re = (double)v.x, im = (double)v.y;
    const double p = re * re + im * im;
    bool t0, t1, t2;
    acc0 += units_of(p, sc[0], t0);
    acc1 += units_of(p, sc[1], t1);
    acc2 += units_of(p, sc[2], t2);
    fl |= (t0 ? 1u : 0u) | (t1 ? 2u : 0u) | (t2 ? 4u : 0u);
  }
  red[0][threadIdx.x] = acc0;
  red[1][threadIdx.x] = acc1;
  red[2][threadIdx.x] = acc2;
  fred[threadIdx.x] = fl;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
      fred[threadIdx.x] |= fred[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) sums[3 * (uint64_t)blockIdx.x + c] = red[c][0];
    flags[blockIdx.x] = fred[0];
  }
}

// One workgroup per job: the block's exact fold values S_i = (m_start + prefix_i) * U from an
// integer prefix sum (binary64, exact below 2^53), and every draw u of the job (sorted, all in
// [S_start, S_end)) gets the first index i with u < S_i. Chunks of 4096 amplitudes go through
// LDS; each thread owns 16 consecutive ones.
template <typename R>
__global__ __launch_bounds__(256) void k_sample_exact_blocks(const typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                             int bb, const ExactJob* __restrict__ jobs,
                                                             const double* __restrict__ u,
                                                             uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int CH = 4096, PER = 16;
  __shared__ double pv[CH + CH / PER];  // one pad word per 16 against bank conflicts
  __shared__ double scan[256];
  const ExactJob job = jobs[blockIdx.x];
  const uint64_t b0 = job.block << bb;
  const uint64_t len = 1ull << bb;
  const uint32_t t = threadIdx.x;
  double carry = job.m_start;  // integer count of units before the chunk
  for (uint64_t c0 = 0; c0 < len; c0 += CH) {
    for (int q = t; q < CH; q += 256) {
      const uint64_t i = b0 + c0 + q;
      double p = 0.0;
      if (c0 + q < len && i < dim) {
        const double re = (double)psi[i].x, im = (double)psi[i].y;
        p = re * re + im * im;
      }
      pv[q + q / PER] = p;
    }
    __syncthreads();
    double r[PER];
    double tsum = 0.0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      bool tie;
      r[j] = units_of(pv[t * (PER + 1) + j], job.scale, tie);
      tsum += r[j];
    }
    // inclusive scan of the thread sums (Hillis-Steele; integer-valued, exact)
    scan[t] = tsum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const double add = t >= (uint32_t)off ? scan[t - off] : 0.0;
      __syncthreads();
      scan[t] += add;
      __syncthreads();
    }
    double run = carry + (scan[t] - tsum);
    const double chunk_total = scan[255];
    // first draw of the job at or above this thread's starting fold value
    const double s_prev = run * job.unit;
    uint64_t lo = job.first, hi = job.last;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (u[mid] < s_prev) lo = mid + 1;
      else hi = mid;
    }
    uint64_t d = lo;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      run += r[j];
      const double s_i = run * job.unit;
      while (d < job.last && u[d] < s_i) {
        out[d] = b0 + c0 + (uint64_t)t * PER + j;
        ++d;
      }
    }
    carry += chunk_total;
    __syncthreads();
  }
}

// =====================================================================================
// Reductions (f64 accumulation; complex64 promoted before multiplying)
// =====================================================================================

template <typename R>
__global__ __launch_bounds__(256) void k_norm_partial(const typename CT<R>::T* __restrict__ psi,
                                                      uint64_t dim, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0, comp = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    const double re = (double)psi[i].x, im = (double)psi[i].y;
    const double v = re * re + im * im;
    const double y = v - comp;
    const double t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Gradient term of one rotation gate: sum_j conj(lam_j) (G psi)_j for G in {X, Y, Z, I} on bit t.
// Returns partial (re, im). Y: (Y psi)_j = -i (-1)^{bit_t(j)} psi_{j^t}; Z: (-1)^{bit_t(j)} psi_j.
template <typename R>
__global__ __launch_bounds__(256) void k_gen_partial(const typename CT<R>::T* __restrict__ lam,
                                                     const typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                     int t, int gen, double2* __restrict__ part) {
  __shared__ double rr[256], ri[256];
  double sr = 0.0, si = 0.0;
  const uint64_t tb = 1ull << t;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += stride) {
    double gr, gi;
    if (gen == 1) {
      gr = (double)psi[j ^ tb].x;
      gi = (double)psi[j ^ tb].y;
    } else if (gen == 2) {
      const double s = (j & tb) ? -1.0 : 1.0;
      const double pr = (double)psi[j ^ tb].x * s, pi = (double)psi[j ^ tb].y * s;
      gr = pi;   // -i * (pr + i pi) = pi - i pr
      gi = -pr;
    } else if (gen == 3) {
      const double s = (j & tb) ? -1.0 : 1.0;
      gr = (double)psi[j].x * s;
      gi = (double)psi[j].y * s;
    } else {
      gr = (double)psi[j].x;
      gi = (double)psi[j].y;
    }
    const double lr = (double)lam[j].x, li = -(double)lam[j].y;
    sr += lr * gr - li * gi;
    si += lr * gi + li * gr;
  }
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rr[threadIdx.x] += rr[threadIdx.x + w];
      ri[threadIdx.x] += ri[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(rr[0], ri[0]);
}

// lam = O psi for a whole Pauli sum in one pass: each output adds one product per X-mask
// group (groups and terms read with uniform loads).
template <typename R>
__global__ __launch_bounds__(256) void k_pauli_sum(typename CT<R>::T* __restrict__ lam,
                                                   const typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                   const PauliGroup* __restrict__ groups, int ngroups,
                                                   const PauliZTerm* __restrict__ terms) {
  using C = typename CT<R>::T;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < dim; o += stride) {
    double yr = 0.0, yi = 0.0;
    for (int g = 0; g < ngroups; ++g) {
      const uint64_t j = o ^ groups[g].xm;
      double fr = 0.0, fi = 0.0;
      const uint32_t t0 = groups[g].first, t1 = t0 + groups[g].count;
      for (uint32_t t = t0; t < t1; ++t) {
        const bool neg = __popcll(j & terms[t].zm) & 1;
        fr += neg ? -terms[t].cr : terms[t].cr;
        fi += neg ? -terms[t].ci : terms[t].ci;
      }
      const double pr = (double)psi[j].x, pi = (double)psi[j].y;
      yr += fr * pr - fi * pi;
      yi += fr * pi + fi * pr;
    }
    C w;
    w.x = (R)yr;
    w.y = (R)yi;
    lam[o] = w;
  }
}

// |0...0> in each of `count` states of dim amplitudes laid out back to back.
template <typename R>
__global__ __launch_bounds__(256) void k_init_instances(typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                        uint64_t total) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    psi[i].x = (R)((i & (dim - 1)) == 0 ? 1 : 0);
    psi[i].y = (R)0;
  }
}

// One workgroup per state (states back to back, `stride` amplitudes apart): its draws
// u[off[y] .. off[y+1]) (ascending) placed by the left fold of its probabilities in index
// order, as k_sample_fold does for one state: probabilities in parallel, the fold in one lane.
template <typename R>
__global__ __launch_bounds__(256) void k_sample_fold_instances(const typename CT<R>::T* __restrict__ psi,
                                                               uint64_t dim, uint64_t stride,
                                                               const double* __restrict__ u,
                                                               const uint64_t* __restrict__ off,
                                                               uint64_t* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int CH = 2048;
  __shared__ double p[CH];
  __shared__ int done;
  const typename CT<R>::T* ps = psi + (uint64_t)blockIdx.x * stride;
  const uint64_t d1 = off[blockIdx.x + 1];
  uint64_t d = off[blockIdx.x];
  double cum = 0.0;
  if (threadIdx.x == 0) done = (d >= d1);
  __syncthreads();
  for (uint64_t base = 0; base < dim && !done; base += CH) {
    for (int q = threadIdx.x; q < CH; q += blockDim.x) {
      const uint64_t i = base + q;
      double v = 0.0;
      if (i < dim) {
        const double re = (double)ps[i].x, im = (double)ps[i].y;
        v = re * re + im * im;
      }
      p[q] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t cnt = (dim - base < (uint64_t)CH) ? (dim - base) : (uint64_t)CH;
      for (uint64_t q = 0; q < cnt; ++q) {
        cum = cum + p[q];
        while (d < d1 && u[d] < cum) out[d++] = base + q;
      }
      done = (d >= d1);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    while (d < d1) out[d++] = ~0ull;
}

// Partial sums of Re <psi|O|psi> = Re sum_o conj(psi_o) (O psi)_o in one pass, one per workgroup.
template <typename R>
__global__ __launch_bounds__(256) void k_pauli_expect(const typename CT<R>::T* __restrict__ psi, uint64_t dim,
                                                      const PauliGroup* __restrict__ groups, int ngroups,
                                                      const PauliZTerm* __restrict__ terms, double* __restrict__ part,
                                                      int accumulate) {
  __shared__ double red[256];
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < dim; o += stride) {
    double yr = 0.0, yi = 0.0;
    for (int g = 0; g < ngroups; ++g) {
      const uint64_t j = o ^ groups[g].xm;
      double fr = 0.0, fi = 0.0;
      const uint32_t t0 = groups[g].first, t1 = t0 + groups[g].count;
      for (uint32_t t = t0; t < t1; ++t) {
        const bool neg = __popcll(j & terms[t].zm) & 1;
        fr += neg ? -terms[t].cr : terms[t].cr;
        fi += neg ? -terms[t].ci : terms[t].ci;
      }
      const double pr = (double)psi[j].x, pi = (double)psi[j].y;
      yr += fr * pr - fi * pi;
      yi += fr * pi + fi * pr;
    }
    acc += (double)psi[o].x * yr + (double)psi[o].y * yi;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = accumulate ? part[blockIdx.x] + red[0] : red[0];
}

// f[r] += (-1)^{popcount(r & hz)} a for the positions r of a chunk: one case per pattern hz,
// so the signs are compile-time constants (adds and subtracts, no selects).
template <int CH, int HZ>
__device__ __forceinline__ void add_signed_fixed(double (&f)[CH], double a) {
#pragma unroll
  for (int r = 0; r < CH; ++r) {
    if (__builtin_popcount((unsigned)(r & HZ)) & 1) f[r] -= a;
    else f[r] += a;
  }
}

template <int CH>
__device__ __forceinline__ void add_signed(double (&f)[CH], double a, uint32_t hz) {
  static_assert(CH == 16, "chunks of 16 positions");
  switch (hz) {
    case 0: add_signed_fixed<CH, 0>(f, a); break;
    case 1: add_signed_fixed<CH, 1>(f, a); break;
    case 2: add_signed_fixed<CH, 2>(f, a); break;
    case 3: add_signed_fixed<CH, 3>(f, a); break;
    case 4: add_signed_fixed<CH, 4>(f, a); break;
    case 5: add_signed_fixed<CH, 5>(f, a); break;
    case 6: add_signed_fixed<CH, 6>(f, a); break;
    case 7: add_signed_fixed<CH, 7>(f, a); break;
    case 8: add_signed_fixed<CH, 8>(f, a); break;
    case 9: add_signed_fixed<CH, 9>(f, a); break;
    case 10: add_signed_fixed<CH, 10>(f, a); break;
    case 11: add_signed_fixed<CH, 11>(f, a); break;
    case 12: add_signed_fixed<CH, 12>(f, a); break;
    case 13: add_signed_fixed<CH, 13>(f, a); break;
    case 14: add_signed_fixed<CH, 14>(f, a); break;
    default: add_signed_fixed<CH, 15>(f, a); break;
  }
}

// The same partial sums for groups whose X masks all lie in one tile: tile bits = the set
// bits of tmask (K of them, the lowest first). Each workgroup walks tiles blockIdx.x,
// +gridDim.x, ...: it loads the tile into LDS once and takes every partner amplitude
// psi[o ^ xm] from there (lxm[g] = xm of group g in tile coordinates), so the whole pass reads
// the state once however many groups it holds.
template <typename R, int K>
__global__ __launch_bounds__(256) void k_pauli_expect_tiles(const typename CT<R>::T* __restrict__ psi,
                                                            uint64_t ntiles, uint64_t tmask, uint64_t nontile,
                                                            const PauliGroup* __restrict__ groups,
                                                            const uint32_t* __restrict__ lxm, int ngroups,
                                                            const PauliZTerm* __restrict__ terms,
                                                            double* __restrict__ part, int accumulate) {
  using T = typename CT<R>::T;
  constexpr int NPT = (1 << K) / 256;  // amplitudes per thread
  __shared__ T lds[1 << K];
  __shared__ double red[256];
  const uint32_t tid = threadIdx.x;
  // global offset of tile position p = tid + 256 r: the thread's part (tile bits 0..7) and the
  // register part (tile bits 8..K-1)
  uint64_t toff = 0, hib[K - 8];
  {
    uint64_t m = tmask;
    for (int b = 0; b < K; ++b) {
      const uint64_t low = m & (0ull - m);
      m ^= low;
      if (b < 8) {
        if ((tid >> b) & 1u) toff |= low;
      } else {
        hib[b - 8] = low;
      }
    }
  }
  double acc = 0.0;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = deposit_bits(t, nontile) | toff;
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      uint64_t off = 0;
#pragma unroll
      for (int q = 0; q < K - 8; ++q)
        if ((r >> q) & 1) off |= hib[q];
      const T* src = &psi[base | off];
      T v;
      v.x = src->x;
      v.y = src->y;
      lds[tid + 256 * r] = v;
    }
    __syncthreads();
    // per group: the coefficient of every position (terms outer, the thread's positions inner),
    // then the partner amplitudes from LDS. The sign (-1)^{popcount(j & zm)}, j = o ^ xm, splits
    // into the thread's part (base ^ xm) and the register part of o: r's bits on the tile bits
    // 8.. that zm holds (hz), a constant per position.
    constexpr int CH = 16;  // positions per chunk (register budget)
#pragma unroll 1
    for (int c0 = 0; c0 < NPT; c0 += CH) {
      double yr[CH], yi[CH];
#pragma unroll
      for (int r = 0; r < CH; ++r) yr[r] = yi[r] = 0.0;
      for (int g = 0; g < ngroups; ++g) {
        const uint64_t bx = base ^ groups[g].xm;
        const uint32_t lx = lxm[g];
        const uint32_t t0 = groups[g].first, t1 = t0 + groups[g].count;
        double fr[CH], fi[CH];
#pragma unroll
        for (int r = 0; r < CH; ++r) fr[r] = fi[r] = 0.0;
        for (uint32_t q = t0; q < t1; ++q) {
          const uint64_t zm = terms[q].zm;
          const double cr = terms[q].cr, ci = terms[q].ci;
          uint32_t hz = 0;
#pragma unroll
          for (int b = 0; b < K - 8; ++b)
            if (zm & hib[b]) hz |= 1u << b;
          // the thread's signed coefficient (its part of the sign, and the chunk's), then each
          // position adds or subtracts it by the parity of its register bits in hz
          const uint32_t bt = ((uint32_t)__popcll(bx & zm) ^ (uint32_t)__popc((uint32_t)c0 & hz)) & 1u;
          const double ar = bt ? -cr : cr;
          add_signed<CH>(fr, ar, hz & (CH - 1));
          if (ci != 0.0) add_signed<CH>(fi, bt ? -ci : ci, hz & (CH - 1));
        }
#pragma unroll
        for (int r = 0; r < CH; ++r) {
          const T w = lds[(tid + 256u * (uint32_t)(c0 + r)) ^ lx];
          const double pr = (double)w.x, pi = (double)w.y;
          yr[r] += fr[r] * pr - fi[r] * pi;
          yi[r] += fr[r] * pi + fi[r] * pr;
        }
      }
#pragma unroll
      for (int r = 0; r < CH; ++r) {
        const T v = lds[tid + 256u * (uint32_t)(c0 + r)];
        acc += (double)v.x * yr[r] + (double)v.y * yi[r];
      }
    }
    __syncthreads();  // the next tile overwrites lds
  }
  red[tid] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = accumulate ? part[blockIdx.x] + red[0] : red[0];
}

// Partial sums of Re sum_j conj(a_j) b_j, one per workgroup.
template <typename R>
__global__ __launch_bounds__(256) void k_dot_re_partial(const typename CT<R>::T* __restrict__ a,
                                                        const typename CT<R>::T* __restrict__ b, uint64_t dim,
                                                        double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += stride)
    acc += (double)a[j].x * (double)b[j].x + (double)a[j].y * (double)b[j].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The streaming ceiling of a sweep: every amplitude read, multiplied by a real factor and
// written back in place, one amplitude per thread in index order (the fastest in-place shape
// measured on MI355X).
template <typename R>
__global__ __launch_bounds__(256) void k_stream_scale(typename CT<R>::T* __restrict__ psi, uint64_t dim, R f) {
  // non-temporal, as the tile sweeps of large states access memory
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dim) {
    typename CT<R>::T v;
    load_nt(v, &psi[i]);
    v.x *= f;
    v.y *= f;
    store_nt(&psi[i], v);
  }
}

__global__ void k_stamp(uint64_t* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = wall_clock64();
}

// Row sums of a partial-sum table in a fixed order (one workgroup per row): each thread adds
// a strided subset, then a tree over the workgroup.
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ part, uint64_t ncols,
                                                  double* __restrict__ out) {
  __shared__ double red[256];
  const double* row = part + (uint64_t)blockIdx.x * ncols;
  double acc = 0.0;
  for (uint64_t j = threadIdx.x; j < ncols; j += 256) acc += row[j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// complex64 <-> binary64 conversion for state upload/download
__global__ void k_f64_to_c64(float2* __restrict__ dst, const double2* __restrict__ src, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}
__global__ void k_c64_to_f64(double2* __restrict__ dst, const float2* __restrict__ src, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = make_double2((double)src[i].x, (double)src[i].y);
}

}  // namespace dev

// =====================================================================================
// Host launchers
// =====================================================================================

namespace {
inline unsigned grid_for(uint64_t work, int block, unsigned cap = 1u << 20) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}
}  // namespace

int tile_max_bits(int prec) { return 13; }
int tile_reg_bits(int K) { return K < 4 ? K : 4; }

// States of at least 512 MB (twice the memory-side cache) are streamed with non-temporal
// accesses: no later pass finds them in a cache anyway.
inline int streamed(int prec, uint64_t dim) { return dim * (prec == 0 ? 16u : 8u) >= (512ull << 20) ? 1 : 0; }

hipError_t launch_init_basis(int prec, void* psi, uint64_t dim, uint64_t one_at, hipStream_t s) {
  const unsigned g = grid_for(dim, 256, 1u << 16);
  const int nt = streamed(prec, dim);
  if (prec == 0)
    dev::k_init_basis<double><<<g, 256, 0, s>>>((double2*)psi, dim, one_at, nt);
  else
    dev::k_init_basis<float><<<g, 256, 0, s>>>((float2*)psi, dim, one_at, nt);
  return hipGetLastError();
}

hipError_t launch_exact(int prec, void* psi, uint64_t ngroups, const ExactArgs& a, const double2* M,
                        hipStream_t s) {
  const unsigned g = grid_for(ngroups, 256, 1u << 16);
#define QR_EXACT_CASE(KR)                                                                   \
  case KR:                                                                                  \
    if (prec == 0)                                                                          \
      dev::k_exact<double, KR><<<g, 256, 0, s>>>((double2*)psi, ngroups, a, M);             \
    else                                                                                    \
      dev::k_exact<float, KR><<<g, 256, 0, s>>>((float2*)psi, ngroups, a, M);               \
    break;
  switch (a.K) {
    QR_EXACT_CASE(1)
    QR_EXACT_CASE(2)
    QR_EXACT_CASE(3)
    QR_EXACT_CASE(4)
    default:
      return hipErrorInvalidValue;
  }
#undef QR_EXACT_CASE
  return hipGetLastError();
}

hipError_t launch_exact_wide(int prec, void* psi, void* tmp, uint64_t tmp_groups, uint64_t ngroups,
                             const ExactWideArgs& a, const double2* M, const uint64_t* off,
                             const uint32_t* order, const uint8_t* flags, hipStream_t s) {
  // one thread per tmp slot: grid * block == tmp_groups
  const int block = tmp_groups >= 256 ? 256 : (int)tmp_groups;
  const unsigned g = (unsigned)(tmp_groups / block);
  if (prec == 0)
    dev::k_exact_wide<double><<<g, block, 0, s>>>((double2*)psi, (double2*)tmp, ngroups, a, M, off, order, flags);
  else
    dev::k_exact_wide<float><<<g, block, 0, s>>>((float2*)psi, (float2*)tmp, ngroups, a, M, off, order, flags);
  return hipGetLastError();
}

hipError_t launch_dense1(int prec, void* psi, uint64_t npairs, int t, uint64_t cmask, const double2* u,
                         hipStream_t s) {
  const unsigned g = grid_for(npairs, 256, 1u << 18);
  if (prec == 0)
    dev::k_dense1<double><<<g, 256, 0, s>>>((double2*)psi, npairs, t, cmask, u[0], u[1], u[2], u[3]);
  else
    dev::k_dense1<float><<<g, 256, 0, s>>>((float2*)psi, npairs, t, cmask, u[0], u[1], u[2], u[3]);
  return hipGetLastError();
}

hipError_t launch_diag1(int prec, void* psi, uint64_t dim, int t, uint64_t cmask, double2 d0, double2 d1,
                        hipStream_t s) {
  const unsigned g = grid_for(dim, 256, 1u << 18);
  if (prec == 0)
    dev::k_diag1<double><<<g, 256, 0, s>>>((double2*)psi, dim, t, cmask, d0, d1);
  else
    dev::k_diag1<float><<<g, 256, 0, s>>>((float2*)psi, dim, t, cmask, d0, d1);
  return hipGetLastError();
}

hipError_t launch_densek_reg(int prec, void* psi, uint64_t ngroups, const DenseKArgs& a, const double2* M,
                             const uint64_t* off, hipStream_t s) {
  const unsigned g = grid_for(ngroups, 256, 1u << 18);
#define QR_DENSEK_CASE(KK)                                                                     \
  case KK:                                                                                     \
    if (prec == 0)                                                                             \
      dev::k_densek_reg<double, KK><<<g, 256, 0, s>>>((double2*)psi, ngroups, a, M, off);      \
    else                                                                                       \
      dev::k_densek_reg<float, KK><<<g, 256, 0, s>>>((float2*)psi, ngroups, a, M, off);        \
    break;
  switch (a.k) {
    QR_DENSEK_CASE(2)
    QR_DENSEK_CASE(3)
    QR_DENSEK_CASE(4)
    default:
      return hipErrorInvalidValue;
  }
#undef QR_DENSEK_CASE
  return hipGetLastError();
}

hipError_t launch_densek(int prec, void* psi, void* tmp, uint64_t tmp_groups, uint64_t ngroups,
                         const DenseKArgs& a, const double2* M, const uint64_t* off, hipStream_t s) {
  const int block = tmp_groups >= 256 ? 256 : (int)tmp_groups;
  const unsigned g = (unsigned)(tmp_groups / block);
  if (prec == 0)
    dev::k_densek<double><<<g, block, 0, s>>>((double2*)psi, (double2*)tmp, ngroups, a, M, off);
  else
    dev::k_densek<float><<<g, block, 0, s>>>((float2*)psi, (float2*)tmp, ngroups, a, M, off);
  return hipGetLastError();
}

hipError_t launch_diagk(int prec, void* psi, uint64_t dim, const DiagKArgs& a, const double2* D,
                        hipStream_t s) {
  const unsigned g = grid_for(dim, 256, 1u << 18);
  if (prec == 0)
    dev::k_diagk<double><<<g, 256, 0, s>>>((double2*)psi, dim, a, D);
  else
    dev::k_diagk<float><<<g, 256, 0, s>>>((float2*)psi, dim, a, D);
  return hipGetLastError();
}


hipError_t launch_sample_fold(int prec, const void* psi, uint64_t dim, const double* u, uint64_t shots,
                              uint64_t* out, hipStream_t s) {
  if (prec == 0)
    dev::k_sample_fold<double><<<1, 1024, 0, s>>>((const double2*)psi, dim, u, shots, out);
  else
    dev::k_sample_fold<float><<<1, 1024, 0, s>>>((const float2*)psi, dim, u, shots, out);
  return hipGetLastError();
}

hipError_t launch_block_mass(int prec, const void* psi, uint64_t dim, int bb, double* mass, hipStream_t s) {
  const uint64_t nb = (dim + (1ull << bb) - 1) >> bb;
  const int nt = streamed(prec, dim);
  if (prec == 0)
    dev::k_block_mass<double><<<(unsigned)nb, 256, 0, s>>>((const double2*)psi, dim, bb, mass, nt);
  else
    dev::k_block_mass<float><<<(unsigned)nb, 256, 0, s>>>((const float2*)psi, dim, bb, mass, nt);
  return hipGetLastError();
}

hipError_t launch_block_exact_sums(int prec, const void* psi, uint64_t dim, int bb, const int32_t* ebase,
                                   double* sums, uint32_t* flags, hipStream_t s) {
  const uint64_t nb = (dim + (1ull << bb) - 1) >> bb;
  const int nt = streamed(prec, dim);
  if (prec == 0)
    dev::k_block_exact_sums<double>
        <<<(unsigned)nb, 256, 0, s>>>((const double2*)psi, dim, bb, ebase, sums, flags, nt);
  else
    dev::k_block_exact_sums<float>
        <<<(unsigned)nb, 256, 0, s>>>((const float2*)psi, dim, bb, ebase, sums, flags, nt);
  return hipGetLastError();
}

hipError_t launch_sample_exact_blocks(int prec, const void* psi, uint64_t dim, int bb, const ExactJob* jobs,
                                      uint64_t njobs, const double* u, uint64_t* out, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  if (bb < 12) return hipErrorInvalidValue;  // blocks are whole 4096-amplitude chunks
  if (prec == 0)
    dev::k_sample_exact_blocks<double><<<(unsigned)njobs, 256, 0, s>>>((const double2*)psi, dim, bb, jobs, u, out);
  else
    dev::k_sample_exact_blocks<float><<<(unsigned)njobs, 256, 0, s>>>((const float2*)psi, dim, bb, jobs, u, out);
  return hipGetLastError();
}

hipError_t launch_norm_partial(int prec, const void* psi, uint64_t dim, double* part, int nblocks,
                               hipStream_t s) {
  if (prec == 0)
    dev::k_norm_partial<double><<<nblocks, 256, 0, s>>>((const double2*)psi, dim, part);
  else
    dev::k_norm_partial<float><<<nblocks, 256, 0, s>>>((const float2*)psi, dim, part);
  return hipGetLastError();
}

hipError_t launch_gen_partial(int prec, const void* lam, const void* psi, uint64_t dim, int t, int gen,
                              double2* part, int nblocks, hipStream_t s) {
  if (prec == 0)
    dev::k_gen_partial<double><<<nblocks, 256, 0, s>>>((const double2*)lam, (const double2*)psi, dim, t, gen, part);
  else
    dev::k_gen_partial<float><<<nblocks, 256, 0, s>>>((const float2*)lam, (const float2*)psi, dim, t, gen, part);
  return hipGetLastError();
}

hipError_t launch_pauli_sum(int prec, void* lam, const void* psi, uint64_t dim, const PauliGroup* groups, int ngroups,
                            const PauliZTerm* terms, hipStream_t s) {
  const unsigned g = grid_for(dim, 256, 1u << 18);
  if (prec == 0)
    dev::k_pauli_sum<double><<<g, 256, 0, s>>>((double2*)lam, (const double2*)psi, dim, groups, ngroups, terms);
  else
    dev::k_pauli_sum<float><<<g, 256, 0, s>>>((float2*)lam, (const float2*)psi, dim, groups, ngroups, terms);
  return hipGetLastError();
}

hipError_t launch_pauli_expect(int prec, const void* psi, uint64_t dim, const PauliGroup* groups, int ngroups,
                               const PauliZTerm* terms, double* part, int nblocks, hipStream_t s, int accumulate) {
  if (prec == 0)
    dev::k_pauli_expect<double>
        <<<nblocks, 256, 0, s>>>((const double2*)psi, dim, groups, ngroups, terms, part, accumulate);
  else
    dev::k_pauli_expect<float>
        <<<nblocks, 256, 0, s>>>((const float2*)psi, dim, groups, ngroups, terms, part, accumulate);
  return hipGetLastError();
}

hipError_t launch_init_instances(int prec, void* psi, uint64_t dim, uint64_t count, hipStream_t s) {
  const uint64_t total = dim * count;
  if (prec == 0)
    dev::k_init_instances<double><<<grid_for(total, 256, 1u << 16), 256, 0, s>>>((double2*)psi, dim, total);
  else
    dev::k_init_instances<float><<<grid_for(total, 256, 1u << 16), 256, 0, s>>>((float2*)psi, dim, total);
  return hipGetLastError();
}

hipError_t launch_sample_fold_instances(int prec, const void* psi, uint64_t dim, uint64_t count, const double* u,
                                        const uint64_t* off, uint64_t* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (count > 0x7fffffffull) return hipErrorInvalidValue;
  if (prec == 0)
    dev::k_sample_fold_instances<double><<<(unsigned)count, 256, 0, s>>>((const double2*)psi, dim, dim, u, off, out);
  else
    dev::k_sample_fold_instances<float><<<(unsigned)count, 256, 0, s>>>((const float2*)psi, dim, dim, u, off, out);
  return hipGetLastError();
}

int pauli_expect_partials(uint64_t dim) { return (int)std::min<uint64_t>(1024, std::max<uint64_t>(1, dim / 256)); }

hipError_t launch_pauli_expect_passes(int prec, const void* psi, uint64_t dim, const PauliGroup* groups,
                                      const uint32_t* lxm, const PauliZTerm* terms, const PauliPass* passes,
                                      int npasses, double* part, hipStream_t s) {
  const int nb = pauli_expect_partials(dim);
  for (int i = 0; i < npasses; ++i) {
    const PauliPass& ps = passes[i];
    const int acc = i > 0 ? 1 : 0;
    if (ps.tmask == 0) {
      const hipError_t e = launch_pauli_expect(prec, psi, dim, groups + ps.first, (int)ps.count, terms, part, nb, s, acc);
      if (e != hipSuccess) return e;
      continue;
    }
    const int K = __builtin_popcountll(ps.tmask);
    if (K != pauli_tile_bits(prec) || (dim >> K) == 0) return hipErrorInvalidValue;
    const uint64_t ntiles = dim >> K, nontile = (dim - 1) & ~ps.tmask;
    if (prec == 0)
      dev::k_pauli_expect_tiles<double, 12><<<nb, 256, 0, s>>>((const double2*)psi, ntiles, ps.tmask, nontile,
                                                              groups + ps.first, lxm + ps.first, (int)ps.count,
                                                              terms, part, acc);
    else
      dev::k_pauli_expect_tiles<float, 13><<<nb, 256, 0, s>>>((const float2*)psi, ntiles, ps.tmask, nontile,
                                                             groups + ps.first, lxm + ps.first, (int)ps.count,
                                                             terms, part, acc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_dot_re_partial(int prec, const void* a, const void* b, uint64_t dim, double* part, int nblocks,
                                 hipStream_t s) {
  if (prec == 0)
    dev::k_dot_re_partial<double><<<nblocks, 256, 0, s>>>((const double2*)a, (const double2*)b, dim, part);
  else
    dev::k_dot_re_partial<float><<<nblocks, 256, 0, s>>>((const float2*)a, (const float2*)b, dim, part);
  return hipGetLastError();
}

hipError_t launch_stream_scale(int prec, void* psi, uint64_t dim, double f, hipStream_t s) {
  const uint64_t g = (dim + 255) / 256;
  if (g > 0x7fffffffull) return hipErrorInvalidValue;
  if (prec == 0)
    dev::k_stream_scale<double><<<(unsigned)g, 256, 0, s>>>((double2*)psi, dim, f);
  else
    dev::k_stream_scale<float><<<(unsigned)g, 256, 0, s>>>((float2*)psi, dim, (float)f);
  return hipGetLastError();
}

hipError_t launch_stamp(uint64_t* out, hipStream_t s) {
  dev::k_stamp<<<1, 64, 0, s>>>(out);
  return hipGetLastError();
}

hipError_t launch_sum_rows(const double* part, uint64_t ncols, uint32_t nrows, double* out, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  dev::k_sum_rows<<<nrows, 256, 0, s>>>(part, ncols, out);
  return hipGetLastError();
}

hipError_t launch_f64_to_c64(float2* dst, const double2* src, uint64_t n, hipStream_t s) {
  dev::k_f64_to_c64<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_c64_to_f64(double2* dst, const float2* src, uint64_t n, hipStream_t s) {
  dev::k_c64_to_f64<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

}  // namespace qr
