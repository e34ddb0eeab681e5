// tile_ops.hpp — register-level gate kernels shared by the interpreted tile sweep
// (kernels.hip) and the runtime-specialised tile sweeps (jit.cpp compiles this file's text
// with hipRTC). Self-contained: no includes, only HIP built-ins.
#pragma once

namespace qr {
namespace dev {

template <typename R>
struct CT;
template <>
struct CT<double> {
  using T = double2;
};
template <>
struct CT<float> {
  using T = float2;
};


template <typename R, int RB>
struct Regs {
  typename CT<R>::T v[1 << RB];
};

// Coefficients travel by value (never through a pointer to a local array, which would put
// them in scratch memory).
struct Coef4 {
  double2 c[4];
};

// The LDS element at byte offset `off`. Remaps address LDS by per-thread byte offsets plus
// constants that fold into the access's immediate offset.
template <typename T>
__device__ __forceinline__ T& lds_at(T* base, uint32_t off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}

// ---- complex helpers -------------------------------------------------------------------
// factor forms of a diagonal entry: 0 general, 1 = -1, 2 = +i, 3 = -i
template <typename R, int FORM>
__device__ __forceinline__ void cscale(typename CT<R>::T& x, R fr, R fi) {
  const R xr = x.x, xi = x.y;
  if (FORM == 1) {
    x.x = -xr;
    x.y = -xi;
  } else if (FORM == 2) {
    x.x = -xi;
    x.y = xr;
  } else if (FORM == 3) {
    x.x = xi;
    x.y = -xr;
  } else {
    x.x = fma(fr, xr, -fi * xi);
    x.y = fma(fr, xi, fi * xr);
  }
}

// Rotations as three shears: x += al*y, y += be*x, x += al*y with al = -tan(phi/2) and
// be = sin(phi) give [[cos phi, -sin phi], [sin phi, cos phi]]. The host reduces phi to
// [-pi/2, pi/2] (so |al|, |be| <= 1) and folds the sign and the magnitude into the stage's
// scalar: 6 fused multiply-adds per amplitude pair and no branch on the angle.
template <typename R, int RB, int J, bool HAS_NS>
__device__ __forceinline__ void rot_real(Regs<R, RB>& a, R al, R be, uint32_t ns) {
#pragma unroll
  for (int m0 = 0; m0 < (1 << RB); ++m0) {
    if (m0 & (1 << J)) continue;
    if (HAS_NS && ((uint32_t)m0 & ns) != ns) continue;
    const int m1 = m0 | (1 << J);
    R xr = a.v[m0].x, xi = a.v[m0].y, yr = a.v[m1].x, yi = a.v[m1].y;
    xr = fma(al, yr, xr); xi = fma(al, yi, xi);
    yr = fma(be, xr, yr); yi = fma(be, xi, yi);
    xr = fma(al, yr, xr); xi = fma(al, yi, xi);
    a.v[m0].x = xr; a.v[m0].y = xi;
    a.v[m1].x = yr; a.v[m1].y = yi;
  }
}
// [[cos, -i sin], [-i sin, cos]] is the real rotation of the pair (x, i y): the same shears
// with y's real and imaginary parts exchanged.
template <typename R, int RB, int J, bool HAS_NS>
__device__ __forceinline__ void rot_x(Regs<R, RB>& a, R al, R be, uint32_t ns) {
#pragma unroll
  for (int m0 = 0; m0 < (1 << RB); ++m0) {
    if (m0 & (1 << J)) continue;
    if (HAS_NS && ((uint32_t)m0 & ns) != ns) continue;
    const int m1 = m0 | (1 << J);
    R xr = a.v[m0].x, xi = a.v[m0].y, yr = a.v[m1].x, yi = a.v[m1].y;
    xr = fma(-al, yi, xr); xi = fma(al, yr, xi);
    yi = fma(-be, xr, yi); yr = fma(be, xi, yr);
    xr = fma(-al, yi, xr); xi = fma(al, yr, xi);
    a.v[m0].x = xr; a.v[m0].y = xi;
    a.v[m1].x = yr; a.v[m1].y = yi;
  }
}

// 2x2 on register slot J for every register pair whose register-held controls (ns) are 1.
template <typename R, int RB, int J, int FORM, bool HAS_NS>
__device__ __forceinline__ void u1_slot(Regs<R, RB>& a, const Coef4 u, uint32_t ns) {
  const R u00r = (R)u.c[0].x, u00i = (R)u.c[0].y, u01r = (R)u.c[1].x, u01i = (R)u.c[1].y;
  const R u10r = (R)u.c[2].x, u10i = (R)u.c[2].y, u11r = (R)u.c[3].x, u11i = (R)u.c[3].y;
  if (FORM == 7 || FORM == 8) {  // rotations as shears: al in c[0], be in c[1]
    if (FORM == 7) rot_real<R, RB, J, HAS_NS>(a, u00r, u01r, ns);
    else rot_x<R, RB, J, HAS_NS>(a, u00r, u01r, ns);
    return;
  }
#pragma unroll
  for (int m0 = 0; m0 < (1 << RB); ++m0) {
    if (m0 & (1 << J)) continue;
    if (HAS_NS && ((uint32_t)m0 & ns) != ns) continue;
    const int m1 = m0 | (1 << J);
    const R xr = a.v[m0].x, xi = a.v[m0].y, yr = a.v[m1].x, yi = a.v[m1].y;
    R o0r, o0i, o1r, o1i;
    if (FORM == 3) {  // X: swap
      o0r = yr; o0i = yi; o1r = xr; o1i = xi;
    } else if (FORM == 4) {  // real diagonal, imaginary off-diagonal (Rx)
      o0r = fma(-u01i, yi, u00r * xr);
      o0i = fma(u01i, yr, u00r * xi);
      o1r = fma(-u10i, xi, u11r * yr);
      o1i = fma(u10i, xr, u11r * yi);
    } else if (FORM == 5) {  // [[1,1],[1,-1]]
      o0r = xr + yr; o0i = xi + yi; o1r = xr - yr; o1i = xi - yi;
    } else if (FORM == 6) {  // [[0,-1],[1,0]]
      o0r = -yr; o0i = -yi; o1r = xr; o1i = xi;
    } else if (FORM == 1) {  // anti-diagonal
      o0r = fma(u01r, yr, -u01i * yi);
      o0i = fma(u01r, yi, u01i * yr);
      o1r = fma(u10r, xr, -u10i * xi);
      o1i = fma(u10r, xi, u10i * xr);
    } else if (FORM == 2) {  // real
      o0r = fma(u00r, xr, u01r * yr);
      o0i = fma(u00r, xi, u01r * yi);
      o1r = fma(u10r, xr, u11r * yr);
      o1i = fma(u10r, xi, u11r * yi);
    } else {
      o0r = u00r * xr; o0i = u00r * xi;
      o0r = fma(-u00i, xi, o0r); o0i = fma(u00i, xr, o0i);
      o0r = fma(u01r, yr, o0r); o0i = fma(u01r, yi, o0i);
      o0r = fma(-u01i, yi, o0r); o0i = fma(u01i, yr, o0i);
      o1r = u10r * xr; o1i = u10r * xi;
      o1r = fma(-u10i, xi, o1r); o1i = fma(u10i, xr, o1i);
      o1r = fma(u11r, yr, o1r); o1i = fma(u11r, yi, o1i);
      o1r = fma(-u11i, yi, o1r); o1i = fma(u11i, yr, o1i);
    }
    a.v[m0].x = o0r; a.v[m0].y = o0i;
    a.v[m1].x = o1r; a.v[m1].y = o1i;
  }
}

// X on register slot J in the lanes where `ok` holds (controls held by threads; compiled
// stages, where J is a constant): a branch on `ok` in which the lanes that take it exchange
// the amplitudes. Moves under the lane mask (30-qubit brickwork 156.6 -> 150.9 ms) cost less
// than a select of every 32-bit half in every lane.
template <typename R, int RB, int J, bool HAS_NS>
__device__ __forceinline__ void u1_xswap(Regs<R, RB>& a, bool ok, uint32_t ns) {
  if (ok) {
#pragma unroll
    for (int m0 = 0; m0 < (1 << RB); ++m0) {
      if (m0 & (1 << J)) continue;
      if (HAS_NS && ((uint32_t)m0 & ns) != ns) continue;
      const int m1 = m0 | (1 << J);
      const typename CT<R>::T x = a.v[m0];
      a.v[m0].x = a.v[m1].x; a.v[m0].y = a.v[m1].y;
      a.v[m1].x = x.x; a.v[m1].y = x.y;
    }
  }
}

template <typename R, int RB, int J, bool HAS_NS>
__device__ __forceinline__ void u1_form(Regs<R, RB>& a, const Coef4 u, uint32_t form, uint32_t ns) {
  switch (form) {
    case 3: u1_slot<R, RB, J, 3, HAS_NS>(a, u, ns); break;
    case 1: u1_slot<R, RB, J, 1, HAS_NS>(a, u, ns); break;
    case 2: u1_slot<R, RB, J, 2, HAS_NS>(a, u, ns); break;
    case 4: u1_slot<R, RB, J, 4, HAS_NS>(a, u, ns); break;
    case 5: u1_slot<R, RB, J, 5, HAS_NS>(a, u, ns); break;
    case 6: u1_slot<R, RB, J, 6, HAS_NS>(a, u, ns); break;
    case 7: u1_slot<R, RB, J, 7, HAS_NS>(a, u, ns); break;
    case 8: u1_slot<R, RB, J, 8, HAS_NS>(a, u, ns); break;
    default: u1_slot<R, RB, J, 0, HAS_NS>(a, u, ns); break;
  }
}

template <typename R, int RB, bool HAS_NS>
__device__ __forceinline__ void u1_dispatch(Regs<R, RB>& a, uint32_t slot, const Coef4 u, uint32_t form,
                                            uint32_t ns) {
  switch (slot) {
    case 0: u1_form<R, RB, 0, HAS_NS>(a, u, form, ns); break;
    case 1: if constexpr (RB > 1) u1_form<R, RB, 1, HAS_NS>(a, u, form, ns); break;
    case 2: if constexpr (RB > 2) u1_form<R, RB, 2, HAS_NS>(a, u, form, ns); break;
    case 3: if constexpr (RB > 3) u1_form<R, RB, 3, HAS_NS>(a, u, form, ns); break;
    case 4: if constexpr (RB > 4) u1_form<R, RB, 4, HAS_NS>(a, u, form, ns); break;
    default: break;
  }
}

// diag on register slot J: registers with bit J = 1 get d1, the others d0 (skipped if ONLY1).
template <typename R, int RB, int J, bool ONLY1, int FORM, bool HAS_NS>
__device__ __forceinline__ void d1_reg(Regs<R, RB>& a, const Coef4 d, uint32_t ns) {
  const R d0r = (R)d.c[0].x, d0i = (R)d.c[0].y, d1r = (R)d.c[1].x, d1i = (R)d.c[1].y;
#pragma unroll
  for (int m = 0; m < (1 << RB); ++m) {
    if (HAS_NS && ((uint32_t)m & ns) != ns) continue;
    const bool bit = (m >> J) & 1;
    if (ONLY1 && !bit) continue;
    if (bit) cscale<R, ONLY1 ? FORM : 0>(a.v[m], d1r, d1i);
    else cscale<R, 0>(a.v[m], d0r, d0i);
  }
}

template <typename R, int RB, int J, bool HAS_NS>
__device__ __forceinline__ void d1_reg_form(Regs<R, RB>& a, const Coef4 d, uint32_t flags, uint32_t ns) {
  if (!(flags & 1)) {
    d1_reg<R, RB, J, false, 0, HAS_NS>(a, d, ns);
    return;
  }
  switch (flags >> 4) {
    case 1: d1_reg<R, RB, J, true, 1, HAS_NS>(a, d, ns); break;
    case 2: d1_reg<R, RB, J, true, 2, HAS_NS>(a, d, ns); break;
    case 3: d1_reg<R, RB, J, true, 3, HAS_NS>(a, d, ns); break;
    default: d1_reg<R, RB, J, true, 0, HAS_NS>(a, d, ns); break;
  }
}

template <typename R, int RB, bool HAS_NS>
__device__ __forceinline__ void d1_reg_dispatch(Regs<R, RB>& a, uint32_t slot, const Coef4 d, uint32_t flags,
                                                uint32_t ns) {
  switch (slot) {
    case 0: d1_reg_form<R, RB, 0, HAS_NS>(a, d, flags, ns); break;
    case 1: if constexpr (RB > 1) d1_reg_form<R, RB, 1, HAS_NS>(a, d, flags, ns); break;
    case 2: if constexpr (RB > 2) d1_reg_form<R, RB, 2, HAS_NS>(a, d, flags, ns); break;
    case 3: if constexpr (RB > 3) d1_reg_form<R, RB, 3, HAS_NS>(a, d, flags, ns); break;
    case 4: if constexpr (RB > 4) d1_reg_form<R, RB, 4, HAS_NS>(a, d, flags, ns); break;
    default: break;
  }
}

// every register (whose register-held controls hold) times one factor
template <typename R, int RB, int FORM, bool HAS_NS>
__device__ __forceinline__ void scale_all(Regs<R, RB>& a, R fr, R fi, uint32_t ns) {
#pragma unroll
  for (int m = 0; m < (1 << RB); ++m) {
    if (HAS_NS && ((uint32_t)m & ns) != ns) continue;
    cscale<R, FORM>(a.v[m], fr, fi);
  }
}

// ---- adjoint sweep: gradient taps ----------------------------------------------------------
// Im(conj(l) p) and Re(conj(l) p), accumulated in binary64 for both precisions.
template <typename C>
__device__ __forceinline__ double imc(const C& l, const C& p) {
  return (double)l.x * (double)p.y - (double)l.y * (double)p.x;
}
template <typename C>
__device__ __forceinline__ double rec(const C& l, const C& p) {
  return (double)l.x * (double)p.x + (double)l.y * (double)p.y;
}

// sum_j Im(conj(l_j) (G p)_j) over this thread's registers, G = X (GEN 1) or Y (GEN 2) on
// register slot J: (X p)_j = p_{j^J}; (Y p)_j = -i (-1)^{bit J of j} p_{j^J}.
template <typename R, int RB, int J, int GEN>
__device__ __forceinline__ double tap_pair(const Regs<R, RB>& l, const Regs<R, RB>& p) {
  double acc = 0.0;
#pragma unroll
  for (int m0 = 0; m0 < (1 << RB); ++m0) {
    if (m0 & (1 << J)) continue;
    const int m1 = m0 | (1 << J);
    if (GEN == 1) acc += imc(l.v[m0], p.v[m1]) + imc(l.v[m1], p.v[m0]);
    else acc += rec(l.v[m1], p.v[m0]) - rec(l.v[m0], p.v[m1]);
  }
  return acc;
}

// sum_j s_j Im(conj(l_j) p_j) with s_j = (-1)^{bit J of j} (J >= 0) or 1 (J < 0).
template <typename R, int RB, int J>
__device__ __forceinline__ double tap_diag(const Regs<R, RB>& l, const Regs<R, RB>& p) {
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < (1 << RB); ++m) {
    const double v = imc(l.v[m], p.v[m]);
    if (J >= 0 && (m & (1 << J))) acc -= v;
    else acc += v;
  }
  return acc;
}

// Sums `acc` over each group of min(TPB, 64) threads (one wavefront, or the whole workgroup
// when it is smaller) and stores one partial per group: row[tile * groups + group].
template <int TPB>
__device__ __forceinline__ void tap_store(double acc, double* row, unsigned long long tile, unsigned int tid) {
  constexpr int W = TPB < 64 ? TPB : 64;
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, W);
  if ((tid & (W - 1)) == 0) row[tile * (unsigned long long)(TPB / W) + tid / W] = acc;
}

// Amplitude loads and stores marked non-temporal (streamed once per sweep, not kept in cache),
// as one 16- or 8-byte access each.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load_nt(double2& d, const double2* p) {
  const nt_d2 v = __builtin_nontemporal_load((const nt_d2*)p);
  d.x = v.x;
  d.y = v.y;
}
__device__ __forceinline__ void load_nt(float2& d, const float2* p) {
  const nt_f2 v = __builtin_nontemporal_load((const nt_f2*)p);
  d.x = v.x;
  d.y = v.y;
}
__device__ __forceinline__ void store_nt(double2* p, const double2& s) {
  nt_d2 v;
  v.x = s.x;
  v.y = s.y;
  __builtin_nontemporal_store(v, (nt_d2*)p);
}
__device__ __forceinline__ void store_nt(float2* p, const float2& s) {
  nt_f2 v;
  v.x = s.x;
  v.y = s.y;
  __builtin_nontemporal_store(v, (nt_f2*)p);
}

// Deposits the bits of `v` into the set bits of `mask`, lowest first (uniform scalar work).
__device__ __forceinline__ unsigned long long deposit_bits(unsigned long long v, unsigned long long mask) {
  unsigned long long out = 0;
  while (mask) {
    const unsigned long long low = mask & (0ull - mask);
    if (v & 1) out |= low;
    v >>= 1;
    mask ^= low;
  }
  return out;
}

}  // namespace dev
}  // namespace qr
