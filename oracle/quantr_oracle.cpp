/*
 * quantr_oracle.cpp — CPU restatement of quantr's simulation algorithm. TEST INFRASTRUCTURE.
 *
 * This file is the parity oracle for the HIP engine. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker (or as the timed
 * CPU baseline); the product never calls it.
 *
 * The reference (quantr 0.6.0) is Rust and cannot be built in this environment (no cargo or
 * rustc, dependencies not vendored), so its algorithm is restated here from the source:
 *
 *   ref_hashmap  faithful restatement of Circuit::apply_gate (simulation.rs:64-135): loop over
 *                every basis state in index order (super_position_iter.rs:56-72), compute
 *                the gate image of that basis state (standard_gate_ops.rs / the custom
 *                table), scatter coeff*amp into a hash map keyed by output index
 *                (simulation.rs:158-180), overwrite with "untouched" states
 *                (simulation.rs:120-133), gather back (super_positions_unchecked.rs:50-60).
 *   ref_sv       in-place state vector with the same per-amplitude arithmetic (products
 *                coeff*amp in the textbook complex form, summed in ascending input index
 *                order, first term inserted then added), OpenMP over independent groups.
 *   measure      SuperPosition::measure (super_positions.rs:332-342) with injected draws.
 *   expectation  <psi|O|psi> for Pauli sums, accumulated in long double.
 *
 * Build with -ffp-contract=off so no product/sum is fused (Rust does not contract).
 *
 * Parity pins (tests/): every hand-computed golden vector of the reference's own tests
 * (circuit.rs:603-982, tests/qft.rs:30-39, tests/grovers.rs:124-133), numpy Kronecker
 * brute force on random circuits, closed forms (QFT matrix, inverse pairs), and
 * ref_hashmap == ref_sv bit for bit. The fastrand draw stream (fastrand ^2.1.0) is not
 * restated: draws are injected, so that part of measure_all is "parity unpinned".
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

struct cplx {
  double re, im;
};

// num-complex Mul for Complex<f64>: (a+bi)(c+di) = (ac - bd) + (ad + bc)i, no fusing.
inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
// num-complex exp of a purely imaginary argument: from_polar(exp(0), x) = (1*cos x, 1*sin x).
inline cplx cexp_i(double x) { return {1.0 * std::cos(x), 1.0 * std::sin(x)}; }

const double FRAC_1_SQRT_2 = 0.70710678118654752440084436210484903928483593768847;
const cplx ZERO = {0.0, 0.0};
inline cplx re(double x) { return {x, 0.0}; }
inline cplx im(double x) { return {0.0, x}; }

// Same field layout as qr_op in include/quantr_hip.h (kept separate on purpose).
struct orc_op {
  uint32_t kind, position, ncontrols, reserved0;
  uint64_t controls_offset;
  double arg;
  int32_t arg_int;
  uint32_t reserved1;
  uint64_t image_offset, some_offset;
};
static_assert(sizeof(orc_op) == 56, "orc_op layout");

enum Kind {
  ID = 0, H, X, Y, Z, S, SDAG, T, TDAG, RX, RY, RZ, X90, Y90, MX90, MY90, PHASE,
  CR, CRK, CZ, CY, CNOT, SWAP, TOFFOLI, CUSTOM
};

thread_local std::string g_err;

// ---- standard_gate_ops.rs restated: image (column) of one local basis state -------------
// Each returns the 2^k amplitudes of the image of local basis state `in` (MSB-first over the
// acting wires), exactly as the corresponding function builds its SuperPosition.

void image_single(int kind, double angle, int q, cplx* out) {
  switch (kind) {
    case H:  // standard_gate_ops.rs:37-43
      out[0] = re(FRAC_1_SQRT_2);
      out[1] = re(q ? -FRAC_1_SQRT_2 : FRAC_1_SQRT_2);
      return;
    case RX: {  // :45-56
      cplx rp = re(std::cos(0.5 * angle));
      cplx ip = im(-std::sin(0.5 * angle));
      if (!q) { out[0] = rp; out[1] = ip; } else { out[0] = ip; out[1] = rp; }
      return;
    }
    case RY: {  // :58-70
      cplx c = re(std::cos(0.5 * angle));
      cplx sp = re(std::sin(0.5 * angle));
      cplx sn = re(-std::sin(0.5 * angle));
      if (!q) { out[0] = c; out[1] = sp; } else { out[0] = sn; out[1] = c; }
      return;
    }
    case RZ: {  // :72-83  exp(i*(-angle*0.5)), exp(i*(angle*0.5))
      cplx ne = cexp_i(-angle * 0.5);
      cplx pe = cexp_i(angle * 0.5);
      if (!q) { out[0] = ne; out[1] = ZERO; } else { out[0] = ZERO; out[1] = pe; }
      return;
    }
    case PHASE: {  // global_phase :85-95
      cplx e = cexp_i(angle * 0.5);
      if (!q) { out[0] = e; out[1] = ZERO; } else { out[0] = ZERO; out[1] = e; }
      return;
    }
    case X90:  // :97-103
      if (!q) { out[0] = ZERO; out[1] = im(-1.0); } else { out[0] = im(-1.0); out[1] = ZERO; }
      return;
    case Y90:  // :105-111
      if (!q) { out[0] = ZERO; out[1] = re(-1.0); } else { out[0] = re(1.0); out[1] = ZERO; }
      return;
    case MX90:  // :113-119
      if (!q) { out[0] = ZERO; out[1] = im(1.0); } else { out[0] = im(1.0); out[1] = ZERO; }
      return;
    case MY90:  // :121-127
      if (!q) { out[0] = ZERO; out[1] = re(1.0); } else { out[0] = re(-1.0); out[1] = ZERO; }
      return;
    case T:  // tgate :129-135
      if (!q) { out[0] = re(1.0); out[1] = ZERO; } else { out[0] = ZERO; out[1] = {FRAC_1_SQRT_2, FRAC_1_SQRT_2}; }
      return;
    case TDAG:  // :137-143
      if (!q) { out[0] = re(1.0); out[1] = ZERO; } else { out[0] = ZERO; out[1] = {FRAC_1_SQRT_2, -FRAC_1_SQRT_2}; }
      return;
    case S:  // phase :145-151
      if (!q) { out[0] = re(1.0); out[1] = ZERO; } else { out[0] = ZERO; out[1] = im(1.0); }
      return;
    case SDAG:  // phasedag :153-159
      if (!q) { out[0] = re(1.0); out[1] = ZERO; } else { out[0] = ZERO; out[1] = im(-1.0); }
      return;
    case X:  // pauli_x :161-167
      if (!q) { out[0] = ZERO; out[1] = re(1.0); } else { out[0] = re(1.0); out[1] = ZERO; }
      return;
    case Y:  // pauli_y :169-175
      if (!q) { out[0] = ZERO; out[1] = im(1.0); } else { out[0] = im(-1.0); out[1] = ZERO; }
      return;
    case Z:  // pauli_z :177-183
      if (!q) { out[0] = re(1.0); out[1] = ZERO; } else { out[0] = ZERO; out[1] = re(-1.0); }
      return;
  }
}

// Double gates: local input (q1 = control-side wire, q2 = position), 4 amplitudes.
void image_double(int kind, double angle, int k, int q1, int q2, cplx* out) {
  for (int i = 0; i < 4; ++i) out[i] = ZERO;
  int in = q1 * 2 + q2;
  switch (kind) {
    case CNOT: {  // :189-197
      static const int img[4] = {0, 1, 3, 2};
      out[img[in]] = re(1.0);
      return;
    }
    case CY:  // :199-207
      if (in == 0) out[0] = re(1.0);
      else if (in == 1) out[1] = re(1.0);
      else if (in == 2) out[3] = im(1.0);
      else out[2] = im(-1.0);
      return;
    case CZ:  // :209-217
      if (in == 3) out[3] = re(-1.0); else out[in] = re(1.0);
      return;
    case SWAP: {  // :219-227
      static const int img[4] = {0, 2, 1, 3};
      out[img[in]] = re(1.0);
      return;
    }
    case CR:  // :229-238  exp(i*angle) on |11>
      if (in == 3) out[3] = cexp_i(angle); else out[in] = re(1.0);
      return;
    case CRK: {  // :240-250  exp(i*(2*PI).div(2f64.powi(k)))
      if (in == 3) {
        double p = 1.0;  // 2f64.powi(k): exact power of two
        p = std::ldexp(1.0, k);
        out[3] = cexp_i((2.0 * M_PI) / p);
      } else {
        out[in] = re(1.0);
      }
      return;
    }
  }
}

void image_toffoli(int in, cplx* out) {  // :256-267
  for (int i = 0; i < 8; ++i) out[i] = ZERO;
  int o = in;
  if (in == 6) o = 7; else if (in == 7) o = 6;
  out[o] = re(1.0);
}

// Acting wires of a gate, MSB first: [controls..., position] (simulation.rs:73-112).
bool acting_wires(const orc_op& op, const uint32_t* controls, std::vector<uint32_t>& w) {
  w.clear();
  switch (op.kind) {
    case CR: case CRK: case CZ: case CY: case CNOT: case SWAP:
      if (op.ncontrols != 1) return false;
      w.push_back(controls[op.controls_offset]);
      break;
    case TOFFOLI:
      if (op.ncontrols != 2) return false;
      w.push_back(controls[op.controls_offset]);
      w.push_back(controls[op.controls_offset + 1]);
      break;
    case CUSTOM:
      for (uint32_t c = 0; c < op.ncontrols; ++c) w.push_back(controls[op.controls_offset + c]);
      break;
    default:
      if (op.ncontrols != 0) return false;
  }
  w.push_back(op.position);
  return true;
}

// Image of local basis state `in` (k acting wires). Returns false for a custom None.
bool gate_image(const orc_op& op, int k, uint64_t in, const double* images, const uint8_t* some,
                cplx* out) {
  if (op.kind == CUSTOM) {
    if (!some[op.some_offset + in]) return false;
    uint64_t dim = 1ull << k;
    const double* p = images + op.image_offset + in * dim * 2;
    for (uint64_t r = 0; r < dim; ++r) out[r] = {p[2 * r], p[2 * r + 1]};
    return true;
  }
  if (k == 1) {
    image_single(op.kind, op.arg, (int)(in & 1), out);
  } else if (k == 2) {
    image_double(op.kind, op.arg, op.arg_int, (int)((in >> 1) & 1), (int)(in & 1), out);
  } else {
    image_toffoli((int)in, out);
  }
  return true;
}

inline uint64_t bit_of_wire(uint32_t n, uint32_t w) { return 1ull << (n - 1 - w); }

// Local index of global index i over acting wires (MSB first).
inline uint64_t local_index(uint64_t i, uint32_t n, const std::vector<uint32_t>& w) {
  uint64_t l = 0;
  for (size_t q = 0; q < w.size(); ++q) l = (l << 1) | ((i & bit_of_wire(n, w[q])) ? 1 : 0);
  return l;
}

// insert_qubits (product_states.rs:112-123): write local bits of s into the acting wires of i.
inline uint64_t insert_local(uint64_t i, uint32_t n, const std::vector<uint32_t>& w, uint64_t s) {
  size_t k = w.size();
  for (size_t q = 0; q < k; ++q) {
    uint64_t b = bit_of_wire(n, w[q]);
    if ((s >> (k - 1 - q)) & 1) i |= b; else i &= ~b;
  }
  return i;
}

bool validate(uint32_t n, const orc_op* ops, uint64_t nops, const uint32_t* controls,
              uint64_t ncontrols, uint64_t nimages, uint64_t nsome) {
  std::vector<uint32_t> w;
  for (uint64_t g = 0; g < nops; ++g) {
    const orc_op& op = ops[g];
    if (op.kind >= 25) { g_err = "unknown gate kind"; return false; }
    if (op.kind == ID) continue;
    if (op.controls_offset + op.ncontrols > ncontrols) { g_err = "controls out of range"; return false; }
    if (!acting_wires(op, controls, w)) { g_err = "wrong number of controls"; return false; }
    std::vector<bool> seen(n, false);
    for (uint32_t x : w) {
      if (x >= n) { g_err = "wire out of range"; return false; }
      if (seen[x]) { g_err = "repeated wire"; return false; }
      seen[x] = true;
    }
    if (op.kind == CUSTOM) {
      if (w.size() > 20) { g_err = "custom gate too wide"; return false; }
      uint64_t dim = 1ull << w.size();
      if (op.some_offset + dim > nsome || op.image_offset + dim * dim * 2 > nimages) {
        g_err = "custom table out of range"; return false;
      }
    }
  }
  return true;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
int orc_op_size() { return (int)sizeof(orc_op); }

/* Faithful hash-map restatement of Circuit::apply_gate, gate by gate.
 * state: 2^n interleaved (re, im) doubles, updated in place. Returns 0 on success. */
int orc_simulate_hashmap(uint32_t n, double* state, const orc_op* ops, uint64_t nops,
                         const uint32_t* controls, uint64_t ncontrols, const double* images,
                         uint64_t nimages, const uint8_t* some, uint64_t nsome) {
  if (n == 0 || n > 30) { g_err = "n out of range"; return 1; }
  if (!validate(n, ops, nops, controls, ncontrols, nimages, nsome)) return 1;
  const uint64_t dim = 1ull << n;
  cplx* amps = reinterpret_cast<cplx*>(state);
  std::vector<uint32_t> w;
  std::vector<cplx> img;
  for (uint64_t g = 0; g < nops; ++g) {
    const orc_op& op = ops[g];
    if (op.kind == ID) continue;  // simulation.rs:38-41
    acting_wires(op, controls, w);
    const int k = (int)w.size();
    img.resize(1ull << k);
    std::unordered_map<uint64_t, cplx> mapped;     // simulation.rs:66
    std::vector<std::pair<uint64_t, cplx>> untouched;  // simulation.rs:67
    mapped.reserve(dim);
    for (uint64_t i = 0; i < dim; ++i) {  // every basis state, zeros included
      const cplx amp = amps[i];
      const uint64_t l = local_index(i, n, w);
      if (gate_image(op, k, l, images, some, img.data())) {
        for (uint64_t s = 0; s < (1ull << k); ++s) {  // simulation.rs:168-179
          const uint64_t key = insert_local(i, n, w, s);
          const cplx prod = cmul(img[s], amp);
          auto it = mapped.find(key);
          if (it == mapped.end()) mapped.emplace(key, prod);
          else it->second = cadd(it->second, prod);
        }
      } else {
        untouched.emplace_back(i, amp);
      }
    }
    for (auto& kv : untouched) mapped[kv.first] = kv.second;  // overwrite, :126-133
    for (uint64_t i = 0; i < dim; ++i) {  // super_positions_unchecked.rs:50-60
      auto it = mapped.find(i);
      amps[i] = (it == mapped.end()) ? ZERO : it->second;
    }
  }
  return 0;
}

/* In-place state-vector restatement with the same arithmetic order (bitwise equal to
 * orc_simulate_hashmap). nthreads <= 0 uses all OpenMP threads. */
int orc_simulate_sv(uint32_t n, double* state, const orc_op* ops, uint64_t nops,
                    const uint32_t* controls, uint64_t ncontrols, const double* images,
                    uint64_t nimages, const uint8_t* some, uint64_t nsome, int nthreads) {
  if (n == 0 || n > 36) { g_err = "n out of range"; return 1; }
  if (!validate(n, ops, nops, controls, ncontrols, nimages, nsome)) return 1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  cplx* amps = reinterpret_cast<cplx*>(state);
  std::vector<uint32_t> w;
  for (uint64_t g = 0; g < nops; ++g) {
    const orc_op& op = ops[g];
    if (op.kind == ID) continue;
    acting_wires(op, controls, w);
    const int k = (int)w.size();
    const uint64_t kd = 1ull << k;
    // Matrix columns: col[j] = image of local j (or none).
    std::vector<cplx> cols(kd * kd);
    std::vector<uint8_t> has(kd);
    for (uint64_t j = 0; j < kd; ++j) has[j] = gate_image(op, k, j, images, some, &cols[j * kd]);
    // Offsets of local states and ascending-global-order of inputs.
    std::vector<uint64_t> off(kd);
    uint64_t amask = 0;
    for (uint64_t j = 0; j < kd; ++j) off[j] = insert_local(0, n, w, j);
    for (uint32_t x : w) amask |= bit_of_wire(n, x);
    std::vector<uint64_t> order(kd);
    for (uint64_t j = 0; j < kd; ++j) order[j] = j;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return off[a] < off[b]; });
    const uint64_t groups = 1ull << (n - k);
    // Free bits (non-acting), in ascending order, to enumerate group bases.
    std::vector<uint64_t> free_bits;
    for (uint32_t b = 0; b < n; ++b)
      if (!(amask & (1ull << b))) free_bits.push_back(1ull << b);
#pragma omp parallel
    {
      std::vector<cplx> in(kd), out(kd);
#pragma omp for schedule(static)
      for (int64_t gi = 0; gi < (int64_t)groups; ++gi) {
        uint64_t base = 0;
        for (size_t f = 0; f < free_bits.size(); ++f)
          if ((uint64_t)gi & (1ull << f)) base |= free_bits[f];
        for (uint64_t j = 0; j < kd; ++j) in[j] = amps[base | off[j]];
        for (uint64_t r = 0; r < kd; ++r) {
          if (!has[r] && op.kind == CUSTOM) {  // untouched input overwrites its slot
            out[r] = in[r];
            continue;
          }
          bool first = true;
          cplx acc = ZERO;
          for (uint64_t oi = 0; oi < kd; ++oi) {
            const uint64_t j = order[oi];
            if (!has[j]) continue;
            const cplx prod = cmul(cols[j * kd + r], in[j]);
            acc = first ? prod : cadd(acc, prod);
            first = false;
          }
          out[r] = first ? ZERO : acc;
        }
        for (uint64_t j = 0; j < kd; ++j) amps[base | off[j]] = out[j];
      }
    }
  }
  return 0;
}

/* SuperPosition::measure with injected draws: per shot, left fold of re*re + im*im in index
 * order, first i with u < cum; UINT64_MAX for None. Literal O(shots * 2^n) form. */
int orc_measure_literal(uint32_t n, const double* state, const double* uniforms, uint64_t shots,
                        uint64_t* out) {
  const uint64_t dim = 1ull << n;
  for (uint64_t s = 0; s < shots; ++s) {
    double cum = 0.0;
    out[s] = UINT64_MAX;
    for (uint64_t i = 0; i < dim; ++i) {
      const double re_ = state[2 * i], im_ = state[2 * i + 1];
      cum += re_ * re_ + im_ * im_;  // norm_sqr, then the fold
      if (uniforms[s] < cum) { out[s] = i; break; }
    }
  }
  return 0;
}

/* Same results as orc_measure_literal from one pass over the state (draws visited in sorted
 * order; the fold values are the same sequence). */
int orc_measure(uint32_t n, const double* state, const double* uniforms, uint64_t shots,
                uint64_t* out) {
  const uint64_t dim = 1ull << n;
  std::vector<uint64_t> ord(shots);
  for (uint64_t s = 0; s < shots; ++s) ord[s] = s;
  std::sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) { return uniforms[a] < uniforms[b]; });
  double cum = 0.0;
  uint64_t p = 0;
  for (uint64_t i = 0; i < dim && p < shots; ++i) {
    const double re_ = state[2 * i], im_ = state[2 * i + 1];
    cum += re_ * re_ + im_ * im_;
    while (p < shots && uniforms[ord[p]] < cum) out[ord[p++]] = i;
  }
  for (; p < shots; ++p) out[ord[p]] = UINT64_MAX;
  return 0;
}

/* <psi|O|psi> for a Pauli sum; same term encoding as qr_observable. Long-double accumulation. */
struct orc_term {
  double coeff;
  uint32_t nfactors, reserved;
  uint64_t offset;
};
int orc_expectation(uint32_t n, const double* state, const orc_term* terms, uint64_t nterms,
                    const uint32_t* wires, const uint8_t* paulis, double* out) {
  const uint64_t dim = 1ull << n;
  long double total = 0.0L;
  for (uint64_t t = 0; t < nterms; ++t) {
    uint64_t xm = 0, zm = 0;
    int ny = 0;
    for (uint32_t f = 0; f < terms[t].nfactors; ++f) {
      const uint32_t w = wires[terms[t].offset + f];
      if (w >= n) { g_err = "observable wire out of range"; return 1; }
      const uint64_t b = bit_of_wire(n, w);
      switch (paulis[terms[t].offset + f]) {
        case 'X': case 1: xm ^= b; break;
        case 'Y': case 2: xm ^= b; zm ^= b; ny++; break;
        case 'Z': case 3: zm ^= b; break;
        default: g_err = "bad pauli"; return 1;
      }
    }
    // P|j> = i^ny (-1)^{popcount(j & zm)} |j ^ xm>  (Y = iXZ per factor)
    long double acc_re = 0.0L, acc_im = 0.0L;
    for (uint64_t j = 0; j < dim; ++j) {
      const uint64_t i = j ^ xm;
      const long double sgn = (__builtin_popcountll(j & zm) & 1) ? -1.0L : 1.0L;
      // conj(psi_i) * psi_j * sgn
      const long double ar = state[2 * i], ai = -state[2 * i + 1];
      const long double br = state[2 * j], bi = state[2 * j + 1];
      acc_re += sgn * (ar * br - ai * bi);
      acc_im += sgn * (ar * bi + ai * br);
    }
    // multiply by i^ny
    long double r = acc_re, m = acc_im;
    for (int q = 0; q < (ny & 3); ++q) { long double tmp = r; r = -m; m = tmp; }
    total += (long double)terms[t].coeff * r;
  }
  *out = (double)total;
  return 0;
}

}  // extern "C"
