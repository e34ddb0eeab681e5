// lower.cpp — validation of gate programs and lowering of quantr gates to matrices.
//
// Gate matrices follow the images defined in standard_gate_ops.rs (column j = image of local
// basis state j over the acting wires [controls..., position], simulation.rs:73-112). The
// coefficients are computed once per gate on the host in binary64 with the C library's
// cos/sin, the same libm a Linux Rust build calls for f64::cos/sin.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>

#include "engine.hpp"

namespace qr {

namespace {

thread_local std::string t_err;

const double kS = 0.70710678118654752440084436210484903928483593768847;  // FRAC_1_SQRT_2

// exp(i x) as num-complex computes it for a purely imaginary argument: (cos x, sin x).
cd expi(double x) { return cd(std::cos(x), std::sin(x)); }

Error arg_error(const std::string& m) { return Error{QR_ERR_ARG, m}; }

int wires_expected(uint32_t kind) {
  switch (kind) {
    case QR_GATE_CR: case QR_GATE_CRK: case QR_GATE_CZ: case QR_GATE_CY:
    case QR_GATE_CNOT: case QR_GATE_SWAP:
      return 1;
    case QR_GATE_TOFFOLI:
      return 2;
    case QR_GATE_CUSTOM:
      return -1;
    default:
      return 0;
  }
}

// Matrix of a standard gate, M[r*dim + j] (standard_gate_ops.rs:37-267).
std::vector<cd> standard_matrix(uint32_t kind, double th, int kk) {
  const cd z(0, 0), one(1, 0), I(0, 1);
  switch (kind) {
    case QR_GATE_H: return {cd(kS, 0), cd(kS, 0), cd(kS, 0), cd(-kS, 0)};
    case QR_GATE_X: return {z, one, one, z};
    case QR_GATE_Y: return {z, cd(0, -1), I, z};
    case QR_GATE_Z: return {one, z, z, cd(-1, 0)};
    case QR_GATE_S: return {one, z, z, I};
    case QR_GATE_SDAG: return {one, z, z, cd(0, -1)};
    case QR_GATE_T: return {one, z, z, cd(kS, kS)};
    case QR_GATE_TDAG: return {one, z, z, cd(kS, -kS)};
    case QR_GATE_RX: {
      const double c = std::cos(0.5 * th), s = std::sin(0.5 * th);
      return {cd(c, 0), cd(0, -s), cd(0, -s), cd(c, 0)};
    }
    case QR_GATE_RY: {
      const double c = std::cos(0.5 * th), s = std::sin(0.5 * th);
      return {cd(c, 0), cd(-s, 0), cd(s, 0), cd(c, 0)};
    }
    case QR_GATE_RZ: return {expi(-th * 0.5), z, z, expi(th * 0.5)};
    case QR_GATE_PHASE: return {expi(th * 0.5), z, z, expi(th * 0.5)};
    case QR_GATE_X90: return {z, cd(0, -1), cd(0, -1), z};
    case QR_GATE_Y90: return {z, one, cd(-1, 0), z};
    case QR_GATE_MX90: return {z, I, I, z};
    case QR_GATE_MY90: return {z, cd(-1, 0), one, z};
    default: break;
  }
  std::vector<cd> m(16, z);
  auto set_perm = [&](const int* img) { for (int j = 0; j < 4; ++j) m[img[j] * 4 + j] = one; };
  switch (kind) {
    case QR_GATE_CNOT: { const int p[4] = {0, 1, 3, 2}; set_perm(p); return m; }
    case QR_GATE_SWAP: { const int p[4] = {0, 2, 1, 3}; set_perm(p); return m; }
    case QR_GATE_CZ: m[0] = one; m[5] = one; m[10] = one; m[15] = cd(-1, 0); return m;
    case QR_GATE_CY: m[0] = one; m[5] = one; m[3 * 4 + 2] = I; m[2 * 4 + 3] = cd(0, -1); return m;
    case QR_GATE_CR: m[0] = one; m[5] = one; m[10] = one; m[15] = expi(th); return m;
    case QR_GATE_CRK: {
      m[0] = one; m[5] = one; m[10] = one;
      m[15] = expi((2.0 * M_PI) / std::ldexp(1.0, kk));  // (2*PI).div(2f64.powi(k))
      return m;
    }
    case QR_GATE_TOFFOLI: {
      std::vector<cd> t(64, z);
      for (int j = 0; j < 8; ++j) {
        int o = j == 6 ? 7 : (j == 7 ? 6 : j);
        t[o * 8 + j] = one;
      }
      return t;
    }
    default: break;
  }
  return {};
}

const char* kind_name(uint32_t k) {
  static const char* names[] = {"Id", "H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx",
                                "Ry", "Rz", "X90", "Y90", "MX90", "MY90", "Phase", "CR",
                                "CRk", "CZ", "CY", "CNot", "Swap", "Toffoli", "Custom"};
  return k < QR_GATE_KIND_COUNT ? names[k] : "?";
}

std::string fmt_double(double x) {
  std::ostringstream os;
  os.precision(17);
  os << x;
  std::string s = os.str();
  if (s.find_first_of(".eE") == std::string::npos && s.find("inf") == std::string::npos &&
      s.find("nan") == std::string::npos)
    s += ".0";
  return s;
}

}  // namespace

void set_error(const std::string& msg) { t_err = msg; }
void clear_error() { t_err.clear(); }
const char* last_error_cstr() { return t_err.c_str(); }

uint64_t FastOp::target_mask() const {
  if (kind == FK_DIAG1 || kind == FK_DIAGK) return 0;
  // a gradient tap reads psi and lam; X/Y generators pair amplitudes across their bit, so
  // they are ordered like a gate on it; Z and the identity commute with every diagonal
  if (kind == FK_GRAD) return (gen == 1 || gen == 2) ? (1ull << bits[0]) : 0;
  uint64_t m = 0;
  for (int i = 0; i < k; ++i) m |= 1ull << bits[i];
  return m;
}

uint64_t FastOp::read_mask() const {
  uint64_t m = cmask;
  if (kind == FK_GRAD) return gen == 3 ? (1ull << bits[0]) : 0;
  if (kind == FK_DIAG1 || kind == FK_DIAGK)
    for (int i = 0; i < k; ++i) m |= 1ull << bits[i];
  return m;
}

std::string gate_debug(const qr_op& op, const uint32_t* controls) {
  std::ostringstream os;
  const uint32_t* c = controls ? controls + op.controls_offset : nullptr;
  switch (op.kind) {
    case QR_GATE_RX: case QR_GATE_RY: case QR_GATE_RZ: case QR_GATE_PHASE:
      os << kind_name(op.kind) << "(" << fmt_double(op.arg) << ")"; break;
    case QR_GATE_CR:
      os << "CR(" << fmt_double(op.arg) << ", " << (c ? c[0] : 0) << ")"; break;
    case QR_GATE_CRK:
      os << "CRk(" << op.arg_int << ", " << (c ? c[0] : 0) << ")"; break;
    case QR_GATE_CZ: case QR_GATE_CY: case QR_GATE_CNOT: case QR_GATE_SWAP:
      os << kind_name(op.kind) << "(" << (c ? c[0] : 0) << ")"; break;
    case QR_GATE_TOFFOLI:
      os << "Toffoli(" << (c ? c[0] : 0) << ", " << (c ? c[1] : 0) << ")"; break;
    case QR_GATE_CUSTOM: {
      os << "Custom(<fn>, [";
      for (uint32_t i = 0; i < op.ncontrols; ++i) os << (i ? ", " : "") << (c ? c[i] : 0);
      os << "], _)";
      break;
    }
    default: os << kind_name(op.kind);
  }
  return os.str();
}

// Mirrors the checks of has_overlapping_controls_and_target (circuit.rs:272-293) and the
// position bound of add_gates_with_positions (circuit.rs:157-166); also checks what the
// reference would only discover by panicking (custom table sizes, circuit.rs has no check).
void validate_program(uint32_t n, const qr_program* p) {
  if (n == 0) throw arg_error("The initialised circuit must have at least one wire.");
  if (!p) throw arg_error("program is NULL");
  if (p->nops && !p->ops) throw arg_error("program ops pointer is NULL");
  for (uint64_t g = 0; g < p->nops; ++g) {
    const qr_op& op = p->ops[g];
    if (op.kind >= QR_GATE_KIND_COUNT) throw arg_error("unknown gate kind " + std::to_string(op.kind));
    if (op.kind == QR_GATE_ID) continue;
    if (op.position >= n) {
      std::ostringstream os;
      os << "The position, " << op.position << ", is out of bounds for the circuit with " << n
         << " qubits.";
      throw arg_error(os.str());
    }
    const int expect = wires_expected(op.kind);
    if (expect >= 0 && (int)op.ncontrols != expect)
      throw arg_error("gate " + std::string(kind_name(op.kind)) + " needs " + std::to_string(expect) +
                      " control wire(s), got " + std::to_string(op.ncontrols));
    if (op.kind == QR_GATE_CUSTOM && op.ncontrols + 1ull > 16)
      throw Error{QR_ERR_UNSUPPORTED, "custom gates act on at most 16 wires"};
    // ranges compared without overflow: offset <= size and count <= size - offset
    auto in_range = [](uint64_t off, uint64_t count, uint64_t size) { return off <= size && count <= size - off; };
    if (op.ncontrols) {
      if (!p->controls || !in_range(op.controls_offset, op.ncontrols, p->ncontrols))
        throw arg_error("control list of gate " + std::to_string(g) + " is out of range");
    }
    const uint32_t* c = op.ncontrols ? p->controls + op.controls_offset : nullptr;
    std::string dbg = gate_debug(op, p->controls);
    // repeated control wires (contains_repeating_values, circuit.rs:296-305)
    for (uint32_t i = 0; i < op.ncontrols; ++i)
      for (uint32_t j = i + 1; j < op.ncontrols; ++j)
        if (c[i] == c[j] && c[i] < n)
          throw arg_error("The gate, " + dbg + ", has overlapping control nodes.");
    for (uint32_t i = 0; i < op.ncontrols; ++i)
      if (c[i] == op.position) {
        std::ostringstream os;
        os << "The gate, " << dbg << ", has a control node that equals the gate's position "
           << op.position << ".";
        throw arg_error(os.str());
      }
    for (uint32_t i = 0; i < op.ncontrols; ++i)
      if (c[i] >= n) {
        std::ostringstream os;
        os << "The control node at position " << c[i]
           << ", is greater than the umnber of qubits " << n << ".";
        throw arg_error(os.str());
      }
    if (op.kind == QR_GATE_CUSTOM) {
      const uint32_t k = op.ncontrols + 1;
      const uint64_t dim = 1ull << k;
      if (!p->some || !in_range(op.some_offset, dim, p->nsome))
        throw arg_error("custom gate " + std::to_string(g) + ": Some/None table out of range");
      if (!p->images || !in_range(op.image_offset, dim * dim * 2, p->nimages))
        throw arg_error("custom gate " + std::to_string(g) + ": image table out of range (needs " +
                        std::to_string(dim) + " images of " + std::to_string(dim) + " amplitudes)");
    }
  }
}

std::vector<ExactGate> lower_exact(uint32_t n, const qr_program* p) {
  validate_program(n, p);
  std::vector<ExactGate> out;
  out.reserve(p->nops);
  for (uint64_t g = 0; g < p->nops; ++g) {
    const qr_op& op = p->ops[g];
    if (op.kind == QR_GATE_ID) continue;  // simulation.rs:38-41
    ExactGate e;
    e.kind = op.kind;
    e.arg = op.arg;
    const uint32_t* c = op.ncontrols ? p->controls + op.controls_offset : nullptr;
    for (uint32_t i = 0; i < op.ncontrols; ++i) e.bits.push_back((int)(n - 1 - c[i]));
    e.bits.push_back((int)(n - 1 - op.position));
    e.K = (int)e.bits.size();
    const uint64_t dim = 1ull << e.K;
    e.some.assign(dim, 1);
    e.keep.assign(dim, 0);
    if (op.kind == QR_GATE_CUSTOM) {
      e.custom = true;
      e.M.assign(dim * dim, cd(0, 0));
      const double* img = p->images + op.image_offset;
      for (uint64_t j = 0; j < dim; ++j) e.some[j] = p->some[op.some_offset + j] ? 1 : 0;
      // Overwrite rule (simulation.rs:120-133): an input whose image is None keeps its
      // amplitude and overwrites whatever the other inputs scattered into its slot.
      for (uint64_t r = 0; r < dim; ++r) {
        if (!e.some[r]) {
          e.keep[r] = 1;
          e.M[r * dim + r] = cd(1, 0);
          continue;
        }
        for (uint64_t j = 0; j < dim; ++j)
          if (e.some[j]) e.M[r * dim + j] = cd(img[(j * dim + r) * 2], img[(j * dim + r) * 2 + 1]);
      }
    } else {
      e.M = standard_matrix(op.kind, op.arg, op.arg_int);
    }
    out.push_back(std::move(e));
  }
  return out;
}

namespace {

bool is_zero(cd x) { return x.real() == 0.0 && x.imag() == 0.0; }
bool is_one(cd x) { return x.real() == 1.0 && x.imag() == 0.0; }

// Restricts matrix A over wires W (MSB first) to the block where local wire q is 1.
std::vector<cd> restrict_to_one(const std::vector<cd>& A, int K, int q) {
  const int dim = 1 << K, sub = 1 << (K - 1);
  const int bq = K - 1 - q;  // bit of wire q in the local index
  auto expand = [&](int s) {  // sub-index -> full index with bit bq = 1
    int lo = s & ((1 << bq) - 1);
    int hi = s >> bq;
    return (hi << (bq + 1)) | (1 << bq) | lo;
  };
  std::vector<cd> R(sub * sub);
  for (int r = 0; r < sub; ++r)
    for (int j = 0; j < sub; ++j) R[r * sub + j] = A[expand(r) * dim + expand(j)];
  return R;
}

// Local wire q is a control of A: A never changes wire q, and acts as the identity
// whenever wire q is 0.
bool is_control_wire(const std::vector<cd>& A, int K, int q) {
  const int dim = 1 << K, bq = K - 1 - q;
  for (int r = 0; r < dim; ++r)
    for (int j = 0; j < dim; ++j) {
      const bool rb = (r >> bq) & 1, jb = (j >> bq) & 1;
      const cd v = A[r * dim + j];
      if (rb != jb) {
        if (!is_zero(v)) return false;
      } else if (!jb) {
        if (r == j ? !is_one(v) : !is_zero(v)) return false;
      }
    }
  return true;
}

bool is_diagonal(const std::vector<cd>& A, int K) {
  const int dim = 1 << K;
  for (int r = 0; r < dim; ++r)
    for (int j = 0; j < dim; ++j)
      if (r != j && !is_zero(A[r * dim + j])) return false;
  return true;
}

}  // namespace

std::vector<FastOp> lower_fast(uint32_t n, const std::vector<ExactGate>& gates, bool want_params) {
  std::vector<FastOp> out;
  out.reserve(gates.size());
  int pcount = 0;
  for (size_t gi = 0; gi < gates.size(); ++gi) {
    const ExactGate& e = gates[gi];
    std::vector<cd> A = e.M;
    std::vector<int> wires(e.bits);
    int K = e.K;
    uint64_t cmask = 0;
    // Peel control wires (CNOT, CY, CZ, CR, CRk, Toffoli, controlled custom tables).
    bool changed = true;
    while (changed && K > 1) {
      changed = false;
      for (int q = 0; q < K; ++q) {
        if (is_control_wire(A, K, q)) {
          cmask |= 1ull << wires[q];
          A = restrict_to_one(A, K, q);
          wires.erase(wires.begin() + q);
          --K;
          changed = true;
          break;
        }
      }
    }
    FastOp f;
    f.cmask = cmask;
    f.gate_index = (int)gi;
    f.k = K;
    for (int i = 0; i < K; ++i) f.bits[i] = wires[i];
    const int dim = 1 << K;
    const int nc = __builtin_popcountll(cmask);
    if (is_diagonal(A, K)) {
      f.kind = K == 1 ? FK_DIAG1 : FK_DIAGK;
      f.m.resize(dim);
      int nontrivial = 0;
      for (int r = 0; r < dim; ++r) {
        f.m[r] = A[r * dim + r];
        if (!is_one(f.m[r])) ++nontrivial;
      }
      f.f_touch = (double)nontrivial / dim / (double)(1ull << nc);
    } else {
      f.kind = K == 1 ? FK_DENSE1 : FK_DENSEK;
      f.m = A;
      f.f_touch = 1.0 / (double)(1ull << nc);
    }
    classify(f);
    if (want_params) {
      switch (e.kind) {
        case QR_GATE_RX: f.param = pcount++; f.gen = 1; break;
        case QR_GATE_RY: f.param = pcount++; f.gen = 2; break;
        case QR_GATE_RZ: f.param = pcount++; f.gen = 3; break;
        case QR_GATE_PHASE: f.param = pcount++; f.gen = 4; break;
        default: break;
      }
      if (f.param >= 0) {  // keep the parameter gate's bit even if the matrix is trivial
        f.k = 1;
        f.bits[0] = e.bits.back();
      }
    }
    out.push_back(std::move(f));
  }
  return out;
}

std::vector<FastOp> adjoint_sweep_ops(const std::vector<FastOp>& ops) {
  std::vector<FastOp> rev;
  rev.reserve(2 * ops.size());
  for (size_t i = ops.size(); i-- > 0;) {
    const FastOp& op = ops[i];
    if (op.param >= 0) {
      FastOp t;
      t.kind = FK_GRAD;
      t.k = 1;
      t.bits[0] = op.bits[0];
      t.gen = op.gen;
      t.param = op.param;
      t.f_touch = 0.0;
      rev.push_back(t);
    }
    rev.push_back(adjoint(op));
  }
  return rev;
}

void classify(FastOp& f) {
  f.form = U1_GENERAL;
  f.sub = SUB_NONE;
  if (f.kind == FK_DIAG1) {
    // d0 is factored out only when it is the larger entry: diag(1, d1/d0) stays finite and
    // bounded (a zero or tiny d0, e.g. a post-selection onto |1>, keeps its general form)
    if (!is_one(f.m[0]) && f.m[0] != cd(0, 0) && std::abs(f.m[0]) >= std::abs(f.m[1])) f.sub = SUB_D0;
    return;
  }
  if (f.kind != FK_DENSE1) return;
  const std::vector<cd>& A = f.m;
  if (is_zero(A[0]) && is_zero(A[3])) {
    f.form = U1_PERM;
    if (A[1] == A[2]) f.sub = SUB_XLIKE;
    else if (A[1] == -A[2]) f.sub = SUB_YLIKE;
  } else if (A[0].imag() == 0 && A[1].imag() == 0 && A[2].imag() == 0 && A[3].imag() == 0) {
    f.form = U1_REAL;
    if (A[0] == A[1] && A[0] == A[2] && A[3] == -A[0]) f.sub = SUB_HAD;
    else if (A[0] == A[3] && A[1] == -A[2]) f.sub = SUB_RROT;
  } else if (A[0].imag() == 0 && A[3].imag() == 0 && A[1].real() == 0 && A[2].real() == 0) {
    f.sub = (A[0] == A[3] && A[1] == A[2]) ? SUB_XROT : SUB_RX;
  }
}

FastOp adjoint(const FastOp& op) {
  FastOp a = op;
  if (op.kind == FK_DIAG1 || op.kind == FK_DIAGK) {
    for (auto& x : a.m) x = std::conj(x);
  } else {
    const int dim = 1 << op.k;
    for (int r = 0; r < dim; ++r)
      for (int j = 0; j < dim; ++j) a.m[r * dim + j] = std::conj(op.m[j * dim + r]);
  }
  classify(a);
  return a;
}

// ---- observables (Pauli sums) ---------------------------------------------------------

PauliMasks parse_term(uint32_t n, const qr_observable* o, uint64_t t) {
  const qr_pauli_term& pt = o->terms[t];
  if (pt.nfactors && (!o->wires || !o->paulis || pt.offset > o->nfactors || pt.nfactors > o->nfactors - pt.offset))
    throw Error{QR_ERR_ARG, "observable term " + std::to_string(t) + " is out of range"};
  PauliMasks tm{pt.coeff, 0, 0, 0};
  uint64_t seen = 0;
  for (uint32_t f = 0; f < pt.nfactors; ++f) {
    const uint32_t w = o->wires[pt.offset + f];
    if (w >= n) throw Error{QR_ERR_ARG, "observable wire " + std::to_string(w) + " is out of range"};
    const uint64_t b = 1ull << (n - 1 - w);
    if (seen & b) throw Error{QR_ERR_ARG, "observable term repeats wire " + std::to_string(w)};
    seen |= b;
    switch (o->paulis[pt.offset + f]) {
      case 'X': case 1: tm.xm |= b; break;
      case 'Y': case 2: tm.xm |= b; tm.zm |= b; tm.ny++; break;
      case 'Z': case 3: tm.zm |= b; break;
      case 'I': case 0: break;
      default: throw Error{QR_ERR_ARG, "observable pauli must be X, Y or Z"};
    }
  }
  return tm;
}

std::vector<PauliMasks> parse_observable(uint32_t n, const qr_observable* o) {
  if (!o) throw Error{QR_ERR_ARG, "observable is NULL"};
  if (o->nterms && !o->terms) throw Error{QR_ERR_ARG, "observable terms are NULL"};
  std::vector<PauliMasks> out;
  for (uint64_t t = 0; t < o->nterms; ++t) out.push_back(parse_term(n, o, t));
  return out;
}

PauliZTerm z_term(const PauliMasks& tm) {
  double cr = tm.coeff, ci = 0.0;
  for (int q = 0; q < (tm.ny & 3); ++q) {  // i^ny folded into the coefficient
    const double t = cr;
    cr = -ci;
    ci = t;
  }
  return PauliZTerm{tm.zm, cr, ci};
}

int pauli_passes(uint32_t n, int prec, PauliGroup* groups, int G, PauliPass* passes, uint32_t* lxm) {
  const int K = pauli_tile_bits(prec), c = pauli_low_bits(prec);
  const uint64_t low = (1ull << c) - 1;
  if (G <= 0) return 0;
  if ((int)n < K + 3) {  // small states: one direct pass
    passes[0] = PauliPass{0, 0, (uint32_t)G};
    for (int i = 0; i < G; ++i) lxm[i] = 0;
    return 1;
  }
  auto hi = [&](const PauliGroup& x) { return __builtin_popcountll(x.xm & ~low); };
  // groups whose X mask fits a tile first, widest first (then by mask): first-fit packs them
  std::sort(groups, groups + G, [&](const PauliGroup& a, const PauliGroup& b) {
    const bool wa = hi(a) > K - c, wb = hi(b) > K - c;
    if (wa != wb) return wb;
    if (hi(a) != hi(b)) return hi(a) > hi(b);
    return a.xm < b.xm;
  });
  int narrow = 0;
  while (narrow < G && hi(groups[narrow]) <= K - c) ++narrow;
  int np = 0, done = 0;
  while (done < narrow) {
    uint64_t uni = groups[done].xm & ~low;
    int end = done + 1;
    for (int i = done + 1; i < narrow; ++i) {
      const uint64_t u2 = uni | (groups[i].xm & ~low);
      if (__builtin_popcountll(u2) <= K - c) {
        std::swap(groups[end], groups[i]);
        ++end;
        uni = u2;
      }
    }
    // tile bits: the low segment bits, the groups' X bits, then the lowest other bits
    uint64_t t = low | uni;
    for (uint32_t b = 0; b < n && __builtin_popcountll(t) < K; ++b) t |= 1ull << b;
    for (int i = done; i < end; ++i) {
      uint32_t lx = 0;
      uint64_t m = t;
      for (int b = 0; b < K; ++b) {
        const uint64_t bit = m & (0ull - m);
        m ^= bit;
        if (groups[i].xm & bit) lx |= 1u << b;
      }
      lxm[i] = lx;
    }
    passes[np++] = PauliPass{t, (uint32_t)done, (uint32_t)(end - done)};
    done = end;
  }
  if (narrow < G) {
    for (int i = narrow; i < G; ++i) lxm[i] = 0;
    passes[np++] = PauliPass{0, (uint32_t)narrow, (uint32_t)(G - narrow)};
  }
  return np;
}

void pauli_tables(const std::vector<PauliMasks>& terms, std::vector<PauliGroup>& groups,
                  std::vector<PauliZTerm>& flat) {
  // terms grouped by X mask; Z masks and i^ny folded into per-term complex coefficients
  std::map<uint64_t, std::vector<PauliZTerm>> by_x;
  for (const auto& tm : terms) by_x[tm.xm].push_back(z_term(tm));
  groups.clear();
  flat.clear();
  for (const auto& g : by_x) {
    groups.push_back(PauliGroup{g.first, (uint32_t)flat.size(), (uint32_t)g.second.size()});
    flat.insert(flat.end(), g.second.begin(), g.second.end());
  }
}

}  // namespace qr
