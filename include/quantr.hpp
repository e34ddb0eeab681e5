// quantr.hpp — C++ mirror of quantr's public Rust API (quantr 0.6.0, src/lib.rs:73-85) on top
// of the C ABI in quantr_hip.h. Names, argument meaning, validation order and error messages
// follow the Rust crate; a Result<_, QuantrError> becomes a thrown quantr::QuantrError.
//
//   Rust                                         C++
//   Circuit::new(n)?                             quantr::Circuit::new_(n)
//   circuit.add_gate(Gate::H, 2)?                circuit.add_gate(Gate::H(), 2)
//   circuit.simulate()                           std::move(circuit).simulate()
//   simulated.measure_all(500)                   simulated.measure_all(500)
//   Gate::Custom(f, vec![0, 1], "QFT".into())    Gate::Custom(f, {0, 1}, "QFT")
//
// Header-only; link with libquantr_hip.so.
#pragma once

#include <algorithm>
#include <charconv>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "quantr_hip.h"

namespace quantr {

using Complex64 = std::complex<double>;  // num_complex::Complex64

// ---- error.rs:17-33 ----------------------------------------------------------------------
class QuantrError : public std::runtime_error {
 public:
  explicit QuantrError(std::string message) : std::runtime_error(message), message_(std::move(message)) {}
  const std::string& message() const { return message_; }
  // Display form of the reference (ANSI red "[Quantr Error] ...").
  std::string display() const { return "\x1b[91m[Quantr Error] " + message_ + "\x1b[0m "; }

 private:
  std::string message_;
};

namespace detail {
inline void check(qr_status s) {
  if (s != QR_OK) throw QuantrError(qr_last_error());
}
// Rust's Display of an f64: the shortest decimal that round-trips, never in exponent form
// (2.0 -> "2", 0.1 -> "0.1").
inline std::string display_f64(double x) {
  char buf[512];
  const auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::fixed);
  return std::string(buf, r.ptr);
}
// Rust's Debug of an f64 as it appears in gate names: Display plus ".0" when integral.
inline std::string debug_f64(double x) {
  std::string s = display_f64(x);
  if (s.find_first_of(".nN") == std::string::npos) s += ".0";
  return s;
}
inline bool is_ascii(const std::string& s) {
  for (unsigned char c : s)
    if (c >= 0x80) return false;
  return true;
}
// Thread-local draw source standing in for fastrand's thread-local generator: wyrand with the
// wyhash v4.2 constants, f64 in [0, 1) from the top 52 bits (fastrand 2.x's published
// construction). Whether its stream equals fastrand's is not checked anywhere.
struct WyRand {
  uint64_t s = 0x2d358dccaa6c78a5ull;
  uint64_t next_u64() {
    s += 0x2d358dccaa6c78a5ull;
    const unsigned __int128 t = (unsigned __int128)s * (unsigned __int128)(s ^ 0x8bb84b93962eacc9ull);
    return (uint64_t)t ^ (uint64_t)(t >> 64);
  }
  double f64() {
    const uint64_t bits = (uint64_t(1) << 62) - (uint64_t(1) << 52) + (next_u64() >> 12);
    double d;
    static_assert(sizeof(d) == sizeof(bits), "f64");
    std::memcpy(&d, &bits, sizeof d);
    return d - 1.0;
  }
};
inline WyRand& rng() {
  thread_local WyRand r;
  return r;
}
}  // namespace detail

// fastrand::seed (used by the reference's tests, tests/qft.rs:23).
inline void seed(uint64_t s) { detail::rng().s = s; }

namespace states {

// ---- qubit.rs:15-19 -------------------------------------------------------------------
enum class Qubit { Zero = 0, One = 1 };

// ---- product_states.rs:17-216 ----------------------------------------------------------
class ProductState {
 public:
  ProductState() = default;
  static ProductState new_(const std::vector<Qubit>& q) {
    if (q.empty()) throw QuantrError("The slice of qubits is empty, it needs to at least have one element.");
    return ProductState(q);
  }
  explicit ProductState(std::vector<Qubit> q) : qubits_(std::move(q)) {}
  explicit ProductState(Qubit q) : qubits_{q} {}
  std::optional<Qubit> get(size_t i) const {
    if (i < qubits_.size()) return qubits_[i];
    return std::nullopt;
  }
  const std::vector<Qubit>& get_qubits() const { return qubits_; }
  std::vector<Qubit>& get_mut_qubits() { return qubits_; }
  size_t num_qubits() const { return qubits_.size(); }
  ProductState& invert_digit(size_t place) {
    if (place >= num_qubits()) {
      std::ostringstream os;
      os << "The position of the binary digit, " << place << ", is out of bounds. The product dimension is "
         << num_qubits() << ", and so the position must be strictly less.";
      throw QuantrError(os.str());
    }
    qubits_[place] = qubits_[place] == Qubit::Zero ? Qubit::One : Qubit::Zero;
    return *this;
  }
  ProductState kronecker_prod(Qubit other) const {
    ProductState p = *this;
    p.qubits_.push_back(other);
    return p;
  }
  // Index in the computational basis, wire 0 most significant (product_states.rs:191-201;
  // 64-bit here).
  uint64_t comp_basis() const {
    uint64_t v = 0;
    for (Qubit q : qubits_) v = (v << 1) | (q == Qubit::One ? 1 : 0);
    return v;
  }
  static ProductState binary_basis(uint64_t index, size_t n) {  // product_states.rs:205-215
    std::vector<Qubit> q(n);
    for (size_t i = 0; i < n; ++i) q[i] = ((index >> (n - 1 - i)) & 1) ? Qubit::One : Qubit::Zero;
    return ProductState(std::move(q));
  }
  std::string to_string() const {
    std::string s;
    for (Qubit q : qubits_) s += q == Qubit::One ? '1' : '0';
    return s;
  }
  bool operator==(const ProductState& o) const { return qubits_ == o.qubits_; }
  bool operator<(const ProductState& o) const {
    if (qubits_.size() != o.qubits_.size()) return qubits_.size() < o.qubits_.size();
    return comp_basis() < o.comp_basis();
  }

 private:
  std::vector<Qubit> qubits_;
};

// ---- super_positions.rs:18-393 ----------------------------------------------------------
class SuperPosition {
 public:
  static constexpr double ZERO_MARGIN = 1e-6;
  SuperPosition() = default;
  static SuperPosition new_(size_t prod_dimension) {
    if (prod_dimension == 0) throw QuantrError("The number of qubits must be non-zero.");
    SuperPosition s;
    s.amplitudes_.assign(size_t(1) << prod_dimension, Complex64(0, 0));
    s.amplitudes_[0] = 1.0;
    s.product_dim_ = prod_dimension;
    return s;
  }
  static SuperPosition new_with_amplitudes(const std::vector<Complex64>& a) {
    if (!equal_within_error(norm_sqr_sum(a), 1.0))
      throw QuantrError(
          "Slice given to set amplitudes in super position does not conserve probability, the absolute square sum "
          "of the coefficents must be one.");
    const size_t len = a.size();
    if ((len & (len - 1)) != 0)
      throw QuantrError("The length of the array must be of the form 2**n where n is an integer.");
    return new_with_amplitudes_unchecked(a);
  }
  static SuperPosition new_with_amplitudes_unchecked(const std::vector<Complex64>& a) {
    SuperPosition s;
    s.amplitudes_ = a;
    size_t tz = 0;
    while (tz < 64 && a.size() && !((a.size() >> tz) & 1)) ++tz;
    s.product_dim_ = tz;
    return s;
  }
  // From<ProductState> (super_positions.rs:360-376)
  static SuperPosition from(const ProductState& p) {
    SuperPosition s;
    s.product_dim_ = p.num_qubits();
    s.amplitudes_.assign(size_t(1) << s.product_dim_, Complex64(0, 0));
    s.amplitudes_[p.comp_basis()] = 1.0;
    return s;
  }
  std::optional<Complex64> get_amplitude(size_t pos) const {
    if (pos < amplitudes_.size()) return amplitudes_[pos];
    return std::nullopt;
  }
  size_t get_num_qubits() const { return product_dim_; }
  size_t get_dimension() const { return amplitudes_.size(); }
  const std::vector<Complex64>& get_amplitudes() const { return amplitudes_; }
  Complex64 get_amplitude_from_state(const ProductState& p) const {
    if ((size_t(1) << p.num_qubits()) != amplitudes_.size()) {
      std::ostringstream os;
      os << "Unable to retreive product state, |\"" << p.to_string() << "\"> with dimension " << p.num_qubits()
         << ". The superposition is a linear combination of states with different dimension. These dimensions "
            "should be equal.";
      throw QuantrError(os.str());
    }
    return amplitudes_[p.comp_basis()];
  }
  SuperPosition& set_amplitudes(const std::vector<Complex64>& a) {
    if (a.size() != amplitudes_.size()) {
      std::ostringstream os;
      os << "The slice given to set the amplitudes in the computational basis has length " << a.size()
         << ", when it should have length " << amplitudes_.size() << ".";
      throw QuantrError(os.str());
    }
    if (!equal_within_error(norm_sqr_sum(a), 1.0))
      throw QuantrError(
          "Slice given to set amplitudes in super position does not conserve probability, the absolute square sum "
          "of the coefficents must be one.");
    amplitudes_ = a;
    return *this;
  }
  SuperPosition& set_amplitudes_unchecked(const std::vector<Complex64>& a) {
    amplitudes_ = a;
    return *this;
  }
  // super_positions.rs:105-131 (keys: product states of one dimension; missing states are 0).
  static SuperPosition new_with_hash_amplitudes(const std::map<ProductState, Complex64>& h) {
    if (h.empty())
      throw QuantrError("An empty HashMap was given. A superposition must have at least one non-zero state.");
    const size_t dim = h.begin()->first.num_qubits();
    double total = 0.0;
    for (const auto& kv : h) {
      if (kv.first.num_qubits() != dim) {
        std::ostringstream os;
        os << "The first state has product dimension of " << dim << ", whilst the state, |" << kv.first.to_string()
           << ">, found as a key in the HashMap has dimension " << kv.first.num_qubits() << ".";
        throw QuantrError(os.str());
      }
      total += std::norm(kv.second);
    }
    if (!equal_within_error(total, 1.0))
      throw QuantrError("The total sum of the absolute square of all amplitudes, " + detail::display_f64(total) +
                        ", does not equal 1. That is, the superpositon does not conserve probability.");
    SuperPosition s;
    s.product_dim_ = dim;
    s.amplitudes_.assign(size_t(1) << dim, Complex64(0, 0));
    for (const auto& kv : h) s.amplitudes_[kv.first.comp_basis()] = kv.second;
    return s;
  }
  // super_positions.rs:270-304
  SuperPosition& set_amplitudes_from_states(const std::map<ProductState, Complex64>& h) {
    if (h.empty())
      throw QuantrError("An empty HashMap was given. A superposition must have at least one non-zero state.");
    double total = 0.0;
    for (const auto& kv : h) {
      if (kv.first.num_qubits() != product_dim_) {
        std::ostringstream os;
        os << "The first state has product dimension of " << product_dim_ << ", whilst the state, |"
           << kv.first.to_string() << ">, found as a key in the HashMap has dimension " << kv.first.num_qubits()
           << ".";
        throw QuantrError(os.str());
      }
      total += std::norm(kv.second);
    }
    if (!equal_within_error(total, 1.0))
      throw QuantrError(
          "The total sum of the absolute square of all amplitudes does not equal 1. That is, the superpositon does "
          "not conserve probability.");
    std::fill(amplitudes_.begin(), amplitudes_.end(), Complex64(0, 0));
    for (const auto& kv : h) amplitudes_[kv.first.comp_basis()] = kv.second;
    return *this;
  }
  // From<Qubit> (super_positions.rs:378-392)
  static SuperPosition from(Qubit q) { return from(ProductState(q)); }
  // super_positions.rs:332-342 on the host: one draw from the thread-local generator, left fold
  // of |amplitude|^2, the first state whose running sum exceeds the draw (none if the total
  // probability is below it).
  std::optional<ProductState> measure() const;
  // Iteration (super_position_iter.rs:56-72): every basis state in index order, zeros included.
  class iterator {
   public:
    iterator(const SuperPosition* s, size_t i) : s_(s), i_(i) {}
    std::pair<ProductState, Complex64> operator*() const {
      return {ProductState::binary_basis(i_, s_->product_dim_), s_->amplitudes_[i_]};
    }
    iterator& operator++() {
      ++i_;
      return *this;
    }
    bool operator!=(const iterator& o) const { return i_ != o.i_; }

   private:
    const SuperPosition* s_;
    size_t i_;
  };
  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, amplitudes_.size()); }
  // Non-zero amplitudes keyed by basis state (super_positions.rs:315-323).
  std::map<ProductState, Complex64> to_hash_map() const {
    std::map<ProductState, Complex64> m;
    for (size_t i = 0; i < amplitudes_.size(); ++i)
      if (!equal_within_error(std::norm(amplitudes_[i]), 0.0))
        m.emplace(ProductState::binary_basis(i, product_dim_), amplitudes_[i]);
    return m;
  }
  bool operator==(const SuperPosition& o) const {
    return product_dim_ == o.product_dim_ && amplitudes_ == o.amplitudes_;
  }

 private:
  static bool equal_within_error(double x, double c) { return x < c + ZERO_MARGIN && x > c - ZERO_MARGIN; }
  static double norm_sqr_sum(const std::vector<Complex64>& a) {
    double s = 0.0;
    for (const auto& z : a) s += z.real() * z.real() + z.imag() * z.imag();
    return s;
  }
  std::vector<Complex64> amplitudes_;
  size_t product_dim_ = 0;
};

inline std::optional<ProductState> SuperPosition::measure() const {
  double cum = 0.0;
  const double u = detail::rng().f64();
  for (size_t i = 0; i < amplitudes_.size(); ++i) {
    cum += amplitudes_[i].real() * amplitudes_[i].real() + amplitudes_[i].imag() * amplitudes_[i].imag();
    if (u < cum) return ProductState::binary_basis(i, product_dim_);
  }
  return std::nullopt;
}

}  // namespace states

// ---- measurement.rs:16-28 ----------------------------------------------------------------
template <class T>
class Measurement {
 public:
  enum class Kind { Observable, NonObservable };
  static Measurement Observable(T v) { return Measurement(Kind::Observable, std::move(v)); }
  static Measurement NonObservable(T v) { return Measurement(Kind::NonObservable, std::move(v)); }
  Kind kind() const { return kind_; }
  bool is_observable() const { return kind_ == Kind::Observable; }
  T take() && { return std::move(value_); }
  const T& get() const { return value_; }

 private:
  Measurement(Kind k, T v) : kind_(k), value_(std::move(v)) {}
  Kind kind_;
  T value_;
};

// ---- gate.rs:14-239 --------------------------------------------------------------------------
using CustomFn = std::function<std::optional<states::SuperPosition>(const states::ProductState&)>;

class Gate {
 public:
  enum class Kind : uint32_t {
    Id = QR_GATE_ID, H = QR_GATE_H, X = QR_GATE_X, Y = QR_GATE_Y, Z = QR_GATE_Z, S = QR_GATE_S,
    Sdag = QR_GATE_SDAG, T = QR_GATE_T, Tdag = QR_GATE_TDAG, Rx = QR_GATE_RX, Ry = QR_GATE_RY,
    Rz = QR_GATE_RZ, X90 = QR_GATE_X90, Y90 = QR_GATE_Y90, MX90 = QR_GATE_MX90, MY90 = QR_GATE_MY90,
    Phase = QR_GATE_PHASE, CR = QR_GATE_CR, CRk = QR_GATE_CRK, CZ = QR_GATE_CZ, CY = QR_GATE_CY,
    CNot = QR_GATE_CNOT, Swap = QR_GATE_SWAP, Toffoli = QR_GATE_TOFFOLI, Custom = QR_GATE_CUSTOM
  };
  Gate() = default;
  static Gate Id() { return Gate(Kind::Id); }
  static Gate H() { return Gate(Kind::H); }
  static Gate X() { return Gate(Kind::X); }
  static Gate Y() { return Gate(Kind::Y); }
  static Gate Z() { return Gate(Kind::Z); }
  static Gate S() { return Gate(Kind::S); }
  static Gate Sdag() { return Gate(Kind::Sdag); }
  static Gate T() { return Gate(Kind::T); }
  static Gate Tdag() { return Gate(Kind::Tdag); }
  static Gate Rx(double a) { return Gate(Kind::Rx, a); }
  static Gate Ry(double a) { return Gate(Kind::Ry, a); }
  static Gate Rz(double a) { return Gate(Kind::Rz, a); }
  static Gate X90() { return Gate(Kind::X90); }
  static Gate Y90() { return Gate(Kind::Y90); }
  static Gate MX90() { return Gate(Kind::MX90); }
  static Gate MY90() { return Gate(Kind::MY90); }
  static Gate Phase(double a) { return Gate(Kind::Phase, a); }
  static Gate CR(double a, size_t c) { return Gate(Kind::CR, a, 0, {c}); }
  static Gate CRk(int32_t k, size_t c) { return Gate(Kind::CRk, 0.0, k, {c}); }
  static Gate CZ(size_t c) { return Gate(Kind::CZ, 0.0, 0, {c}); }
  static Gate CY(size_t c) { return Gate(Kind::CY, 0.0, 0, {c}); }
  static Gate CNot(size_t c) { return Gate(Kind::CNot, 0.0, 0, {c}); }
  static Gate Swap(size_t c) { return Gate(Kind::Swap, 0.0, 0, {c}); }
  static Gate Toffoli(size_t c1, size_t c2) { return Gate(Kind::Toffoli, 0.0, 0, {c1, c2}); }
  static Gate Custom(CustomFn f, std::vector<size_t> controls, std::string name) {
    Gate g(Kind::Custom, 0.0, 0, std::move(controls));
    g.fn_ = std::make_shared<CustomFn>(std::move(f));
    g.name_ = std::move(name);
    return g;
  }

  Kind kind() const { return kind_; }
  double angle() const { return angle_; }
  int32_t k() const { return k_; }
  const std::vector<size_t>& controls() const { return nodes_; }
  const std::string& custom_name() const { return name_; }
  const std::shared_ptr<CustomFn>& custom_fn() const { return fn_; }

  // gate.rs:110-138
  std::optional<std::vector<size_t>> get_nodes() const {
    if (is_single_gate()) return std::nullopt;
    return nodes_;
  }
  // gate.rs:173-201
  bool is_single_gate() const {
    switch (kind_) {
      case Kind::CNot: case Kind::Swap: case Kind::CZ: case Kind::CY: case Kind::CR: case Kind::CRk:
      case Kind::Toffoli: case Kind::Custom:
        return false;
      default:
        return true;
    }
  }
  bool is_custom_gate() const { return kind_ == Kind::Custom; }
  // gate.rs:210-238
  std::string get_name() const {
    switch (kind_) {
      case Kind::Id: return "";
      case Kind::X: return "X";
      case Kind::H: return "H";
      case Kind::S: return "S";
      case Kind::Sdag: return "S*";
      case Kind::T: return "T";
      case Kind::Tdag: return "T*";
      case Kind::Y: return "Y";
      case Kind::Z: return "Z";
      case Kind::Rx: return "Rx";
      case Kind::Ry: return "Ry";
      case Kind::Rz: return "Rz";
      case Kind::Phase: return "P";
      case Kind::X90: return "X90";
      case Kind::Y90: return "Y90";
      case Kind::MX90: return "X90*";
      case Kind::MY90: return "Y90*";
      case Kind::CR: return "CR";
      case Kind::CRk: return "CRk";
      case Kind::Swap: return "Sw";
      case Kind::CZ: return "Z";
      case Kind::CY: return "Y";
      case Kind::CNot: return "X";
      case Kind::Toffoli: return "X";
      case Kind::Custom: return name_;
    }
    return "";
  }
  // Rust Debug formatting, as used in QuantrError messages.
  std::string debug() const {
    static const char* names[] = {"Id", "H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx", "Ry", "Rz", "X90",
                                  "Y90", "MX90", "MY90", "Phase", "CR", "CRk", "CZ", "CY", "CNot", "Swap",
                                  "Toffoli", "Custom"};
    std::ostringstream os;
    const char* nm = names[(uint32_t)kind_];
    switch (kind_) {
      case Kind::Rx: case Kind::Ry: case Kind::Rz: case Kind::Phase: os << nm << "(" << fmt(angle_) << ")"; break;
      case Kind::CR: os << "CR(" << fmt(angle_) << ", " << nodes_[0] << ")"; break;
      case Kind::CRk: os << "CRk(" << k_ << ", " << nodes_[0] << ")"; break;
      case Kind::CZ: case Kind::CY: case Kind::CNot: case Kind::Swap: os << nm << "(" << nodes_[0] << ")"; break;
      case Kind::Toffoli: os << "Toffoli(" << nodes_[0] << ", " << nodes_[1] << ")"; break;
      case Kind::Custom: {
        os << "Custom(<fn>, [";
        for (size_t i = 0; i < nodes_.size(); ++i) os << (i ? ", " : "") << nodes_[i];
        os << "], \"" << name_ << "\")";
        break;
      }
      default: os << nm;
    }
    return os.str();
  }
  bool operator==(const Gate& o) const {
    return kind_ == o.kind_ && angle_ == o.angle_ && k_ == o.k_ && nodes_ == o.nodes_ && name_ == o.name_ &&
           fn_ == o.fn_;
  }
  bool operator!=(const Gate& o) const { return !(*this == o); }

 private:
  explicit Gate(Kind k, double a = 0.0, int32_t kk = 0, std::vector<size_t> nodes = {})
      : kind_(k), angle_(a), k_(kk), nodes_(std::move(nodes)) {}
  static std::string fmt(double x) { return detail::debug_f64(x); }
  Kind kind_ = Kind::Id;
  double angle_ = 0.0;
  int32_t k_ = 0;
  std::vector<size_t> nodes_;
  std::shared_ptr<CustomFn> fn_;
  std::string name_;
};

// A Pauli-sum observable for expectation values and gradients (no reference counterpart).
struct PauliTerm {
  double coeff;
  std::vector<std::pair<size_t, char>> factors;  // (wire, 'X' | 'Y' | 'Z')
};
using Observable = std::vector<PauliTerm>;

namespace detail {

// Owns the arrays behind a qr_program.
struct Program {
  std::vector<qr_op> ops;
  std::vector<uint32_t> controls;
  std::vector<double> images;
  std::vector<uint8_t> some;
  qr_program view() const {
    return qr_program{ops.data(), ops.size(), controls.data(), controls.size(),
                      images.data(), images.size(), some.data(), some.size()};
  }
  bool operator==(const Program& o) const {
    return ops.size() == o.ops.size() &&
           (ops.empty() || std::memcmp(ops.data(), o.ops.data(), ops.size() * sizeof(qr_op)) == 0) &&
           controls == o.controls && images == o.images && some == o.some;
  }
};

// Device state owned through a shared pointer (SimulatedCircuit copies share it).
struct DeviceState {
  qr_state* st = nullptr;
  explicit DeviceState(uint32_t n, qr_precision p) { check(qr_state_create(n, p, &st)); }
  ~DeviceState() {
    if (st) qr_state_destroy(st);
  }
  DeviceState(const DeviceState&) = delete;
  DeviceState& operator=(const DeviceState&) = delete;
};

struct ObservableArrays {
  std::vector<qr_pauli_term> terms;
  std::vector<uint32_t> wires;
  std::vector<uint8_t> paulis;
  explicit ObservableArrays(const Observable& o) {
    for (const auto& t : o) {
      terms.push_back(qr_pauli_term{t.coeff, (uint32_t)t.factors.size(), 0, (uint64_t)wires.size()});
      for (const auto& f : t.factors) {
        wires.push_back((uint32_t)f.first);
        paulis.push_back((uint8_t)f.second);
      }
    }
  }
  qr_observable view() const {
    return qr_observable{terms.data(), terms.size(), wires.data(), paulis.data(), wires.size()};
  }
};

}  // namespace detail

class Circuit;

// ---- simulated_circuit.rs:19-188 ---------------------------------------------------------------
class SimulatedCircuit {
 public:
  using Bins = std::map<states::ProductState, size_t>;

  // measure_all (simulated_circuit.rs:63-73) with draws from the thread-local generator.
  Measurement<Bins> measure_all(size_t shots) const {
    std::vector<double> u(shots);
    for (auto& x : u) x = detail::rng().f64();
    return measure_all_with_draws(u);
  }
  // Same, with the uniform draws injected (one per shot) — the parity-testable form.
  Measurement<Bins> measure_all_with_draws(const std::vector<double>& uniforms) const {
    if (std::any_of(gates_.begin(), gates_.end(), [](const Gate& g) { return g.is_custom_gate(); }) &&
        !disable_warnings_)
      std::fprintf(stderr, "%s\n",
                   "\x1b[93m[Quantr Warning] Custom gates were detected in the circuit. Measurements will be taken "
                   "from a cached register in memory, and so if the Custom gate does NOT implement a unitary "
                   "mapping, the measure_all method will most likely lead to wrong results. To simulate a circuit "
                   "without cache, see SimulatedCircuit::measure_all_without_cache.\x1b[0m");
    Bins bins;
    bin_draws(uniforms, bins);
    return Measurement<Bins>::Observable(std::move(bins));
  }
  // measure_all_without_cache (simulated_circuit.rs:81-114): the first shot from the current
  // register, every later shot from the circuit re-simulated on |0...0> with its custom
  // functions called again (2^k calls per custom gate and shot, where the reference makes 2^n),
  // all re-simulations batched on the device. The simulated state is left as it was.
  Measurement<Bins> measure_all_without_cache(size_t shots);

  Measurement<const states::SuperPosition*> get_state() const {
    download();
    return Measurement<const states::SuperPosition*>::NonObservable(&*host_);
  }
  Measurement<states::SuperPosition> take_state() && {
    download();
    return Measurement<states::SuperPosition>::NonObservable(std::move(*host_));
  }
  // simulated_circuit.rs:163-165 assigns `printing` to disable_warnings; kept as is.
  void print_warnings(bool printing) { disable_warnings_ = printing; }
  const std::vector<Gate>& get_circuit_gates() const { return gates_; }
  size_t get_num_qubits() const { return num_qubits_; }
  void set_print_progress(bool printing) { config_progress_ = printing; }

  // Extensions (no reference counterpart).
  double expectation(const Observable& o) const {
    detail::ObservableArrays arr(o);
    qr_observable v = arr.view();
    double e = 0.0;
    detail::check(qr_expectation(dev_->st, &v, &e));
    return e;
  }
  std::vector<uint64_t> sample_indices(const std::vector<double>& uniforms) const {
    std::vector<uint64_t> idx(uniforms.size());
    if (!uniforms.empty()) detail::check(qr_sample(dev_->st, uniforms.data(), uniforms.size(), idx.data()));
    return idx;
  }
  qr_state* device_state() const { return dev_->st; }

 private:
  friend class Circuit;
  SimulatedCircuit(std::vector<Gate> g, size_t n, std::shared_ptr<detail::DeviceState> d, bool progress,
                   std::shared_ptr<detail::Program> prog, qr_precision prec)
      : gates_(std::move(g)), num_qubits_(n), dev_(std::move(d)), config_progress_(progress),
        program_(std::move(prog)), precision_(prec) {}
  void download() const {
    if (host_) return;
    std::vector<Complex64> a(size_t(1) << num_qubits_);
    detail::check(qr_state_get_amplitudes(dev_->st, reinterpret_cast<double*>(a.data()), a.size()));
    host_ = states::SuperPosition::new_with_amplitudes_unchecked(a);
  }
  void bin_draws(const std::vector<double>& u, Bins& bins) const { bin_indices(sample_indices(u), bins); }
  void bin_indices(const std::vector<uint64_t>& idx, Bins& bins) const {
    for (uint64_t i : idx) {  // add_to_bin (simulated_circuit.rs:116-130)
      if (i == UINT64_MAX) {
        if (!disable_warnings_)
          std::fprintf(stderr, "%s\n",
                       "\x1b[93m[Quantr Warning] The superposition failed to collapse to a state during repeat "
                       "measurements. This is likely due to the use of Gate::Custom where the mapping is not "
                       "unitary.\x1b[0m");
        continue;
      }
      bins[states::ProductState::binary_basis(i, num_qubits_)] += 1;
    }
  }

  std::vector<Gate> gates_;
  size_t num_qubits_;
  std::shared_ptr<detail::DeviceState> dev_;
  bool config_progress_ = false;
  bool disable_warnings_ = false;
  std::shared_ptr<detail::Program> program_;
  qr_precision precision_ = QR_C128;
  mutable std::optional<states::SuperPosition> host_;
};

// ---- circuit.rs:26-474 ----------------------------------------------------------------------
class Circuit {
 public:
  static Circuit new_(size_t num_qubits) {
    if (num_qubits == 0) throw QuantrError("The initialised circuit must have at least one wire.");
    return Circuit(num_qubits);
  }
  size_t get_num_qubits() const { return num_qubits_; }
  void set_print_progress(bool progress) { config_progress_ = progress; }
  const std::vector<Gate>& get_gates() const { return gates_; }
  // Precision of the device state (complex128 unless changed; no reference counterpart).
  Circuit& set_precision(qr_precision p) {
    precision_ = p;
    return *this;
  }
  Circuit& set_options(const qr_options& o) {
    options_ = o;
    have_options_ = true;
    return *this;
  }

  Circuit& add_gate(Gate gate, size_t position) {
    std::vector<std::pair<size_t, Gate>> m{{position, std::move(gate)}};
    return add_gates_with_positions(m);
  }

  // circuit.rs:152-188. The Rust argument is a HashMap<usize, Gate>; pairs with distinct keys here.
  Circuit& add_gates_with_positions(const std::vector<std::pair<size_t, Gate>>& gates_with_positions) {
    for (const auto& kv : gates_with_positions)
      if (kv.first >= num_qubits_) {
        std::ostringstream os;
        os << "The position, " << kv.first << ", is out of bounds for the circuit with " << num_qubits_
           << " qubits.";
        throw QuantrError(os.str());
      }
    std::vector<Gate> column(num_qubits_, Gate::Id());
    for (const auto& kv : gates_with_positions) column[kv.first] = kv.second;
    has_overlapping_controls_and_target(column, num_qubits_);
    push_multi_gates(column);
    gates_.insert(gates_.end(), column.begin(), column.end());
    return *this;
  }

  // circuit.rs:210-226
  Circuit& add_gates(const std::vector<Gate>& gates) {
    if (gates.size() != num_qubits_) {
      std::ostringstream os;
      os << "The number of gates, " << gates.size() << ", does not match the number of wires, " << num_qubits_
         << ". All wires must have gates added.";
      throw QuantrError(os.str());
    }
    has_overlapping_controls_and_target(gates, num_qubits_);
    std::vector<Gate> v = gates;
    push_multi_gates(v);
    gates_.insert(gates_.end(), v.begin(), v.end());
    return *this;
  }

  // circuit.rs:324-341
  Circuit& add_repeating_gate(const Gate& gate, const std::vector<size_t>& positions) {
    for (size_t p : positions)
      if (p >= num_qubits_) {  // the reference indexes out of bounds here (a panic)
        std::ostringstream os;
        os << "The position, " << p << ", is out of bounds for the circuit with " << num_qubits_ << " qubits.";
        throw QuantrError(os.str());
      }
    if (contains_repeating_values(num_qubits_, positions)) {
      std::ostringstream os;
      os << "Attempted to add more than one gate onto a single wire. The positions in [";
      for (size_t i = 0; i < positions.size(); ++i) os << (i ? ", " : "") << positions[i];
      os << "] must all differ.";
      throw QuantrError(os.str());
    }
    std::vector<Gate> gates(num_qubits_, Gate::Id());
    for (size_t p : positions) gates[p] = gate;
    return add_gates(gates);
  }

  // circuit.rs:463-473
  Circuit& change_register(states::SuperPosition super_pos) {
    if (super_pos.get_num_qubits() != num_qubits_) {
      std::ostringstream os;
      os << "The custom register has a product state dimension of " << super_pos.get_num_qubits()
         << ", while the number of qubits is " << num_qubits_ << ". These must equal each other.";
      throw QuantrError(os.str());
    }
    register_ = std::move(super_pos);
    return *this;
  }

  // circuit.rs:364-388 (consumes the circuit) and :411-435 (keeps it).
  SimulatedCircuit simulate() && { return run(std::move(gates_), std::move(register_)); }
  SimulatedCircuit clone_and_simulate() const { return run(gates_, register_); }

  // Id-free gate program of this circuit (custom gates tabulated), as handed to the C ABI.
  std::shared_ptr<detail::Program> lower() const { return lower_gates(gates_, num_qubits_); }

  static std::shared_ptr<detail::Program> lower_gates(const std::vector<Gate>& gates, size_t n) {
    auto p = std::make_shared<detail::Program>();
    size_t counter = 0;
    for (const Gate& g : gates) {
      const size_t pos = counter % n;  // simulation.rs:43
      ++counter;
      if (g.kind() == Gate::Kind::Id) continue;
      qr_op op{};
      op.kind = (uint32_t)g.kind();
      op.position = (uint32_t)pos;
      op.ncontrols = (uint32_t)g.controls().size();
      op.controls_offset = p->controls.size();
      for (size_t c : g.controls()) p->controls.push_back((uint32_t)c);
      op.arg = g.angle();
      op.arg_int = g.k();
      if (g.is_custom_gate()) tabulate(g, p.get(), op);
      p->ops.push_back(op);
    }
    return p;
  }

 private:
  explicit Circuit(size_t n) : num_qubits_(n) {}

  // Calls the user function once per local basis state [controls..., position]
  // (simulation.rs:137-156) and stores Some/None and the images.
  static void tabulate(const Gate& g, detail::Program* p, qr_op& op) {
    const size_t k = g.controls().size() + 1;
    if (k > 16) throw QuantrError("custom gates act on at most 16 wires");
    const size_t dim = size_t(1) << k;
    op.image_offset = p->images.size();
    op.some_offset = p->some.size();
    p->images.resize(p->images.size() + dim * dim * 2, 0.0);
    p->some.resize(p->some.size() + dim, 0);
    for (size_t j = 0; j < dim; ++j) {
      const auto ps = states::ProductState::binary_basis(j, k);
      std::optional<states::SuperPosition> img = (*g.custom_fn())(ps);
      if (!img) continue;
      const auto& a = img->get_amplitudes();
      if (a.size() != dim) {
        std::ostringstream os;
        os << "The custom gate " << g.debug() << " returned a superposition of " << a.size()
           << " amplitudes for the product state |" << ps.to_string() << ">, it must have " << dim << ".";
        throw QuantrError(os.str());
      }
      p->some[op.some_offset + j] = 1;
      for (size_t r = 0; r < dim; ++r) {
        p->images[op.image_offset + (j * dim + r) * 2] = a[r].real();
        p->images[op.image_offset + (j * dim + r) * 2 + 1] = a[r].imag();
      }
    }
  }

  SimulatedCircuit run(std::vector<Gate> gates, std::optional<states::SuperPosition> reg) const {
    auto prog = lower_gates(gates, num_qubits_);
    auto dev = std::make_shared<detail::DeviceState>((uint32_t)num_qubits_, precision_);
    if (have_options_) detail::check(qr_state_set_options(dev->st, &options_));
    if (reg) {
      const auto& a = reg->get_amplitudes();
      detail::check(qr_state_set_amplitudes(dev->st, reinterpret_cast<const double*>(a.data()), a.size()));
    }
    if (config_progress_) std::printf("Starting circuit simulation...\n");
    qr_program v = prog->view();
    detail::check(qr_apply_program(dev->st, &v));
    if (config_progress_) {
      size_t counter = 0;
      for (const Gate& g : gates) {  // print_circuit_log (simulation.rs:183-200)
        if (g.kind() != Gate::Kind::Id)
          std::printf("Applying %s on wire %zu # %zu/%zu \n", g.debug().c_str(), counter % num_qubits_,
                      counter + 1, gates.size());
        if (counter + 1 == gates.size()) std::printf("Finished circuit simulation.\n");
        ++counter;
      }
    }
    return SimulatedCircuit(std::move(gates), num_qubits_, std::move(dev), config_progress_, std::move(prog),
                            precision_);
  }

  // circuit.rs:230-270
  static void push_multi_gates(std::vector<Gate>& gates) {
    std::vector<Gate> extended;
    bool found_multi = false, found_second = false;
    for (const Gate& g : gates) {
      if (g.is_custom_gate() && !detail::is_ascii(g.custom_name()))
        throw QuantrError("The custom function name, " + g.custom_name() +
                          ", does not only use ASCII chars. This could lead to problems in printing the circuit "
                          "diagram. This warning will be promoted to an Error in the next major release.");
      if (g.kind() != Gate::Kind::Id) {
        if (found_multi) {
          found_second = true;
          break;
        }
        found_multi = true;
      }
    }
    if (!found_second) return;
    std::vector<size_t> multi_pos;
    for (size_t pos = 0; pos < gates.size(); ++pos)
      if (!gates[pos].is_single_gate()) {
        std::vector<Gate> col(gates.size(), Gate::Id());
        col[pos] = gates[pos];
        extended.insert(extended.end(), col.begin(), col.end());
        multi_pos.push_back(pos);
      }
    for (size_t i : multi_pos) gates[i] = Gate::Id();
    gates.insert(gates.end(), extended.begin(), extended.end());
  }

  // circuit.rs:272-293
  static void has_overlapping_controls_and_target(const std::vector<Gate>& gates, size_t size) {
    for (size_t pos = 0; pos < gates.size(); ++pos) {
      auto nodes = gates[pos].get_nodes();
      if (!nodes) continue;
      for (size_t node : *nodes)  // the reference indexes out of bounds first (a panic)
        if (node >= size) {
          std::ostringstream os;
          os << "The control node at position " << node << ", is greater than the umnber of qubits " << size
             << ".";
          throw QuantrError(os.str());
        }
      if (contains_repeating_values(size, *nodes))
        throw QuantrError("The gate, " + gates[pos].debug() + ", has overlapping control nodes.");
      if (std::find(nodes->begin(), nodes->end(), pos) != nodes->end()) {
        std::ostringstream os;
        os << "The gate, " << gates[pos].debug() << ", has a control node that equals the gate's position " << pos
           << ".";
        throw QuantrError(os.str());
      }
    }
  }

  // circuit.rs:296-305
  static bool contains_repeating_values(size_t n, const std::vector<size_t>& a) {
    std::vector<bool> seen(n, false);
    for (size_t j : a) {
      if (j < n && seen[j]) return true;
      if (j < n) seen[j] = true;
    }
    return false;
  }

  std::vector<Gate> gates_;
  size_t num_qubits_;
  std::optional<states::SuperPosition> register_;
  bool config_progress_ = false;
  qr_precision precision_ = QR_C128;
  qr_options options_{1, 0, 0, 2, 0, 0, 2, 0, 0};
  bool have_options_ = false;
};

inline Measurement<SimulatedCircuit::Bins> SimulatedCircuit::measure_all_without_cache(size_t shots) {
  Bins bins;
  if (shots == 0) return Measurement<Bins>::Observable(std::move(bins));
  bin_draws({detail::rng().f64()}, bins);
  if (config_progress_) std::printf("Measured state # 1/%zu\n", shots);
  // Every later shot re-simulates the circuit on |0...0>, calling the custom functions again
  // (simulated_circuit.rs:91-111). The shots' programs are tabulated here in shot order (one
  // program per run of equal consecutive tables) and simulated and measured together on the
  // device (qr_sample_programs): programs of one structure run each sweep as one launch.
  constexpr size_t kPrograms = 1024;  // programs per device call
  std::vector<std::shared_ptr<detail::Program>> progs;
  std::vector<uint64_t> counts;
  std::vector<double> draws;
  auto flush = [&]() {
    if (progs.empty()) return;
    std::vector<qr_program> views;
    views.reserve(progs.size());
    for (const auto& p : progs) views.push_back(p->view());
    std::vector<uint64_t> idx(draws.size());
    detail::check(qr_sample_programs((uint32_t)num_qubits_, precision_, views.data(), views.size(), counts.data(),
                                     draws.data(), idx.data()));
    bin_indices(idx, bins);
    progs.clear();
    counts.clear();
    draws.clear();
  };
  for (size_t i = 0; i + 1 < shots; ++i) {
    auto prog = Circuit::lower_gates(gates_, num_qubits_);
    if (progs.empty() || !(*prog == *progs.back())) {
      if (progs.size() >= kPrograms) flush();
      progs.push_back(std::move(prog));
      counts.push_back(0);
    }
    ++counts.back();
    if (config_progress_) std::printf("Register reset to zero state\n");
    draws.push_back(detail::rng().f64());
    if (config_progress_) std::printf("Measured state # %zu/%zu\n", i + 2, shots);
  }
  flush();
  return Measurement<Bins>::Observable(std::move(bins));
}

// ---- printer.rs:17-377: text diagram of a circuit ----------------------------------------------
// One block of four text rows per wire (box top, wire with the gate name, box bottom,
// connection row) and one block column per gate column; a column holding a multi-wire gate
// draws its controls (█), crossings (┼) and the connecting line (│).
class Printer {
 public:
  explicit Printer(const Circuit& c) : circuit_(&c) {}
  void print_diagram() {
    if (columns() > 14 && !disable_warnings_)
      std::fprintf(stderr, "%s\n",
                   "\x1b[93m[Quantr Warning] The string displaying the circuit diagram exceeds 72 chars, which could "
                   "cause the circuit to render incorrectly in terminals (due to the wrapping). Instead, consider "
                   "saving the string to a .txt file by using Printer::save_diagram.\x1b[0m");
    std::printf("%s\n", get_diagram().c_str());
  }
  bool save_diagram(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const std::string d = get_diagram();
    const bool ok = std::fwrite(d.data(), 1, d.size(), f) == d.size();
    return std::fclose(f) == 0 && ok;
  }
  bool print_and_save_diagram(const std::string& path) {
    std::printf("%s\n", get_diagram().c_str());
    return save_diagram(path);
  }
  std::string get_diagram() {
    if (!diagram_) diagram_ = make_diagram();
    return *diagram_;
  }
  void set_warnings(bool printing) { disable_warnings_ = printing; }

 private:
  struct Rows {
    std::string top, name, bottom, link;
  };
  static std::string rep(const char* s, size_t n) {
    std::string out;
    for (size_t i = 0; i < n; ++i) out += s;
    return out;
  }
  size_t columns() const { return circuit_->get_gates().size() / circuit_->get_num_qubits(); }
  std::string make_diagram() const {
    const size_t n = circuit_->get_num_qubits();
    const auto& gates = circuit_->get_gates();
    std::vector<std::string> rows(4 * n + 1);
    auto put = [&](size_t wire, const Rows& r) {
      rows[4 * wire] += r.top;
      rows[4 * wire + 1] += r.name;
      rows[4 * wire + 2] += r.bottom;
      rows[4 * wire + 3] += r.link;
    };
    for (size_t col = 0; col < columns(); ++col) {
      const Gate* g = &gates[col * n];
      size_t widest = 1, multi = n;
      for (size_t w = 0; w < n; ++w) {
        widest = std::max(widest, g[w].get_name().size());
        if (multi == n && !g[w].is_single_gate()) multi = w;
      }
      if (multi < n) {
        draw_multi(put, g[multi], multi, n);
        continue;
      }
      for (size_t w = 0; w < n; ++w) {
        const std::string name = g[w].get_name();
        const size_t pad = widest - name.size();
        Rows r;
        if (g[w].kind() == Gate::Kind::Id) {
          r = {rep(" ", widest + 4), rep("\u2500", widest + 4), rep(" ", widest + 4), rep(" ", widest + 4)};
        } else {
          r.top = "\u250f\u2501" + rep("\u2501", name.size()) + "\u2501\u2513" + rep(" ", pad);
          r.name = "\u2528 " + name + " \u2520" + rep("\u2500", pad);
          r.bottom = "\u2517\u2501" + rep("\u2501", name.size()) + "\u2501\u251b" + rep(" ", pad);
          r.link = rep(" ", widest + 4);
        }
        put(w, r);
      }
    }
    std::string out;
    for (const auto& r : rows) out += r + "\n";
    return out;
  }
  template <class Put>
  static void draw_multi(Put& put, const Gate& g, size_t pos, size_t n) {
    std::vector<size_t> nodes = *g.get_nodes();
    nodes.push_back(pos);
    const size_t lo = *std::min_element(nodes.begin(), nodes.end());
    const size_t hi = *std::max_element(nodes.begin(), nodes.end());
    const std::string name = g.get_name();
    const size_t L = name.size();
    const std::string bar = "  \u2502  " + rep(" ", L - 1);
    for (size_t w = 0; w < n; ++w) {
      Rows r;
      if (w == pos) {
        r.top = std::string("\u250f\u2501") + (pos > lo ? "\u2537" : "\u2501") + rep("\u2501", L - 1) + "\u2501\u2513";
        r.name = "\u2528 " + name + " \u2520";
        r.bottom = std::string("\u2517\u2501") + (pos < hi ? "\u252f" : "\u2501") + rep("\u2501", L - 1) + "\u2501\u251b";
        r.link = std::string("  ") + (pos < hi ? "\u2502" : " ") + rep(" ", L + 1);
      } else if (w == lo) {
        r = {rep(" ", L + 4), "\u2500\u2500\u2588\u2500\u2500" + rep("\u2500", L - 1), bar, bar};
      } else if (w == hi) {
        r = {bar, "\u2500\u2500\u2588\u2500\u2500" + rep("\u2500", L - 1), rep(" ", L + 4), rep(" ", L + 4)};
      } else if (std::find(nodes.begin(), nodes.end(), w) != nodes.end()) {
        r = {bar, "\u2500\u2500\u2588\u2500\u2500" + rep("\u2500", L - 1), bar, bar};
      } else if (w > lo && w < hi) {
        r = {bar, "\u2500\u2500\u253c\u2500\u2500" + rep("\u2500", L - 1), bar, bar};
      } else {
        r = {rep(" ", L + 4), rep("\u2500", 5) + rep("\u2500", L - 1), " " + rep(" ", L + 3), " " + rep(" ", L + 3)};
      }
      put(w, r);
    }
  }

  const Circuit* circuit_;
  std::optional<std::string> diagram_;
  bool disable_warnings_ = false;
};

// ---- adjoint differentiation (no reference counterpart) -------------------------------------
// forward(): E = <psi|O|psi> for psi = U(theta)|0...0>; backward(g): g * dE/dtheta for every
// Rx/Ry/Rz/Phase gate, in circuit order, from one reverse sweep.
class AdjointFunction {
 public:
  AdjointFunction(size_t num_qubits, qr_precision p = QR_C128) {
    detail::check(qr_grad_create((uint32_t)num_qubits, p, &g_));
  }
  ~AdjointFunction() {
    if (g_) qr_grad_destroy(g_);
  }
  AdjointFunction(const AdjointFunction&) = delete;
  AdjointFunction& operator=(const AdjointFunction&) = delete;
  double forward(const Circuit& c, const Observable& o) {
    auto prog = c.lower();
    detail::ObservableArrays arr(o);
    qr_program pv = prog->view();
    qr_observable ov = arr.view();
    double e = 0.0;
    detail::check(qr_grad_forward(g_, &pv, &ov, &e));
    return e;
  }
  std::vector<double> backward(double grad_out = 1.0) {
    uint64_t n = 0;
    detail::check(qr_grad_num_params(g_, &n));
    std::vector<double> g(n);
    detail::check(qr_grad_backward(g_, grad_out, g.data(), n));
    return g;
  }

 private:
  qr_grad* g_ = nullptr;
};

}  // namespace quantr
