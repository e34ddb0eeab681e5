// api_mirror.cpp — the C++ mirror of quantr's API (include/quantr.hpp) exercised the way the
// reference's own tests use the Rust API. Test infrastructure only.
//
//   api_mirror cpu   builder, layout, error and host-state tests (no device calls)
//   api_mirror gpu   simulations through the engine on the GPU
//   api_mirror diagrams   the Printer's diagrams of the reference's printer tests
//
// Each test prints "ok <name>" or "FAIL <name>: <why>"; the exit status is the failure count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "quantr.hpp"

using quantr::Circuit;
using quantr::Complex64;
using quantr::Gate;
using quantr::QuantrError;
using quantr::states::ProductState;
using quantr::states::Qubit;
using quantr::states::SuperPosition;

namespace {

constexpr double kMargin = 1e-8;  // ERROR_MARGIN of tests/qft.rs:20 and tests/grovers.rs:22
const double kS = 0.70710678118654752440;

int g_fail = 0;

void check(bool ok, const std::string& name, const std::string& why) {
  if (!ok) throw std::runtime_error(why);
  (void)name;
}

void run(const char* name, const std::function<void()>& f) {
  try {
    f();
    std::printf("ok %s\n", name);
  } catch (const QuantrError& e) {
    std::printf("FAIL %s: QuantrError %s\n", name, e.message().c_str());
    ++g_fail;
  } catch (const std::exception& e) {
    std::printf("FAIL %s: %s\n", name, e.what());
    ++g_fail;
  }
  std::fflush(stdout);
}

// The call must throw a QuantrError whose message equals `want` (or contains it, if partial).
void expect_error(const std::function<void()>& f, const std::string& want, bool partial = false) {
  try {
    f();
  } catch (const QuantrError& e) {
    const bool ok = partial ? e.message().find(want) != std::string::npos : e.message() == want;
    if (!ok) throw std::runtime_error("message was: " + e.message());
    return;
  }
  throw std::runtime_error("no QuantrError raised, expected: " + want);
}

void expect_amplitudes(const SuperPosition& got, const std::vector<Complex64>& want) {
  const auto& a = got.get_amplitudes();
  if (a.size() != want.size()) throw std::runtime_error("dimension mismatch");
  for (size_t i = 0; i < a.size(); ++i)
    if (std::abs(a[i].real() - want[i].real()) > kMargin || std::abs(a[i].imag() - want[i].imag()) > kMargin) {
      char b[160];
      std::snprintf(b, sizeof b, "amplitude %zu: got (%.12g, %.12g) want (%.12g, %.12g)", i, a[i].real(), a[i].imag(),
                    want[i].real(), want[i].imag());
      throw std::runtime_error(b);
    }
}

std::vector<Gate::Kind> kinds(const std::vector<Gate>& g) {
  std::vector<Gate::Kind> k;
  for (const auto& x : g) k.push_back(x.kind());
  return k;
}

// ---- custom functions used by the reference's tests --------------------------------------

// circuit.rs:504-512: a CNOT that leaves |00>, |01> untouched (None).
std::optional<SuperPosition> example_cnot(const ProductState& p) {
  const Qubit a = p.get_qubits()[0], b = p.get_qubits()[1];
  if (a == Qubit::Zero) return std::nullopt;
  std::vector<Complex64> v(4, 0.0);
  v[b == Qubit::Zero ? 3 : 2] = 1.0;
  return SuperPosition::new_with_amplitudes(v);
}

// tests/grovers.rs:132-148: flips the last qubit when all others are one.
template <size_t N>
std::optional<SuperPosition> multicnot(const ProductState& in) {
  ProductState copy = in;
  std::vector<Qubit> all_one(N, Qubit::One);
  std::vector<Qubit> last_zero = all_one;
  last_zero[N - 1] = Qubit::Zero;
  if (copy.get_qubits() == all_one) {
    copy.get_mut_qubits()[N - 1] = Qubit::Zero;
    return SuperPosition::from(copy);
  }
  if (copy.get_qubits() == last_zero) {
    copy.get_mut_qubits()[N - 1] = Qubit::One;
    return SuperPosition::from(copy);
  }
  return std::nullopt;
}

// tests/qft.rs:51-67: the QFT of the input, computed by simulating a nested circuit.
std::optional<SuperPosition> qft(const ProductState& input) {
  const size_t n = input.num_qubits();
  Circuit mini = Circuit::new_(n);
  for (size_t pos = 0; pos < n; ++pos) {
    mini.add_gate(Gate::H(), pos);
    for (size_t k = 2; k <= n - pos; ++k) mini.add_gate(Gate::CRk((int32_t)k, pos + k - 1), pos);
  }
  mini.change_register(SuperPosition::from(input));
  return std::move(std::move(mini).simulate()).take_state().take();
}

// examples/post_select.rs:39-49: keeps |0> scaled by sqrt(2), drops |1> (non-unitary).
std::optional<SuperPosition> post_select(const ProductState& in) {
  if (in.get_qubits()[0] == Qubit::Zero)
    return SuperPosition::new_with_amplitudes_unchecked({Complex64(std::sqrt(2.0), 0.0), Complex64(0.0, 0.0)});
  return SuperPosition::new_with_amplitudes_unchecked({Complex64(0.0, 0.0), Complex64(0.0, 0.0)});
}

// ---- CPU tests ----------------------------------------------------------------------------

void cpu_tests() {
  using K = Gate::Kind;
  run("pushes_multi_gates", [] {  // circuit.rs:553-566
    Circuit qc = Circuit::new_(3);
    qc.add_gates({Gate::CNot(2), Gate::CNot(0), Gate::H()}).add_gates({Gate::Toffoli(1, 2), Gate::H(), Gate::CNot(0)});
    const std::vector<K> want{K::Id, K::Id, K::H, K::CNot, K::Id, K::Id, K::Id, K::CNot, K::Id,
                              K::Id, K::H, K::Id, K::Toffoli, K::Id, K::Id, K::Id, K::Id, K::CNot};
    check(kinds(qc.get_gates()) == want, "", "layout differs");
    check(qc.get_gates()[3] == Gate::CNot(2) && qc.get_gates()[7] == Gate::CNot(0) &&
              qc.get_gates()[12] == Gate::Toffoli(1, 2),
          "", "gate arguments differ");
  });
  run("pushes_multi_gates_using_vec", [] {  // circuit.rs:568-590
    Circuit qc = Circuit::new_(3);
    qc.add_gates_with_positions({{2, Gate::H()}, {0, Gate::CNot(2)}, {1, Gate::CNot(0)}})
        .add_gates_with_positions({{2, Gate::CNot(0)}, {0, Gate::Toffoli(1, 2)}, {1, Gate::H()}});
    const std::vector<K> want{K::Id, K::Id, K::H, K::CNot, K::Id, K::Id, K::Id, K::CNot, K::Id,
                              K::Id, K::H, K::Id, K::Toffoli, K::Id, K::Id, K::Id, K::Id, K::CNot};
    check(kinds(qc.get_gates()) == want, "", "layout differs");
  });
  run("single_multi_gate_column_kept", [] {  // circuit.rs:236-250: a lone multi gate stays
    Circuit qc = Circuit::new_(3);
    qc.add_gate(Gate::CNot(0), 2);
    check(kinds(qc.get_gates()) == std::vector<K>{K::Id, K::Id, K::CNot}, "", "layout differs");
  });
  run("catches_overlapping_control_nodes", [] {  // circuit.rs:526-532
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gates({Gate::CNot(0), Gate::Id(), Gate::Id()}); },
                 "The gate, CNot(0), has a control node that equals the gate's position 0.");
  });
  run("catches_overlapping_nodes_custom_gate", [] {  // circuit.rs:516-523
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gates({Gate::Id(), Gate::Custom(example_cnot, {1}, "X"), Gate::Id()}); },
                 "has a control node that equals the gate's position 1.", true);
  });
  run("catches_repeated_control_nodes", [] {
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gate(Gate::Toffoli(0, 0), 2); }, "The gate, Toffoli(0, 0), has overlapping control nodes.");
  });
  run("control_node_greater_than_circuit_size", [] {  // circuit.rs:603-612
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gates_with_positions({{2, Gate::H()}, {0, Gate::CNot(2)}, {1, Gate::CNot(3)}}); },
                 "The control node at position 3, is greater than the umnber of qubits 3.");
  });
  run("position_out_of_bounds", [] {
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gate(Gate::X(), 3); }, "The position, 3, is out of bounds for the circuit with 3 qubits.");
  });
  run("add_gates_wrong_length", [] {
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_gates({Gate::X(), Gate::H()}); },
                 "The number of gates, 2, does not match the number of wires, 3. All wires must have gates added.");
  });
  run("repeating_gate_repeated_position", [] {
    Circuit qc = Circuit::new_(3);
    expect_error([&] { qc.add_repeating_gate(Gate::X(), {0, 0}); },
                 "Attempted to add more than one gate onto a single wire. The positions in [0, 0] must all differ.");
  });
  run("zero_wire_circuit", [] {
    expect_error([] { Circuit::new_(0); }, "The initialised circuit must have at least one wire.");
  });
  run("custom_register_wrong_dimension", [] {  // circuit.rs:984-991
    Circuit qc = Circuit::new_(3);
    qc.add_gate(Gate::X(), 1);
    expect_error([&] { qc.change_register(SuperPosition::from(ProductState::new_({Qubit::One, Qubit::Zero}))); },
                 "The custom register has a product state dimension of 2, while the number of qubits is 3. These "
                 "must equal each other.");
  });
  run("custom_name_not_ascii", [] {
    Circuit qc = Circuit::new_(2);
    expect_error([&] { qc.add_gate(Gate::Custom(example_cnot, {0}, "\xce\xbb"), 1); },
                 "does not only use ASCII chars", true);
  });
  run("gate_names_and_debug", [] {  // gate.rs:210-238 names, Debug form in messages
    check(Gate::Sdag().get_name() == "S*" && Gate::MX90().get_name() == "X90*" && Gate::Swap(1).get_name() == "Sw" &&
              Gate::Phase(0.5).get_name() == "P" && Gate::CNot(0).get_name() == "X",
          "", "name");
    check(Gate::Rx(1.0).debug() == "Rx(1.0)" && Gate::CRk(3, 1).debug() == "CRk(3, 1)" &&
              Gate::Toffoli(0, 2).debug() == "Toffoli(0, 2)",
          "", "debug: " + Gate::Rx(1.0).debug());
    check(!Gate::CZ(1).is_single_gate() && Gate::Y90().is_single_gate() && !Gate::Id().get_nodes().has_value(),
          "", "single");
  });
  run("product_state_api", [] {  // product_states.rs:275-321
    const ProductState p = ProductState::new_({Qubit::One, Qubit::Zero, Qubit::One, Qubit::One});
    check(p.to_string() == "1011" && p.comp_basis() == 11, "", "basis");
    check(ProductState::binary_basis(6, 4).to_string() == "0110", "", "binary_basis");
    ProductState q = p;
    q.invert_digit(0);
    check(q.to_string() == "0011", "", "invert");
    expect_error([&] { q.invert_digit(4); },
                 "The position of the binary digit, 4, is out of bounds. The product dimension is 4, and so the "
                 "position must be strictly less.");
    check(p.kronecker_prod(Qubit::Zero).to_string() == "10110", "", "kron");
  });
  run("super_position_hash_api", [] {  // super_positions.rs:92-131, 255-304 (doc examples)
    const ProductState p01 = ProductState::new_({Qubit::Zero, Qubit::One});
    SuperPosition a = SuperPosition::new_with_hash_amplitudes({{p01, Complex64(1, 0)}});
    check(a.get_amplitudes() == std::vector<Complex64>{0, 1, 0, 0}, "", "new_with_hash_amplitudes");
    SuperPosition b = SuperPosition::new_(2);
    b.set_amplitudes_from_states({{p01, Complex64(1, 0)}});
    check(b.get_amplitudes() == std::vector<Complex64>{0, 1, 0, 0}, "", "set_amplitudes_from_states");
    expect_error([] { SuperPosition::new_with_hash_amplitudes({}); },
                 "An empty HashMap was given. A superposition must have at least one non-zero state.");
    expect_error([&] { SuperPosition::new_with_hash_amplitudes({{p01, Complex64(1, 1)}}); },
                 "The total sum of the absolute square of all amplitudes, 2, does not equal 1. That is, the "
                 "superpositon does not conserve probability.");
    expect_error([&] {
      SuperPosition::new_with_hash_amplitudes({{p01, Complex64(kS, 0)}, {ProductState::new_({Qubit::One}), Complex64(kS, 0)}});
    }, ">, found as a key in the HashMap has dimension ", true);  // which key is "first" follows map order
    expect_error([&] { b.set_amplitudes_from_states({{ProductState::new_({Qubit::One}), Complex64(1, 0)}}); },
                 "The first state has product dimension of 2, whilst the state, |1>, found as a key in the HashMap has "
                 "dimension 1.");
  });
  run("super_position_iteration_and_measure", [] {  // super_position_iter.rs:56-72, super_positions.rs:332-342
    SuperPosition s = SuperPosition::new_(2);
    s.set_amplitudes({Complex64(0, 0), Complex64(0, 0), Complex64(0, 1), Complex64(0, 0)});
    std::vector<std::string> seen;
    for (const auto& [state, amp] : s) seen.push_back(state.to_string() + (std::abs(amp) > 0 ? "*" : ""));
    check(seen == std::vector<std::string>{"00", "01", "10*", "11"}, "", "iteration order");
    quantr::seed(3);
    for (int i = 0; i < 20; ++i) check(s.measure()->to_string() == "10", "", "measure of a basis state");
    SuperPosition z = SuperPosition::new_with_amplitudes_unchecked({Complex64(0, 0), Complex64(0, 0)});
    check(!z.measure().has_value(), "", "measure with total probability 0");
    check(SuperPosition::from(Qubit::One).get_amplitudes() == std::vector<Complex64>{0, 1}, "", "from qubit");
  });
  run("gate_debug_floats", [] {  // Rust Debug of f64 in messages
    check(Gate::Rx(0.1).debug() == "Rx(0.1)" && Gate::Phase(2.0).debug() == "Phase(2.0)" &&
              Gate::CR(-1.5, 0).debug() == "CR(-1.5, 0)",
          "", Gate::Rx(0.1).debug());
  });
  run("super_position_api", [] {  // super_positions.rs
    expect_error([] { SuperPosition::new_with_amplitudes({Complex64(1, 0), Complex64(1, 0)}); },
                 "Slice given to set amplitudes in super position does not conserve probability, the absolute "
                 "square sum of the coefficents must be one.");
    expect_error([] { SuperPosition::new_with_amplitudes({Complex64(1, 0), Complex64(0, 0), Complex64(0, 0)}); },
                 "The length of the array must be of the form 2**n where n is an integer.");
    SuperPosition s = SuperPosition::new_(2);
    s.set_amplitudes({Complex64(0, 0), Complex64(kS, 0), Complex64(0, kS), Complex64(0, 0)});
    check(s.get_amplitude_from_state(ProductState::new_({Qubit::One, Qubit::Zero})) == Complex64(0, kS), "", "get");
    auto m = s.to_hash_map();
    check(m.size() == 2 && m.count(ProductState::new_({Qubit::Zero, Qubit::One})), "", "hash map");
  });
}

// ---- GPU tests ----------------------------------------------------------------------------

void gpu_tests() {
  run("simple_qft", [] {  // tests/qft.rs:22-47, custom gate simulating a nested circuit
    quantr::seed(0);
    Circuit qc = Circuit::new_(3);
    qc.add_repeating_gate(Gate::X(), {1, 2}).add_gate(Gate::Custom(qft, {0, 1}, "QFT"), 2);
    const double h = kS * 0.5;
    const std::vector<Complex64> want{{h, 0},       {-h, 0},       {0, -h},      {0, h},
                                      {-0.25, 0.25}, {0.25, -0.25}, {0.25, 0.25}, {-0.25, -0.25}};
    expect_amplitudes(std::move(std::move(qc).simulate()).take_state().take(), want);
  });
  run("grovers_3qubit", [] {  // tests/grovers.rs:24-70
    quantr::seed(0);
    Circuit c = Circuit::new_(3);
    c.add_repeating_gate(Gate::H(), {0, 1, 2});
    c.add_gate(Gate::CZ(1), 2);
    c.add_repeating_gate(Gate::H(), {0, 1, 2})
        .add_repeating_gate(Gate::X(), {0, 1, 2})
        .add_gate(Gate::H(), 2)
        .add_gate(Gate::Toffoli(0, 1), 2)
        .add_gate(Gate::H(), 2)
        .add_repeating_gate(Gate::X(), {0, 1, 2})
        .add_repeating_gate(Gate::H(), {0, 1, 2});
    auto sim = std::move(c).simulate();
    expect_amplitudes(*sim.get_state().get(), {0, 0, 0, -kS, 0, 0, 0, -kS});
    auto bins = sim.measure_all(500);
    check(bins.is_observable(), "", "observable");
    size_t total = 0;
    for (const auto& [state, count] : bins.get()) {
      total += count;
      if (state.to_string() == "011" || state.to_string() == "111") check(count > 200, "", "bin " + state.to_string());
      else check(count == 0, "", "stray bin " + state.to_string());
    }
    check(total == 500, "", "shot count");
  });
  run("x3sudoko", [] {  // tests/grovers.rs:72-130
    quantr::seed(0);
    Circuit qc = Circuit::new_(10);
    qc.add_repeating_gate(Gate::H(), {0, 1, 2, 3, 4, 5}).add_gate(Gate::X(), 8).add_gate(Gate::X(), 9).add_gate(Gate::H(), 9);
    auto oracle = [&qc]() {
      for (size_t i = 0; i <= 2; ++i) qc.add_gate(Gate::Toffoli(i, i + 3), 8);
      qc.add_gate(Gate::Custom(multicnot<4>, {0, 1, 2}, "X"), 6);
      for (size_t i = 0; i <= 2; ++i) qc.add_gate(Gate::CNot(i), 6);
      qc.add_gate(Gate::Custom(multicnot<4>, {3, 4, 5}, "X"), 7);
      for (size_t i = 3; i <= 5; ++i) qc.add_gate(Gate::CNot(i), 7);
    };
    oracle();
    qc.add_gate(Gate::Custom(multicnot<4>, {6, 7, 8}, "X"), 9);
    oracle();
    qc.add_repeating_gate(Gate::H(), {0, 1, 2, 3, 4, 5})
        .add_repeating_gate(Gate::X(), {0, 1, 2, 3, 4, 5})
        .add_gate(Gate::H(), 5)
        .add_gate(Gate::Custom(multicnot<6>, {0, 1, 2, 3, 4}, "X"), 5)
        .add_gate(Gate::H(), 5)
        .add_repeating_gate(Gate::X(), {0, 1, 2, 3, 4, 5})
        .add_repeating_gate(Gate::H(), {0, 1, 2, 3, 4, 5});
    auto sim = std::move(qc).simulate();
    sim.print_warnings(true);
    auto bins = sim.measure_all(5000);
    for (const auto& [state, count] : bins.get()) {
      const std::string six = state.to_string().substr(0, 6);
      const bool sol = six == "001100" || six == "001010" || six == "010100" || six == "010001" || six == "100010" ||
                       six == "100001";
      check(sol ? count > 150 : count < 150, "", "bin " + state.to_string());
    }
  });
  run("custom_gates", [] {  // circuit.rs:629-641
    Circuit qc = Circuit::new_(3);
    qc.add_gate(Gate::H(), 2).add_gate(Gate::Custom(example_cnot, {2}, "cNot"), 1);
    expect_amplitudes(std::move(std::move(qc).simulate()).take_state().take(), {kS, 0, 0, kS, 0, 0, 0, 0});
  });
  run("toffoli_gates", [] {  // circuit.rs:644-658
    Circuit qc = Circuit::new_(4);
    qc.add_gate(Gate::X(), 0).add_gate(Gate::H(), 3).add_gate(Gate::Y(), 3).add_gate(Gate::Toffoli(3, 0), 1);
    std::vector<Complex64> want(16, 0.0);
    want[8] = Complex64(0, -kS);
    want[13] = Complex64(0, kS);
    expect_amplitudes(std::move(std::move(qc).simulate()).take_state().take(), want);
  });
  run("custom_register", [] {  // circuit.rs:970-982
    Circuit c = Circuit::new_(3);
    c.add_gate(Gate::X(), 1).change_register(SuperPosition::from(ProductState::new_({Qubit::One, Qubit::Zero, Qubit::One})));
    expect_amplitudes(std::move(std::move(c).simulate()).take_state().take(), {0, 0, 0, 0, 0, 0, 0, 1});
  });
  run("clone_and_simulate_keeps_circuit", [] {  // circuit.rs:411-435
    Circuit c = Circuit::new_(2);
    c.add_gate(Gate::H(), 0).add_gate(Gate::CNot(0), 1);
    auto a = c.clone_and_simulate();
    auto b = c.clone_and_simulate();
    expect_amplitudes(*a.get_state().get(), {kS, 0, 0, kS});
    expect_amplitudes(*b.get_state().get(), {kS, 0, 0, kS});
    check(c.get_gates().size() == 4, "", "circuit kept");
  });
  run("post_select", [] {  // examples/post_select.rs: non-unitary custom gate
    Circuit qc = Circuit::new_(3);
    qc.add_repeating_gate(Gate::H(), {0, 1, 2}).add_gate(Gate::Custom(post_select, {}, "P"), 1);
    auto sim = std::move(qc).simulate();
    const double q = 0.5;  // (1/sqrt 8) * sqrt 2
    expect_amplitudes(*sim.get_state().get(), {q, q, 0, 0, q, q, 0, 0});
    sim.print_warnings(true);
    auto bins = sim.measure_all(400);
    size_t total = 0;
    for (const auto& [state, count] : bins.get()) {
      check(state.to_string()[1] == '0', "", "post-selected qubit measured as one");
      total += count;
    }
    check(total == 400, "", "shots");  // total probability is 1: no None shots
  });
  run("measure_all_without_cache", [] {  // simulated_circuit.rs:81-114
    quantr::seed(7);
    Circuit c = Circuit::new_(3);
    c.add_repeating_gate(Gate::H(), {0, 2}).add_gate(Gate::Custom(example_cnot, {0}, "X"), 1);
    auto sim = std::move(c).simulate();
    sim.print_warnings(true);
    auto bins = sim.measure_all_without_cache(600);
    size_t total = 0;
    for (const auto& [state, count] : bins.get()) {
      const std::string s = state.to_string();
      check(s == "000" || s == "001" || s == "110" || s == "111", "", "impossible state " + s);
      check(count > 90, "", "bin " + s + " too small");
      total += count;
    }
    check(total == 600, "", "shots");
  });
  run("adjoint_function", [] {  // extension: gradients of <Z0> for Ry(t) |0> = cos t
    Circuit c = Circuit::new_(2);
    c.add_gate(Gate::Ry(0.3), 0).add_gate(Gate::Rx(-0.7), 1);
    quantr::AdjointFunction f(2);
    const double e = f.forward(c, {{1.0, {{0, 'Z'}}}, {0.5, {{1, 'Z'}}}});
    check(std::abs(e - (std::cos(0.3) + 0.5 * std::cos(-0.7))) < 1e-12, "", "energy");
    const auto g = f.backward(2.0);
    check(g.size() == 2 && std::abs(g[0] + 2.0 * std::sin(0.3)) < 1e-12 &&
              std::abs(g[1] + 2.0 * 0.5 * std::sin(-0.7)) < 1e-12,
          "", "gradient");
  });
}

}  // namespace

// Diagrams of the reference's Printer tests (printer.rs:395-440), printed between markers for
// comparison with the fixture.
void diagrams() {
  {
    Circuit qc = Circuit::new_(4);
    qc.add_gate(Gate::H(), 3)
        .add_repeating_gate(Gate::Y(), {0, 1})
        .add_gate(Gate::Toffoli(0, 3), 1)
        .add_gate(Gate::CNot(1), 3)
        .add_gate(Gate::CNot(2), 0)
        .add_gate(Gate::CNot(2), 1);
    quantr::Printer p(qc);
    std::printf("<<<producing_string_circuit\n%s>>>\n", p.get_diagram().c_str());
  }
  {
    Circuit qc = Circuit::new_(4);
    qc.add_gate(Gate::H(), 3);
    qc.add_gates({Gate::H(), Gate::Custom(example_cnot, {3}, "Custom CNot"), Gate::Id(), Gate::X()})
        .add_repeating_gate(Gate::Y(), {0, 1})
        .add_gate(Gate::Toffoli(0, 3), 1)
        .add_gate(Gate::CNot(1), 3)
        .add_gate(Gate::CNot(2), 0)
        .add_gate(Gate::CNot(2), 1);
    quantr::Printer p(qc);
    std::printf("<<<producing_string_circuit_custom\n%s>>>\n", p.get_diagram().c_str());
  }
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  if (mode == "cpu") cpu_tests();
  else if (mode == "gpu") gpu_tests();
  else if (mode == "diagrams") diagrams();
  else {
    std::fprintf(stderr, "usage: api_mirror cpu|gpu\n");
    return 2;
  }
  return g_fail;
}
