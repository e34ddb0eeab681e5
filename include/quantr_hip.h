/*
 * quantr_hip.h — C ABI of the MI355X (gfx950) state-vector engine behind quantr's
 * Circuit / Gate / simulate() / measure_all() API.
 *
 * quantr (Rust crate, v0.6.0) has no FFI of its own: its public Rust API is the boundary.
 * Each entry point below names the reference item it replaces (file:line in the quantr
 * source tree). A Rust `-sys` crate binds these symbols one-to-one; the C++ mirror of the
 * Rust API in quantr.hpp is built on them too.
 *
 * Conventions
 *  - Wire 0 is the most significant bit of an amplitude index: wire w is bit (n-1-w)
 *    (product_states.rs:191-215).
 *  - Amplitudes cross the boundary as interleaved (re, im) binary64 pairs, index order.
 *  - Every function returns qr_status; nothing aborts or unwinds across the boundary.
 *    On failure qr_last_error() returns a thread-local message; messages that have a
 *    QuantrError counterpart use its text (error.rs:17-33, circuit.rs:157-305).
 *  - Calls on one qr_state are serialised by the caller; distinct states may be used from
 *    distinct threads. Host buffers are borrowed only for the duration of a call.
 */
#ifndef QUANTR_HIP_H
#define QUANTR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: qr_options gained stage_ops (a struct of 8 words; version 1 had 7).
 * 3: qr_plan_stats gained max_stage_gates and restore_sweeps.
 * 4: qr_options gained tune (9 words); qr_plan_stats gained tuning and xcd_order_sweeps. */
#define QR_ABI_VERSION 4u

typedef enum qr_status {
  QR_OK = 0,
  QR_ERR_ARG = 1,          /* invalid argument (QuantrError in the reference)      */
  QR_ERR_OOM = 2,          /* device allocation failed                             */
  QR_ERR_HIP = 3,          /* HIP runtime error                                    */
  QR_ERR_UNSUPPORTED = 4,  /* valid request outside this engine's limits           */
  QR_ERR_NO_DEVICE = 5     /* no GPU visible                                       */
} qr_status;

typedef enum qr_precision {
  QR_C128 = 0, /* complex128 amplitudes (the reference's Complex64 = 2 x f64) */
  QR_C64 = 1   /* complex64 amplitudes (2 x f32); reductions stay in f64     */
} qr_precision;

/* One value per variant of quantr::Gate (gate.rs:19-106), in declaration order. */
typedef enum qr_gate_kind {
  QR_GATE_ID = 0,
  QR_GATE_H = 1,
  QR_GATE_X = 2,
  QR_GATE_Y = 3,
  QR_GATE_Z = 4,
  QR_GATE_S = 5,
  QR_GATE_SDAG = 6,
  QR_GATE_T = 7,
  QR_GATE_TDAG = 8,
  QR_GATE_RX = 9,      /* arg = angle */
  QR_GATE_RY = 10,     /* arg = angle */
  QR_GATE_RZ = 11,     /* arg = angle */
  QR_GATE_X90 = 12,
  QR_GATE_Y90 = 13,
  QR_GATE_MX90 = 14,
  QR_GATE_MY90 = 15,
  QR_GATE_PHASE = 16,  /* arg = angle (global phase exp(i*arg/2)) */
  QR_GATE_CR = 17,     /* arg = angle, 1 control */
  QR_GATE_CRK = 18,    /* arg_int = k, 1 control */
  QR_GATE_CZ = 19,     /* 1 control */
  QR_GATE_CY = 20,     /* 1 control */
  QR_GATE_CNOT = 21,   /* 1 control */
  QR_GATE_SWAP = 22,   /* 1 "control" wire = the other swapped wire */
  QR_GATE_TOFFOLI = 23,/* 2 controls */
  QR_GATE_CUSTOM = 24, /* any number of controls; tabulated images */
  QR_GATE_KIND_COUNT = 25
} qr_gate_kind;

/*
 * One non-identity gate of a circuit, i.e. one (GateCategory, position) pair that
 * Circuit::simulate_with_register hands to Circuit::apply_gate (simulation.rs:37-56).
 *
 *  position        wire the gate sits on (qubit_counter % num_qubits, simulation.rs:43)
 *  ncontrols       number of control wires, stored at controls[controls_offset ...];
 *                  order as in the Gate variant (Toffoli(c1, c2), Custom(_, vec![...], _))
 *  arg / arg_int   angle of Rx/Ry/Rz/Phase/CR, k of CRk
 *  image_offset    Custom only: the gate acts on k = ncontrols + 1 wires ordered
 *                  [controls..., position], most significant first
 *                  (simulation.rs:102-112,137-156). images[image_offset + (j*2^k + r)*2 + {0,1}]
 *                  is amplitude r of the user function's image of local basis state j.
 *  some_offset     Custom only: some[some_offset + j] is 1 if the user function returned
 *                  Some(image) for local basis state j, 0 for None (simulation.rs:108-133).
 * QR_GATE_ID entries are skipped, as in simulation.rs:38-41.
 */
typedef struct qr_op {
  uint32_t kind;
  uint32_t position;
  uint32_t ncontrols;
  uint32_t reserved0;
  uint64_t controls_offset;
  double arg;
  int32_t arg_int;
  uint32_t reserved1;
  uint64_t image_offset;
  uint64_t some_offset;
} qr_op; /* 56 bytes */

/* A gate program: the flattened, Id-free gate list of a Circuit, in application order. */
typedef struct qr_program {
  const qr_op* ops;
  uint64_t nops;
  const uint32_t* controls;
  uint64_t ncontrols;
  const double* images;
  uint64_t nimages; /* number of doubles */
  const uint8_t* some;
  uint64_t nsome;
} qr_program;

/* Pauli-sum observable O = sum_t coeff_t * P_t. Term t has nfactors (wire, pauli) factors at
 * wires[offset ...] / paulis[offset ...]; pauli is 'X', 'Y' or 'Z' (or 1, 2, 3).
 * No reference counterpart (quantr has no observables). */
typedef struct qr_pauli_term {
  double coeff;
  uint32_t nfactors;
  uint32_t reserved;
  uint64_t offset;
} qr_pauli_term;

typedef struct qr_observable {
  const qr_pauli_term* terms;
  uint64_t nterms;
  const uint32_t* wires;
  const uint8_t* paulis;
  uint64_t nfactors;
} qr_observable;

/* Execution options of a state (defaults: fused = 1, tile_bits = 0, low_bits = 0, exact_sample = 2,
 * tile_variant = 0, reg_bits = 0, jit = 2). */
typedef struct qr_options {
  uint32_t fused;        /* 1: fused tile sweeps (FMA, <=1e-12 vs the reference order);
                            0: one sweep per gate in the reference's exact arithmetic order */
  uint32_t tile_bits;    /* fused tile width k (0 = auto)                              */
  uint32_t low_bits;     /* contiguous low bits per tile segment c (0 = auto)          */
  uint32_t exact_sample; /* sampler; every mode reproduces the reference's sequential left
                            fold bit for bit. 0 and 2 (default): block-parallel (exact
                            integer sums inside each binade of the fold, sequential folds
                            only where it crosses a binade); 1: one lane folds the whole
                            state in order (a checker; slow for large states)           */
  uint32_t tile_variant; /* 0: auto; 1: one tile in registers per workgroup;
                            2: next tile prefetched into registers (double buffer)       */
  uint32_t reg_bits;     /* amplitudes per thread = 2^reg_bits (0 = auto, 3, 4 or 5)    */
  uint32_t jit;          /* tile sweeps compiled for the circuit at run time (hipRTC):
                            0 never; 1 always; 2 (default) from the second application of
                            the same circuit structure in this process, or at once for
                            n >= 28. Plans are cached per structure either way.          */
  uint32_t stage_ops;    /* gates fused into one tile sweep at most (0 = auto). A capped
                            stage leaves its tail to later sweeps, which balances ALU work
                            between sweeps; results do not depend on it.                 */
  uint32_t tune;         /* 0 (default): plans chosen by estimate, identical on every call.
                            1: measured plans. The first applications of a circuit
                            structure time candidate plans (the planner's best estimates)
                            and, per tile sweep, two orders of dealing tiles to workgroups;
                            later applications use the fastest. Every application computes
                            the circuit (candidates agree within the fused mode's
                            tolerance); qr_plan_stats.tuning marks the measuring ones.
                            Needs jit != 0 and tile_variant != 2; ignored otherwise.     */
} qr_options;

typedef struct qr_state qr_state;
typedef struct qr_grad qr_grad;

/* ---- library -------------------------------------------------------------------------- */

uint32_t qr_abi_version(void);
const char* qr_version(void);
/* Thread-local message of the last failing call on this thread ("" if none). */
const char* qr_last_error(void);
qr_status qr_device_count(int* count);

/* ---- state (SuperPosition resident in HBM) -------------------------------------------- */

/* Allocates 2^num_qubits amplitudes on the current device and sets |0...0>.
 * Replaces SuperPosition::new_unchecked (super_positions_unchecked.rs:39-46) as used by
 * Circuit::simulate (circuit.rs:376-378). num_qubits == 0 is QR_ERR_ARG
 * ("The initialised circuit must have at least one wire.", circuit.rs:49-53). */
qr_status qr_state_create(uint32_t num_qubits, qr_precision precision, qr_state** out);
qr_status qr_state_destroy(qr_state* state);
qr_status qr_state_num_qubits(const qr_state* state, uint32_t* out);
qr_status qr_state_precision(const qr_state* state, qr_precision* out);
/* Resets to |0...0> (simulated_circuit.rs:92-97). */
qr_status qr_state_reset(qr_state* state);
/* Sets the register: Circuit::change_register (circuit.rs:463-473). len = 2^n complex
 * values; a length mismatch is QR_ERR_ARG with the reference's dimension message. */
qr_status qr_state_set_amplitudes(qr_state* state, const double* re_im, uint64_t len);
/* Reads the state back: SimulatedCircuit::get_state / take_state
 * (simulated_circuit.rs:158-160,185-187), SuperPosition::get_amplitudes
 * (super_positions.rs:191-193). Chunked variant for states larger than a host buffer. */
qr_status qr_state_get_amplitudes(qr_state* state, double* re_im_out, uint64_t len);
qr_status qr_state_get_amplitudes_range(qr_state* state, uint64_t first, uint64_t count,
                                        double* re_im_out);
qr_status qr_state_set_options(qr_state* state, const qr_options* options);
qr_status qr_state_get_options(const qr_state* state, qr_options* out);
qr_status qr_state_synchronize(qr_state* state);

/* ---- the hot path ---------------------------------------------------------------------- */

/* Checks a program against num_qubits without touching a GPU (circuit.rs:272-305 rules,
 * plus custom-table dimensions). */
qr_status qr_validate_program(uint32_t num_qubits, const qr_program* program);

/* Applies every gate of the program to the state, in order:
 * Circuit::simulate_with_register + Circuit::apply_gate (simulation.rs:21-135). */
qr_status qr_apply_program(qr_state* state, const qr_program* program);

/* Measures the state once per uniform draw u in [0,1): shot s yields the first index i with
 * u_s < sum_{j<=i} |a_j|^2 (left fold in index order), or UINT64_MAX when no such index
 * exists (None). SuperPosition::measure (super_positions.rs:332-342) with the draw
 * injected instead of fastrand::f64(). Used by measure_all (simulated_circuit.rs:63-130). */
qr_status qr_sample(qr_state* state, const double* uniforms, uint64_t shots, uint64_t* idx_out);

/* Many programs, each simulated from |0...0> and measured with its own draws: program j takes
 * draws_per_program[j] uniforms (consecutive in `uniforms`, programs in order) and idx_out
 * receives, per draw, what qr_sample returns on program j's state. The per-shot loop of
 * SimulatedCircuit::measure_all_without_cache (simulated_circuit.rs:81-114), which
 * re-simulates the circuit for every shot because custom functions may answer differently
 * (examples/post_select.rs:39-49), done for all shots at once: programs of one structure share
 * a plan and their sweeps run as one launch over all of their states. Fused mode; states use
 * the current device and are freed before returning. */
qr_status qr_sample_programs(uint32_t num_qubits, qr_precision precision, const qr_program* programs,
                             uint64_t count, const uint64_t* draws_per_program, const double* uniforms,
                             uint64_t* idx_out);

/* sum_i |a_i|^2 in f64. */
qr_status qr_state_norm_sqr(qr_state* state, double* out);

/* <psi|O|psi> in f64 (no reference counterpart). */
qr_status qr_expectation(qr_state* state, const qr_observable* observable, double* out);

/* qr_expectation with no allocation of any kind (no hipMalloc, hipFree or allocating library
 * call): no counterpart in the reference; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The kernel's working
 * set (observable tables and partial sums) lives in a fixed 1 MiB device array per GPU, held
 * until the process ends. At most 16384 terms (more is QR_ERR_ARG). The result is bitwise equal
 * to qr_expectation's. Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents (nothing in it has to start at zero); calls would race on it if they overlapped, so
 * the library serialises them with a host mutex and each returns only after its kernel ends. */
qr_status qr_expectation_noalloc(qr_state* state, const qr_observable* observable, double* out);

/* ---- adjoint differentiation (no reference counterpart) -------------------------------- */

/* A forward/backward pair with the contract of an autograd Function:
 *   forward:  psi = U(theta)|0...0>, E = <psi|O|psi>; keeps psi, U and O (no intermediates)
 *   backward: grads[p] = grad_out * dE/dtheta_p for every Rx/Ry/Rz/Phase gate p of the
 *             program, in program order, from one reverse sweep (adjoint method).
 * Parameters are the `arg` of those gates; other gates are constants. */
qr_status qr_grad_create(uint32_t num_qubits, qr_precision precision, qr_grad** out);
qr_status qr_grad_destroy(qr_grad* grad);
qr_status qr_grad_forward(qr_grad* grad, const qr_program* program,
                          const qr_observable* observable, double* expval_out);
qr_status qr_grad_num_params(const qr_grad* grad, uint64_t* out);
qr_status qr_grad_backward(qr_grad* grad, double grad_out, double* grads_out, uint64_t nparams);
/* Engine options of both states of the context (as qr_state_set_options). With fused != 0
 * and jit != 0 (the default) the reverse sweep runs as fused tile stages carrying psi and
 * lam together, with gradient taps evaluated inside the tiles; otherwise one launch per
 * op and state. */
qr_status qr_grad_set_options(qr_grad* grad, const qr_options* options);

/* ---- batched serving (no reference counterpart) ---------------------------------------- */

/* Many independent parameterised programs of uneven qubit counts, each with an observable:
 * qr_batch_run computes every expectation value and full adjoint gradient (as qr_grad_forward
 * + qr_grad_backward with grad_out = 1) in ONE launch of a HIP graph captured at creation.
 * Creation plans every instance, compiles the tile stages of every distinct structure and
 * allocates all states (2 x 2^n amplitudes per instance) up front.
 * Runs may pass new programs of the same structure (new angles, new custom tables of the same
 * shape); NULL reuses the programs given at creation. Observables are fixed at creation. */
typedef struct qr_batch qr_batch;
/* Batch options (NULL = defaults: all zero).
 *  autotune        1: for every qubit count >= 16 in the batch, time the candidate tile widths
 *                  ("fusion widths") of the forward sweep, then of the reverse sweep, on one
 *                  representative instance and keep the fastest (cached per process; widths
 *                  set below are kept); 0: built-in widths
 *  fwd_tile_bits   forward-sweep tile width (0 = auto)
 *  dual_tile_bits  reverse-sweep tile width (0 = auto or tuned)
 *  streams         concurrent streams inside the graph (0 = 4) */
typedef struct qr_batch_options {
  uint32_t autotune;
  uint32_t fwd_tile_bits;
  uint32_t dual_tile_bits;
  uint32_t streams;
} qr_batch_options;
qr_status qr_batch_create(uint64_t ninstances, const uint32_t* num_qubits,
                          const qr_program* programs, const qr_observable* observables,
                          qr_precision precision, const qr_batch_options* options,
                          qr_batch** out);
qr_status qr_batch_destroy(qr_batch* batch);
/* Parameters (Rx/Ry/Rz/Phase gates) of each instance: counts_out[ninstances]. */
qr_status qr_batch_param_counts(const qr_batch* batch, uint64_t* counts_out);
/* expvals_out[i]; grads_out[grad_offsets[i] ...] (grad_offsets NULL: instances packed in
 * order) receives instance i's gradient; latency_us_out (optional) receives each instance's
 * completion time in microseconds after the graph began, from the device clock. */
qr_status qr_batch_run(qr_batch* batch, const qr_program* programs, double* expvals_out,
                       double* grads_out, const uint64_t* grad_offsets, double* latency_us_out);

/* One-shot form: create, run once, destroy. */
qr_status qr_batch_grad(uint64_t ninstances, const uint32_t* num_qubits,
                        const qr_program* programs, const qr_observable* observables,
                        qr_precision precision, double* expvals_out, double* grads_out,
                        const uint64_t* grad_offsets, double* latency_us_out);

/* ---- measurement support --------------------------------------------------------------- */

/* Kernel timing on the state's own stream with HIP events (enable = 1 starts a new record).
 * qr_profile_read returns, per kernel family, launches and total milliseconds. */
typedef struct qr_profile_entry {
  char name[48];
  uint64_t launches;
  double total_ms;
  double bytes; /* algorithmic HBM bytes moved by those launches */
} qr_profile_entry;
qr_status qr_profile_enable(qr_state* state, int enable);
/* The streaming ceiling of a sweep of this state: `reps` launches of an in-place kernel that
 * reads every amplitude, multiplies it by 1 and writes it back (one amplitude per thread, the
 * fastest in-place shape measured); ms_out = average milliseconds per launch. The state's
 * values are unchanged. No reference counterpart. */
qr_status qr_measure_stream(qr_state* state, uint32_t reps, double* ms_out);
qr_status qr_profile_read(qr_state* state, qr_profile_entry* entries, uint64_t capacity,
                          uint64_t* count_out);
/* Statistics of the last qr_apply_program: sweeps launched and algorithmic bytes. */
typedef struct qr_plan_stats {
  uint64_t gates;           /* non-identity gates in the program          */
  uint64_t sweeps;          /* kernel launches that touch the state        */
  uint64_t tile_sweeps;     /* of which fused tile sweeps                  */
  double bytes;             /* sum over sweeps of 2 * 2^n * B * f_touch    */
  uint64_t max_stage_gates; /* most gates fused into one tile sweep        */
  uint64_t restore_sweeps;  /* gate-free tile sweeps that restore the
                               state's bit order (moving layout)           */
  uint64_t tuning;          /* 1 if this application measured a candidate
                               (qr_options.tune), else 0                    */
  uint64_t xcd_order_sweeps;/* tile sweeps whose tiles were dealt to the
                               XCDs in contiguous eighths                  */
} qr_plan_stats;
qr_status qr_state_last_plan(const qr_state* state, qr_plan_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* QUANTR_HIP_H */
