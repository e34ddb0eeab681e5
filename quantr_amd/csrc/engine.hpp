// engine.hpp — internal declarations of the gfx950 state-vector engine.
#pragma once

#include <cstdint>
#include <complex>
#include <string>
#include <vector>

#include "quantr_hip.h"
#include "kernels.hpp"

namespace qr {

using cd = std::complex<double>;

// ---- errors -----------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

struct Error {
  qr_status status;
  std::string msg;
};

// ---- gate lowering ------------------------------------------------------------------------

// A gate exactly as quantr applies it: the matrix over its acting wires
// [controls..., position] (MSB first), column j = image of local basis state j, with the
// custom-gate overwrite rule folded in (simulation.rs:108-133).
struct ExactGate {
  int kind = 0;
  int K = 0;                  // acting wires
  std::vector<int> bits;      // global bit positions (n-1-wire), local MSB first
  std::vector<cd> M;          // M[r*dim + j]
  std::vector<uint8_t> some;  // input j contributes (always 1 for standard gates)
  std::vector<uint8_t> keep;  // output r is overwritten by its input (custom None rows)
  bool custom = false;
  double arg = 0.0;
};

// Fast micro-op kinds used by the fused and unfused-fast paths.
enum FastKind : int {
  FK_DENSE1 = 0,  // 2x2 on one target bit (+ controls)
  FK_DIAG1 = 1,   // diag(d0, d1) on one bit (+ controls)
  FK_DENSEK = 2,  // 2^k x 2^k on k target bits (+ controls)
  FK_DIAGK = 3,   // 2^k phases on k bits (+ controls)
  FK_GRAD = 4,    // adjoint sweep only: Im <lam| G |psi> of parameter `param` (G by `gen`)
};

// Specialisation hints of a 2x2 (kernels pick a cheaper form when exact).
enum U1Form : int {
  U1_GENERAL = 0,
  U1_PERM = 1,   // anti-diagonal [[0,b],[c,0]] (X, Y, X90, ...)
  U1_REAL = 2,   // real entries (H, Ry)
};

// Kernel forms of a register-slot 2x2 (TileOpHost::b of TC_U1): the U1Form values, then
enum TileForm : int {
  TF_X = 3,       // [[0,1],[1,0]]: a relabel
  TF_RX = 4,      // real diagonal, imaginary off-diagonal (Rx)
  TF_HAD = 5,     // [[1,1],[1,-1]] (H with its scalar factored out)
  TF_NEGSWAP = 6, // [[0,-1],[1,0]] (Y, Y90, MY90 with their scalar factored out)
  TF_RROT = 7,    // real rotation as three shears, al = c[0], be = c[1] (Ry; magnitude folded out)
  TF_XROT = 8,    // the same for [[c,-is],[-is,c]], the rotation of (x, i y) (Rx)
};

// Exact value classes of an op, used to factor scalars out of uncontrolled gates.
enum OpSub : int {
  SUB_NONE = 0,
  SUB_HAD = 1,     // DENSE1: m00 = m01 = m10 = -m11, real
  SUB_XLIKE = 2,   // DENSE1: anti-diagonal, m01 = m10 (X90 = -iX, MX90 = iX, X)
  SUB_YLIKE = 3,   // DENSE1: anti-diagonal, m01 = -m10 (Y = i[[0,-1],[1,0]], Y90, MY90)
  SUB_RX = 4,      // DENSE1: real diagonal, imaginary off-diagonal
  SUB_D0 = 5,      // DIAG1: m0 != 1 (factor m0 out, leaving diag(1, m1/m0))
  SUB_RROT = 6,    // DENSE1: real rotation [[a,-b],[b,a]]
  SUB_XROT = 7,    // DENSE1: [[a,-ib],[-ib,a]], a and b real
};

struct FastOp {
  int kind = FK_DENSE1;
  int k = 1;
  int bits[8] = {0};    // target bits, local MSB first
  uint64_t cmask = 0;   // control bits that must be 1
  std::vector<cd> m;    // DENSE1: 4 row-major; DIAG1: 2; DENSEK: dim*dim; DIAGK: dim
  int form = U1_GENERAL;
  int sub = SUB_NONE;
  // Differentiation: index of the angle parameter (-1 if constant) and its generator
  // (1 = X, 2 = Y, 3 = Z, 4 = identity / global phase): U = exp(-i theta G / 2), or
  // exp(+i theta / 2) I for the global phase.
  int param = -1;
  int gen = 0;
  int gate_index = -1;  // position in the program (for dependency order)
  double f_touch = 1.0; // fraction of amplitudes this op must read and write
  uint64_t target_mask() const;
  uint64_t read_mask() const;  // controls and diagonal bits
};

// One Pauli-sum term as bit masks: coeff * i^ny * X^xm Z^zm (index bits, wire 0 = MSB).
struct PauliMasks {
  double coeff;
  uint64_t xm, zm;
  int ny;
};
// Checked observable terms (wires in range, no repeated wire, X/Y/Z only). Throws Error.
std::vector<PauliMasks> parse_observable(uint32_t n, const qr_observable* o);
PauliMasks parse_term(uint32_t n, const qr_observable* o, uint64_t t);
// A term as the device's Z-mask term: the coefficient times i^ny.
PauliZTerm z_term(const PauliMasks& tm);
// The observable as device tables: terms grouped by X mask (i^ny folded into coefficients).
void pauli_tables(const std::vector<PauliMasks>& terms, std::vector<PauliGroup>& groups,
                  std::vector<PauliZTerm>& flat);
// Orders an observable's G groups into expectation passes (no allocation): groups whose X
// masks fit one tile are packed first-fit, widest first, into passes of one state read each;
// groups too wide for a tile form a last, direct pass. Writes each group's X mask in its
// pass's tile coordinates to lxm[] and returns the number of passes (at most G + 1).
int pauli_passes(uint32_t n, int prec, PauliGroup* groups, int G, PauliPass* passes, uint32_t* lxm);

// Validates and lowers a program. Throws Error.
void validate_program(uint32_t n, const qr_program* p);
std::vector<ExactGate> lower_exact(uint32_t n, const qr_program* p);
std::vector<FastOp> lower_fast(uint32_t n, const std::vector<ExactGate>& gates, bool want_params);
// Human-readable form of an op (Rust Debug style, e.g. "CNot(2)").
std::string gate_debug(const qr_op& op, const uint32_t* controls);

// Sets form and sub of a DENSE1 / DIAG1 op from its matrix.
void classify(FastOp& op);
// Adjoint of a fast op (for the reverse sweep).
FastOp adjoint(const FastOp& op);
// The reverse sweep of a parameterised program as one op list over (psi, lam): from the
// last op to the first, a gradient tap (FK_GRAD) for each parameter op, then the op's adjoint.
std::vector<FastOp> adjoint_sweep_ops(const std::vector<FastOp>& ops);

// ---- planning -------------------------------------------------------------------------------

struct TileOpHost {
  uint32_t code = 0;
  uint32_t a = 0;        // U1: register slot; D1: global bit; REMAP/SWAPSLOTS: new map index
  uint32_t b = 0;        // U1: form; D1: register slot holding the bit, or 0xFF;
                         // REMAP with flags bit 0 (split): the tile position splitting rounds
  uint32_t flags = 0;    // D1: bit 0 = only the |1> half changes; bits 4-7 = factor form
  uint32_t ns = 0;       // register slots that must be 1 (controls held in registers)
  uint64_t lane_cmask = 0;  // control bits held by thread slots or outside the tile
  uint32_t swz = 0;      // REMAP: index of the LDS swizzle table
  uint32_t param = 0;    // GRAD: parameter row of the partial-sum table
  int src = -1;          // U1 / D1: index of the op whose matrix supplies the coefficients
  int fold = 0;          // scalar folding (FoldMode)
  cd coef[4];
};

// How an op takes part in a stage's scalar folding: its factor is divided out of its
// coefficients (FOLD_M00: m00, FOLD_M10: m10, FOLD_D0: d0), and the product of the stage's
// factors is multiplied into one general-form op (FOLD_ABSORB) or applied by TC_SCALE.
// FOLD_RROT / FOLD_XROT divide out the magnitude and a sign, leaving a rotation by at most
// pi/2 that the kernel applies as three shears.
enum FoldMode : int {
  FOLD_NONE = 0, FOLD_M00 = 1, FOLD_M10 = 2, FOLD_D0 = 3, FOLD_ABSORB = 4, FOLD_SCALE = 5, FOLD_RROT = 6,
  FOLD_XROT = 7
};

struct Stage {
  bool tile = true;            // fused tile sweep, or a standalone op
  int k = 0;                   // tile bits
  std::vector<int> tbits;      // tile bits ascending (global positions)
  std::vector<int> maps;       // flat: maps[i*k + s] = tile position held by slot s
  std::vector<uint16_t> swz;   // flat, per remap: LDS pattern of each slot, old map then new map
  std::vector<TileOpHost> ops;
  int standalone = -1;         // index into the op list for standalone stages
  double bytes = 0.0;          // algorithmic bytes of this sweep (2 * 2^n * B * f_touch)
  int ngates = 0;
  // physical index bit of each tile position when the tile is loaded and when it is stored
  // (they differ when the stage moves bits into the physical low bits); maps address the
  // load layout except map store_map (the last one), which addresses the store layout
  std::vector<int> pload, pstore;
  int store_map = -1;
  int phys_bit(int map_index, int pos) const {
    return map_index == store_map ? pstore[pos] : pload[pos];
  }
};

struct PlanParams {
  int tile_bits = 12;
  int reg_bits = 4;
  int low_bits = 4;
  int max_ops = 0;    // ops per tile stage at most (0 = no limit)
  int lds_split = 0;  // remaps exchange the tile through LDS in two halves (half the LDS)
  int dyn_low = 0;    // moving layout: stages choose which bits the physical low bits hold
  uint32_t seed = 0;  // 0: ties between equally good bits go to the lowest; else a seeded pick
  int tie_slack = 0;  // with a seed: bits up to tie_slack ops short of the best also tie
};

// A remap done in two halves ("split", TileOpHost::flags bit 0) moves the amplitudes whose
// tile position `b` is 0 in the first round and those where it is 1 in the second; position b
// carries the top LDS address bit in the pattern table, which rounds drop. In each round a
// thread writes the registers it holds of that half (all of them, or none, when b is a thread
// slot) and then reads the ones it holds of that half in the new map. A first-round read
// whose register is still to be written in the second round goes to a temporary.
//   old_slot / new_slot: the slot holding position b in the old / new map; nts: thread slots.
inline bool split_read_needs_tmp(int old_slot, int new_slot, int nts, int m) {
  if (old_slot < nts) return !(new_slot == old_slot);       // written by all or no registers later
  return ((m >> (old_slot - nts)) & 1) != 0;                 // register m is written in round 1
}

// Stage op limit when one tile covers the whole state: such a stage could otherwise absorb the
// entire program into one compiled kernel (long compiles, large code); a split costs one launch
// on a cache-resident state.
constexpr int kWholeStateStageOps = 96;

std::vector<Stage> plan_fused(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                              int bytes_per_amp, bool dual);
// States of at least 2^kBalanceMinQubits amplitudes get balanced forward plans.
constexpr uint32_t kBalanceMinQubits = 25;
// plan_fused with the ops per stage chosen by estimated duration (HBM vs ALU time per stage)
// among a few caps, for forward plans of large states without a caller-set cap; otherwise
// plan_fused itself.
std::vector<Stage> plan_fused_balanced(uint32_t n, const std::vector<FastOp>& ops, const PlanParams& pp,
                                       int bytes_per_amp, bool dual);
// The balanced planner's distinct candidate plans (at most `count`), lowest estimate first;
// one plan (plan_fused's) where plan_fused_balanced would not balance.
std::vector<std::vector<Stage>> plan_fused_candidates(uint32_t n, const std::vector<FastOp>& ops,
                                                     const PlanParams& pp, int bytes_per_amp, bool dual,
                                                     int count, std::vector<PlanParams>* chosen = nullptr);
// Appends what the planner's output depends on for one op (kind, bits, controls and the
// value-derived forms), so that programs with equal keys share one plan; coefficients are
// then taken from each program's own ops (TileOpHost::src).
void plan_key_append(const FastOp& op, std::vector<uint64_t>& key);
// The 4 coefficients of every op of a tile stage (c[4i..4i+3]), taken from the program's own
// ops with the stage's scalar folding applied.
void stage_coefficients(const Stage& s, const std::vector<FastOp>& ops, std::vector<cd>& c);

}  // namespace qr
