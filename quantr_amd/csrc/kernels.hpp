// kernels.hpp — argument blocks and host launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qr {

// Exact per-gate sweep (K <= 4 acting wires, amplitudes held in registers).
struct ExactArgs {
  int K;
  uint8_t sorted_bits[8];  // acting bits ascending
  uint64_t off[16];        // offset of local basis state j (MSB-first local index)
  uint8_t order[16];       // local states in ascending global offset order
  uint32_t some;           // bit j: input j contributes
  uint32_t keep;           // bit r: output r keeps its input (custom None)
};

struct ExactWideArgs {
  int K;
  uint8_t sorted_bits[20];
};

struct DenseKArgs {
  int k;
  uint8_t sorted_bits[20];
  uint64_t cmask;
};

struct DiagKArgs {
  int k;
  uint8_t bits[20];  // local MSB first
  uint64_t cmask;
};

// Tile op codes (device interpreter).
enum TileCode : uint32_t {
  TC_END = 0,
  TC_REMAP = 1,      // a = index of the new slot map (data moves through LDS)
  TC_U1 = 2,         // a = register slot, coef -> 4 complex, b = form
  TC_D1 = 3,         // a = global bit, b = register slot or 0xFF, coef -> 2 complex, c = only |1>
  TC_SWAPSLOTS = 4,  // a = index of the new slot map (relabel, no data movement)
  TC_GRAD = 5,       // compiled dual-state stages only: gradient partial of one parameter
  TC_SCALE = 6,      // every amplitude times coef 0 (the stage's folded scalars)
};

// One interpreted op of a tile program (32 bytes, read with scalar loads).
struct TileOpDev {
  uint32_t w0;          // code | a << 8 | b << 16 | flags << 24
  uint32_t ns;          // register slots that must be 1
  uint64_t lane_cmask;  // control bits on thread slots / outside the tile
  uint32_t map;         // REMAP / SWAPSLOTS: new map index
  uint32_t swz;         // REMAP: swizzle table index
  uint32_t pad[2];
};
static_assert(sizeof(TileOpDev) == 32, "TileOpDev layout");

struct TileArgs {
  const TileOpDev* ops;
  const double2* coefs;   // 4 per op, op o at coefs[4 o]
  const uint32_t* maps;   // per map, K words: global bit | tile position << 8
  const uint32_t* swz;    // per remap, 2K words: LDS pattern of each slot, old map then new map
  uint64_t ntiles;
  uint64_t nontile_mask;  // index bits outside the tile; tile number bits are deposited there
  uint32_t nops;
  uint32_t pad;
};

// A block of the exact block-parallel sampler whose fold stays in one binade: its draws are
// u[first .. last), its fold starts at m_start units of `unit` = 2^(e-52), scale = 1 / unit.
struct ExactJob {
  uint64_t block;
  double m_start;
  double unit, scale;
  uint64_t first, last;
};

// Tile kernel variants available: tile bits K and register bits RB per precision.
int tile_max_bits(int prec);  // widest tile variant compiled (13)
int tile_reg_bits(int K);

// ---- launchers (prec: 0 = complex128, 1 = complex64) ----
hipError_t launch_init_basis(int prec, void* psi, uint64_t dim, uint64_t one_at, hipStream_t s);
hipError_t launch_exact(int prec, void* psi, uint64_t ngroups, const ExactArgs& a, const double2* M,
                        hipStream_t s);
hipError_t launch_exact_wide(int prec, void* psi, void* tmp, uint64_t tmp_groups, uint64_t ngroups,
                             const ExactWideArgs& a, const double2* M, const uint64_t* off,
                             const uint32_t* order, const uint8_t* flags, hipStream_t s);
hipError_t launch_dense1(int prec, void* psi, uint64_t npairs, int t, uint64_t cmask, const double2* u,
                         hipStream_t s);
hipError_t launch_diag1(int prec, void* psi, uint64_t dim, int t, uint64_t cmask, double2 d0, double2 d1,
                        hipStream_t s);
hipError_t launch_densek(int prec, void* psi, void* tmp, uint64_t tmp_groups, uint64_t ngroups,
                         const DenseKArgs& a, const double2* M, const uint64_t* off, hipStream_t s);
// Dense 2^k x 2^k, 2 <= k <= 4, each group of amplitudes in registers (no staging).
hipError_t launch_densek_reg(int prec, void* psi, uint64_t ngroups, const DenseKArgs& a, const double2* M,
                             const uint64_t* off, hipStream_t s);
hipError_t launch_diagk(int prec, void* psi, uint64_t dim, const DiagKArgs& a, const double2* D,
                        hipStream_t s);
hipError_t launch_tile(int prec, int K, int RB, bool prefetch, void* psi, const TileArgs& a, hipStream_t s);
hipError_t launch_sample_fold(int prec, const void* psi, uint64_t dim, const double* u, uint64_t shots,
                              uint64_t* out, hipStream_t s);
hipError_t launch_block_mass(int prec, const void* psi, uint64_t dim, int bb, double* mass, hipStream_t s);
// Exact sums of rint(p_i * 2^(52-e)) per block for e = ebase[b] + 0..2, and halfway flags.
hipError_t launch_block_exact_sums(int prec, const void* psi, uint64_t dim, int bb, const int32_t* ebase,
                                   double* sums, uint32_t* flags, hipStream_t s);
// Resolves each job's draws with its block's exact fold values (bb >= 12).
hipError_t launch_sample_exact_blocks(int prec, const void* psi, uint64_t dim, int bb, const ExactJob* jobs,
                                      uint64_t njobs, const double* u, uint64_t* out, hipStream_t s);
hipError_t launch_norm_partial(int prec, const void* psi, uint64_t dim, double* part, int nblocks,
                               hipStream_t s);
hipError_t launch_gen_partial(int prec, const void* lam, const void* psi, uint64_t dim, int t, int gen,
                              double2* part, int nblocks, hipStream_t s);
// One Pauli-sum term of an X-mask group: coefficient (with i^ny folded in) times
// (-1)^{popcount(j & zm)}.
struct PauliZTerm {
  uint64_t zm;
  double cr, ci;
};
// An X-mask group of a Pauli sum: terms[first .. first + count) share X mask xm.
struct PauliGroup {
  uint64_t xm;
  uint32_t first, count;
};
// lam[o] = sum over groups g, terms t of g: c_t (-1)^{popcount(j & zm_t)} psi[j], j = o ^ xm_g
// (the whole observable in one pass).
hipError_t launch_pauli_sum(int prec, void* lam, const void* psi, uint64_t dim, const PauliGroup* groups, int ngroups,
                            const PauliZTerm* terms, hipStream_t s);
// |0...0> in each of `count` states of dim amplitudes stored back to back.
hipError_t launch_init_instances(int prec, void* psi, uint64_t dim, uint64_t count, hipStream_t s);
// For each of `count` states stored back to back: its draws u[off[y] .. off[y+1]) (ascending)
// placed by the left fold of its probabilities in index order (out = first index whose
// cumulative probability exceeds the draw, UINT64_MAX if none).
hipError_t launch_sample_fold_instances(int prec, const void* psi, uint64_t dim, uint64_t count, const double* u,
                                        const uint64_t* off, uint64_t* out, hipStream_t s);
// part[b] = partial sums of Re <psi| O |psi> over nblocks blocks, O as in launch_pauli_sum
// (added to part[b] with `accumulate`). Every group costs a read of the state.
hipError_t launch_pauli_expect(int prec, const void* psi, uint64_t dim, const PauliGroup* groups, int ngroups,
                               const PauliZTerm* terms, double* part, int nblocks, hipStream_t s,
                               int accumulate = 0);
// Groups of an observable evaluated together in one read of the state: groups[first ..
// first + count) whose X masks lie in the tile bits tmask (pauli_tile_bits of them), or with
// tmask == 0 groups read directly (X masks too wide for a tile).
struct PauliPass {
  uint64_t tmask;
  uint32_t first, count;
};
// Tile width of the grouped expectation passes: 2^12 complex128 / 2^13 complex64 amplitudes
// (64 KB of LDS), of which the lowest 3 / 4 bits (128-byte segments) are always tile bits.
inline int pauli_tile_bits(int prec) { return prec == 0 ? 12 : 13; }
inline int pauli_low_bits(int prec) { return prec == 0 ? 3 : 4; }
// Partial sums (workgroups) of every expectation pass over a state of dim amplitudes.
int pauli_expect_partials(uint64_t dim);
// part[0 .. pauli_expect_partials(dim)) = partial sums of Re <psi|O|psi>, pass after pass;
// lxm[g] = group g's X mask in the tile coordinates of its pass.
hipError_t launch_pauli_expect_passes(int prec, const void* psi, uint64_t dim, const PauliGroup* groups,
                                      const uint32_t* lxm, const PauliZTerm* terms, const PauliPass* passes,
                                      int npasses, double* part, hipStream_t s);
// part[b] = partial sums of Re sum_j conj(a_j) b_j over nblocks blocks.
hipError_t launch_dot_re_partial(int prec, const void* a, const void* b, uint64_t dim, double* part, int nblocks,
                                 hipStream_t s);
// psi[i] *= f for every amplitude, in place (the streaming ceiling of a sweep).
hipError_t launch_stream_scale(int prec, void* psi, uint64_t dim, double f, hipStream_t s);
// out[0] = the device wall clock (constant-rate counter).
hipError_t launch_stamp(uint64_t* out, hipStream_t s);
// out[r] = sum of part[r * ncols + 0 .. ncols-1] (fixed order), r < nrows.
hipError_t launch_sum_rows(const double* part, uint64_t ncols, uint32_t nrows, double* out, hipStream_t s);
hipError_t launch_f64_to_c64(float2* dst, const double2* src, uint64_t n, hipStream_t s);
hipError_t launch_c64_to_f64(double2* dst, const float2* src, uint64_t n, hipStream_t s);

}  // namespace qr
