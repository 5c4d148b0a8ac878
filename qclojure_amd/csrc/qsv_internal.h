// qsv_internal.h — shared between the kernel file and the C-ABI file.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/qsv.h"

struct qsv_state;

namespace qsv {

constexpr int kBlock = 256;          // threads per workgroup (4 waves of 64)
constexpr int kReduceBlocks = 1024;  // fixed grid for FP64 reductions
constexpr int kMaxPauliPerPass = 8;  // Pauli terms evaluated per pass over the state
constexpr int kSampleChunkLog = 10;  // amplitudes per chunk in the sampling prefix

// Zero-bit insertion plan: enumerate `count` indices j and map each to the
// full index by inserting zeros at the ascending bit positions ins[0..nins)
// and OR-ing `orMask` (bits forced to one, e.g. controls).
struct InsArgs {
  uint64_t count;
  uint64_t orMask;
  int32_t nins;
  int32_t ins[5];
};

template <typename R>
struct Mat2 {
  R m[8];  // row-major 2x2, (re, im) interleaved
};

template <typename R>
struct Cplx {
  R re, im;
};

// A per-gate launch: high bits are inserted (ins), low bits (< threshold) are
// lane bits: lo_pred must all be set for the gate to act, lo_t is a low
// target. tbit is a high target (pair kernel) or the target whose value picks
// d0/d1 (phase kernel).
struct GateArgs {
  InsArgs ins;
  uint64_t lo_pred;
  uint64_t tbit;
  uint64_t lo_t;
};

// Two-bit exchange: hA/hB are high bits (0 when low), lA/lB low bits.
// Exactly one high bit is always passed as hA with the low one as lB.
struct SwapArgs {
  InsArgs ins;
  uint64_t lo_pred;
  uint64_t hA, hB;
  uint64_t lA, lB;
};

// ---- launchers ---------------------------------------------------------------
// All take a raw device pointer to the amplitude array (interleaved complex128
// or complex64, passed as void*) and launch on `st`.
// qsv_gates.hip
void launch_gate1(int prec, void* psi, const GateArgs& g, const double m[8], hipStream_t st);
void launch_phase(int prec, void* psi, const GateArgs& g, const double d0[2], const double d1[2], hipStream_t st);
void launch_swap2(int prec, void* psi, const SwapArgs& s, const double ph[2], hipStream_t st);
void launch_blockade(int prec, void* psi, uint64_t S, uint64_t bmask, const double ph[2], hipStream_t st);
// qsv_kernels.hip
void launch_fill_basis(int prec, void* psi, uint64_t S, uint64_t idx, hipStream_t st);
void launch_scale(int prec, void* psi, uint64_t S, double f, hipStream_t st);
// normalize-state without a host round trip: FP64 norm into out[0], then
// the scale by 1/norm when the norm exceeds 1e-12 (partials: kReduceBlocks doubles)
void launch_renorm(int prec, void* psi, uint64_t S, double* partials, double* out, hipStream_t st);
void launch_copy(int prec, const void* src, void* dst, uint64_t S, hipStream_t st);
// out[0..4] = Re <a|b>, Im <a|b>, sum |a|^2, sum |b|^2, max |a_i - b_i|;
// partials holds 5 * kReduceBlocks doubles.
void launch_compare(int pa, const void* a, int pb, const void* b, uint64_t S, double* partials, double* out,
                    hipStream_t st);
void launch_f64_to_c64(const double* src, void* dst, uint64_t count, hipStream_t st);
void launch_c64_to_f64(const void* src, double* dst, uint64_t count, hipStream_t st);

// FP64 reductions. `partials` holds kReduceBlocks * nterms * 2 doubles; the
// final values land in out[0..nterms*2) (re, im) on the device.
void launch_norm2(int prec, const void* psi, uint64_t S, double* partials, double* out, hipStream_t st);
void launch_inner(int prec, const void* a, const void* b, uint64_t S, double* partials, double* out,
                  hipStream_t st);
void launch_pauli(int prec, const void* psi, uint64_t S, int nterms, const uint64_t* xmask, const uint64_t* smask,
                  double* partials, double* out, hipStream_t st);
void launch_expect1q(int prec, const void* psi, uint64_t S, uint64_t tbit, const double m[8], double* partials,
                     double* out, hipStream_t st);
void launch_probs(int prec, const void* psi, uint64_t offset, uint64_t count, double* out, hipStream_t st);

// Sampling: chunk sums -> inclusive prefix -> one wave per shot.
void launch_sample(int prec, const void* psi, uint64_t S, double* chunk_prefix, uint64_t shots, uint64_t seed,
                   uint64_t* out, hipStream_t st);

// Measurement of a qubit subset: outcome histogram (bits gathered at positions
// pos[0..m)), then collapse.
void launch_outcome_hist(int prec, const void* psi, uint64_t S, int m, const int* pos, double* hist,
                         hipStream_t st);
void launch_collapse(int prec, void* psi, uint64_t S, int m, const int* pos, uint64_t sel, double scale,
                     hipStream_t st);

// ---- fused tiled passes: planning (qsv_api.hip) ------------------------------
constexpr int kTileBitsC128 = 12;   // 4096 x 16 B = 64 KiB of LDS per tile
constexpr int kTileBitsC64 = 12;    // 4096 x  8 B = 32 KiB: three workgroups per CU
constexpr int kMaxPassGates = 96;
constexpr int kMaxPassStages = 32;

// FG_RENORM: divide the state by its FP64 norm when that exceeds 1e-12, as
// normalize-state does after a Kraus operator (a record of batched LDS
// circuits, a barrier between fused passes; never inside a pass).
enum { FG_DENSE = 0, FG_PHASE = 1, FG_SWAP = 2, FG_BLOCKADE = 3, FG_RENORM = 4 };
// Structure of a dense 2x2, which decides the arithmetic it needs per pair:
// general (16 real multiply-adds), real entries (8), real diagonal with
// imaginary off-diagonal, as RX (8), or the exchange [[0,1],[1,0]] (none).
enum { DM_GENERAL = 0, DM_REAL = 1, DM_RX = 2, DM_EXCHANGE = 3 };

// A gate in tile-local bit coordinates.
struct FGate {
  int32_t kind;        // FG_*
  int32_t a, b;        // dense/phase target a; swap bits a, b
  uint32_t cmask;      // local control bits (all must be 1)
  uint32_t bmask;      // blockade bits
  int32_t sub;         // FG_PHASE: 1 when d0 == 1; FG_DENSE: DM_* structure of m
  double m[8];         // dense: row-major 2x2; phase: d0 = m[0..1], d1 = m[6..7]; swap/blockade: phase m[0..1]
};

struct FStage {
  uint32_t regmask;    // tile-local bits held in registers during this stage
  int32_t g0, g1;      // gates [g0, g1)
};

// Complex64 passes compute in FP32 by default. The rounded 2x2 matrices are
// unitary only to ~1e-7, so the norm would drift by that much per gate; it is
// removed pass by pass instead: a pass sums |amplitude|^2 per tile in FP64
// (FP_RENORM_OUT, and FP_RENORM_IN for the input of the first pass of a run),
// a one-workgroup kernel turns the sums into scale = sqrt(target / norm^2),
// and the next pass multiplies its loads by it (FP_RENORM_APPLY).
enum { FP_RENORM_APPLY = 1, FP_RENORM_IN = 2, FP_RENORM_OUT = 4 };
// A pass program flag beside them: the pass starts from |0..0> instead of
// loading the state (the first pass round of a grouped batch).
enum { FP_INIT_ZERO = 8 };

struct TileNorm {
  double* part_in;    // per-tile input norm^2 (first pass of a run)
  double* part_out;   // per-tile output norm^2
  double* ctl;        // [0] target norm^2, [1] scale for the next pass
};

struct FPass {
  int32_t nstages;
  int32_t ngates;
  uint32_t load_mask, store_mask;   // register masks used for the global load / store
  int32_t renorm;                   // FP_RENORM_* bits
  int32_t pad;
  int32_t pos[16];                  // global index bit of tile-local bit i
  int32_t sorted_pos[16];           // the same bits in ascending order
  FStage stages[kMaxPassStages];
  FGate gates[kMaxPassGates];
};

// ---- fused passes as micro-op programs (qsv_pass.hip) ------------------------
// A pass compiled against the layouts it runs in: every op names the register
// bit, slot mask and thread-id mask it acts on (compile_pass, qsv_api.hip).
constexpr int kPassTileBits = 12;
constexpr int kMaxPassOps = kMaxPassGates + kMaxPassStages + 1;
// Dense structures share the DM_* numbering.
enum { PM_GENERAL = DM_GENERAL, PM_REAL = DM_REAL, PM_RX = DM_RX, PM_EXCHANGE = DM_EXCHANGE };
enum {
  PO_DENSE = 0,       // + 4 * register bit + structure (PM_*): 0..15
  PO_PHASE = 16,      // + 4 * only1 + register bit: diag(d0, d1) on a register bit, 16..23
  PO_PHASE_T = 24,    // diag(d0, d1) on a thread bit (aux = its thread-id bit)
  PO_BLOCKADE = 25,   // phase where exactly one of the bits aux (thread) / kc (slot) is set
  PO_RELAYOUT = 26,   // layouts tmask -> kc (PProg::lay), new register mask aux; only1: wave-local
  PO_SWAP = 27,       // exchange local bits kc and aux where the local bits tmask are set
  PO_BLOCK = 28,      // dense 16x16 on the four register bits (FP32 passes only): table at byte offset aux
};
struct POp {
  int32_t code;
  uint32_t tmask;     // thread-id bits that must be set (thread-bit controls)
  uint32_t kc;        // slot bits that must be set (register-bit controls)
  uint32_t aux;
  int32_t only1;      // PO_PHASE_T: d0 == 1; PO_RELAYOUT: wave-local
  int32_t pad;
  double m[8];        // as FGate
  float f[8];         // m rounded to FP32, for FP32-arithmetic passes
};
// A register layout of a tile resolved on the host. Thread t, slot k holds
// tile-local index pdep(t, ~reg) | pdep(k, reg); with t_j the bits of t, its
// LDS byte offset is XOR_j (t_j ? lt[j] : 0) XOR ls[k] (the swizzle is
// XOR-linear) and its global element offset is
// tile base | OR_j (t_j ? gt[j] : 0) | gs[k].
struct PLayout {
  uint32_t lt[8];
  uint32_t ls[16];
  uint64_t gt[8];
  uint64_t gs[16];
};
constexpr int kMaxPassLayouts = kMaxPassStages + 2;
// PO_BLOCK tables per pass: 4 steps x 64 lanes of (re, im) FP32, lane l of
// step kk holding U[l % 16][4 kk + l / 16] (the A operand of
// v_mfma_f32_16x16x4f32), U indexed by slot.
constexpr int kMaxPassBlocks = 8;
struct PHead {
  int32_t nops;
  int32_t renorm;                   // FP_RENORM_* bits
  uint32_t load_mask, store_mask;   // register masks of the global load / store
  int32_t store_local;              // the layout change before the store is wave-local
  int32_t load_lay, store_lay;      // their entries in the layout table
  int32_t nlay;
  int32_t lay_off;                  // byte offset of the layout table from the program start
  int32_t pad;
  int32_t sorted_pos[16];           // the tile's global bits in ascending order
};
// Ops come right after the header; the layout table sits at h.lay_off, at
// `lay` here or packed right after the last op (grouped batches).
struct PProg {
  PHead h;
  POp ops[kMaxPassOps];
  PLayout lay[kMaxPassLayouts];
  float blk[kMaxPassBlocks][4 * 64 * 2];
};
// f32: complex64 arithmetic in FP32 (ignored for complex128, always FP64)
void launch_pass(int prec, bool f32, void* psi, const PProg* dprog, uint64_t ntiles, const TileNorm& nb,
                 hipStream_t st);
// A group of `count` states of S amplitudes (S >= one tile) back to back: one
// pass round (offs[y] = byte offset of instance y's program in `progs`, or -1).
void launch_pass_group(int prec, void* base, uint64_t S, unsigned count, const char* progs, const int64_t* offs,
                       hipStream_t st);

// The scale the next pass of a complex64 FP32 run applies (from the per-tile norms).
void launch_norm_ctl(const TileNorm& nb, uint64_t ntiles, bool first, hipStream_t st);
void launch_fill_group(int prec, void* base, uint64_t S, uint64_t count, hipStream_t st);

// ---- batched small circuits (qsv_batch.hip) -----------------------------------
constexpr int kBatchThreads = 256;

// One decoded gate of a batched circuit, in index-bit coordinates.
struct BRec {
  int32_t kind;        // FG_DENSE, FG_PHASE, FG_SWAP or FG_BLOCKADE
  int32_t a, b;        // target bit / swap bits
  int32_t phase_only;
  uint64_t cmask;      // control bits (all must be 1)
  uint64_t bmask;      // blockade bits
  double m[8];         // as FGate
};

struct BInst {
  int32_t n;
  int32_t pad;
  uint64_t g0, g1;       // gate records [g0, g1)
  uint64_t t0, t1;       // Z terms [t0, t1)
  uint64_t out_offset;   // amplitude offset of the final state in the output arena, ~0 for none
};

int setup_lds_batch(int prec, size_t lds_bytes);
void launch_lds_batch(int prec, unsigned nblocks, size_t lds_bytes, const BInst* inst, const BRec* gates,
                      const uint64_t* tz, const double* tc, const int32_t* order, double* energies,
                      void* out_states, hipStream_t st);
void launch_zexpect(int prec, const void* psi, uint64_t S, const uint64_t* tz, const double* tc, uint64_t nterms,
                    double* partials, double* out, hipStream_t st);
// Energies of `count` states of S amplitudes back to back: trange[2y..2y+1]
// = term range of state y, partials >= count * blocks_per_state doubles,
// out[dest[y]] = <H> of state y.
void launch_zexpect_group(int prec, const void* base, uint64_t S, unsigned count, const uint64_t* tz,
                          const double* tc, const uint64_t* trange, double* partials, unsigned blocks_per_state,
                          const uint64_t* dest, double* out, hipStream_t st);

// Shots for many small states stored back to back in an arena; norms
// (optional) receives each state's FP64 sum of |a|^2.
struct SampleJob {
  uint64_t state_offset;   // amplitude offset in the arena
  uint64_t out_offset;     // outcome offset
  uint64_t shots;
  uint64_t seed;
  int32_t n;
  int32_t pad;
};
void launch_batch_sample(int prec, const void* arena, const void* jobs, unsigned njobs, uint64_t* outcomes,
                         double* norms, hipStream_t st);
// Shared with the linear-algebra module (qsv_la.hip): set the calling
// thread's error message and return `code`; make sure the runtime is
// initialised; reach a state's amplitude buffer and stream.
int api_fail(int code, const char* msg);
int api_ensure_init();
int api_state_raw(const struct ::qsv_state* s, void** amps, hipStream_t* st, int* n, int* prec);

}  // namespace qsv
