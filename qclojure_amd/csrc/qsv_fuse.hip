// qsv_fuse.hip — fused tiled passes: many gates per read/write of the state.
//
// A pass owns a set of TB = 12 "tile" bits. A tile crosses LDS in 32 KiB: a
// complex64 tile whole, a complex128 tile one plane (real, then imaginary) at
// a time, so three complex128 workgroups fit a CU (their register count
// allows three) and four complex64 ones with FP32 arithmetic. The 2^(n-TB)
// tiles are independent: a workgroup
// of 2^(TB-4) threads loads one tile, each thread holding 16 amplitudes in
// registers, applies the pass's gates and stores it back, so the pass moves
// the state through HBM once however many gates it holds.
//
// Inside a tile, a layout (a register mask of 4 tile bits) decides which bits
// a thread holds in registers: local index l = pdep(k, regmask) |
// pdep(t, ~regmask) for register slot k of thread t. A dense 2x2 needs its
// target bit in the register mask; controls, diagonal phases and blockade
// phases act on any bits (per-slot predicates). Switching layouts between
// stages is one LDS transpose; a swap-type gate is one LDS round trip whose
// write addresses are permuted. Tile bits 0..L-1 are index bits 0..L-1, so the
// lanes of a wave address whole 128 B runs of global memory.
//
// Arithmetic type A: FP64 for complex128; FP32 or FP64 for complex64. With
// FP32 the pass also reports per-tile norms and applies the previous pass's
// renormalisation scale at load (TileNorm in qsv_internal.h).
#include "qsv_internal.h"

namespace qsv {

template <typename R> struct FV;
template <> struct FV<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct FV<float> { typedef float T __attribute__((ext_vector_type(2))); };

constexpr int kRegBits = 4;
constexpr int kRegs = 1 << kRegBits;

__device__ __forceinline__ uint32_t pdep32(uint32_t src, uint32_t mask)
{
  uint32_t out = 0;
  for (uint32_t m = mask; m; m &= m - 1) {
    if (src & 1u) out |= m & (~m + 1u);
    src >>= 1;
  }
  return out;
}

// Bank-spreading swizzle of a tile-local index: a bijection that only
// rewrites the low 4 bits.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ (((l >> 4) ^ (l >> 8)) & 15u); }

template <typename A>
struct M8 {
  A m0, m1, m2, m3, m4, m5, m6, m7;
};

template <typename T, typename R, typename A>
__device__ __forceinline__ void cmad_pair(const M8<A>& m, T& a, T& b)
{
  const A ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y0, y1;
  y0.x = (R)(m.m0 * ax - m.m1 * ay + m.m2 * bx - m.m3 * by);
  y0.y = (R)(m.m0 * ay + m.m1 * ax + m.m2 * by + m.m3 * bx);
  y1.x = (R)(m.m4 * ax - m.m5 * ay + m.m6 * bx - m.m7 * by);
  y1.y = (R)(m.m4 * ay + m.m5 * ax + m.m6 * by + m.m7 * bx);
  a = y0;
  b = y1;
}

template <typename T, typename R, typename A>
__device__ __forceinline__ T cscale(A re, A im, T a)
{
  const A ax = a.x, ay = a.y;
  T y;
  y.x = (R)(re * ax - im * ay);
  y.y = (R)(re * ay + im * ax);
  return y;
}

// Real matrix (DM_REAL): each output is a real combination of a and b.
template <typename T, typename R, typename A>
__device__ __forceinline__ void rmad_pair(const M8<A>& m, T& a, T& b)
{
  const A ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y0, y1;
  y0.x = (R)(m.m0 * ax + m.m2 * bx);
  y0.y = (R)(m.m0 * ay + m.m2 * by);
  y1.x = (R)(m.m4 * ax + m.m6 * bx);
  y1.y = (R)(m.m4 * ay + m.m6 * by);
  a = y0;
  b = y1;
}

// Real diagonal, imaginary off-diagonal (DM_RX): m01 = i m3, m10 = i m5.
template <typename T, typename R, typename A>
__device__ __forceinline__ void xmad_pair(const M8<A>& m, T& a, T& b)
{
  const A ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y0, y1;
  y0.x = (R)(m.m0 * ax - m.m3 * by);
  y0.y = (R)(m.m0 * ay + m.m3 * bx);
  y1.x = (R)(m.m6 * bx - m.m5 * ay);
  y1.y = (R)(m.m6 * by + m.m5 * ax);
  a = y0;
  b = y1;
}

// Dense 2x2 on register bit RBIT; slot k acts when (k & kc) == kc (ALL: no
// register-bit controls, every slot acts, and no per-slot selects are
// generated). DM is the matrix structure (DM_*), known for the whole gate.
template <int RBIT, int DM, bool ALL, typename T, typename R, typename A>
__device__ __forceinline__ void reg_dense(T (&x)[kRegs], const M8<A>& m, uint32_t kc)
{
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (!(k & (1 << RBIT))) {
      if (ALL || (k & kc) == kc) {
        T& a = x[k];
        T& b = x[k | (1 << RBIT)];
        if constexpr (DM == DM_EXCHANGE) {
          const T t = a;
          a = b;
          b = t;
        } else if constexpr (DM == DM_REAL) {
          rmad_pair<T, R, A>(m, a, b);
        } else if constexpr (DM == DM_RX) {
          xmad_pair<T, R, A>(m, a, b);
        } else {
          cmad_pair<T, R, A>(m, a, b);
        }
      }
    }
  }
}

template <int RBIT, bool ALL, typename T, typename R, typename A>
__device__ __forceinline__ void reg_dense_dm(T (&x)[kRegs], const M8<A>& m, uint32_t kc, int dm)
{
  if (dm == DM_EXCHANGE) reg_dense<RBIT, DM_EXCHANGE, ALL, T, R, A>(x, m, kc);
  else if (dm == DM_REAL) reg_dense<RBIT, DM_REAL, ALL, T, R, A>(x, m, kc);
  else if (dm == DM_RX) reg_dense<RBIT, DM_RX, ALL, T, R, A>(x, m, kc);
  else reg_dense<RBIT, DM_GENERAL, ALL, T, R, A>(x, m, kc);
}

// Exchange of the slot pairs on register bit RBIT whose register-bit
// controls kc are set, in the lanes where `act` holds (thread-bit controls):
// selects only, so no lane-divergent branch.
template <int RBIT, typename T>
__device__ __forceinline__ void reg_xchg(T (&x)[kRegs], uint32_t kc, bool act)
{
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (!(k & (1 << RBIT))) {
      const bool sw = act && ((k & kc) == kc);
      const T a = x[k], b = x[k | (1 << RBIT)];
      x[k] = sw ? b : a;
      x[k | (1 << RBIT)] = sw ? a : b;
    }
  }
}

// Diagonal diag(d0, d1) on register bit RBIT (skip0: d0 == 1); slot k acts
// when (k & kc) == kc, every slot when ALL.
template <int RBIT, bool ALL, typename T, typename R, typename A>
__device__ __forceinline__ void reg_phase(T (&x)[kRegs], uint32_t kc, A d0r, A d0i, A d1r, A d1i, bool skip0)
{
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (ALL || (k & kc) == kc) {
      if (k & (1 << RBIT)) x[k] = cscale<T, R, A>(d1r, d1i, x[k]);
      else if (!skip0) x[k] = cscale<T, R, A>(d0r, d0i, x[k]);
    }
  }
}

template <bool ALL, typename T, typename R, typename A>
__device__ __forceinline__ void reg_phase_any(T (&x)[kRegs], int rb, uint32_t kc, A d0r, A d0i, A d1r, A d1i,
                                              bool skip0)
{
  if (rb == 0) reg_phase<0, ALL, T, R, A>(x, kc, d0r, d0i, d1r, d1i, skip0);
  else if (rb == 1) reg_phase<1, ALL, T, R, A>(x, kc, d0r, d0i, d1r, d1i, skip0);
  else if (rb == 2) reg_phase<2, ALL, T, R, A>(x, kc, d0r, d0i, d1r, d1i, skip0);
  else reg_phase<3, ALL, T, R, A>(x, kc, d0r, d0i, d1r, d1i, skip0);
}

template <int RBIT, typename T, typename R, typename A>
__device__ __forceinline__ void reg_dense_any(T (&x)[kRegs], const M8<A>& m, uint32_t kc, int dm)
{
  if (kc == 0) reg_dense_dm<RBIT, true, T, R, A>(x, m, kc, dm);
  else reg_dense_dm<RBIT, false, T, R, A>(x, m, kc, dm);
}

// Sum of v over the workgroup in a fixed order (valid in thread 0).
__device__ __forceinline__ double block_sum(double v, double* red)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const uint32_t t = threadIdx.x;
  if ((t & 63u) == 0) red[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) s += red[w];
  return s;
}

template <typename T>
__device__ __forceinline__ double norm2_slots(const T (&x)[kRegs])
{
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    const double re = x[k].x, im = x[k].y;
    s += re * re + im * im;
  }
  return s;
}

// One tile of one pass per workgroup: load, the pass's stages, store.
// GROUP = true: many states of S amplitudes back to back, one pass round;
// instance blockIdx.y runs the pass at byte offset offs[y] of `pass` (a
// negative offset: no pass this round). Pass descriptors of a group are stored
// truncated after their last gate. Group passes do not renormalise.
template <typename R, typename A, int TB, bool GROUP>
__global__ __launch_bounds__(1 << (TB - kRegBits), (sizeof(R) == 4 && sizeof(A) == 4) ? 4 : 3) void k_tile(typename FV<R>::T* __restrict__ psi,
                                                               const FPass* __restrict__ pass, TileNorm nb,
                                                               uint64_t S, const int64_t* __restrict__ offs)
{
  using T = typename FV<R>::T;
  constexpr uint32_t FULL = (1u << TB) - 1u;
  // complex128 tiles go through LDS one plane (real, then imaginary) at a
  // time: 32 KiB instead of 64, so three workgroups fit a CU
  constexpr bool SPLIT = sizeof(T) == 16;
  __shared__ __align__(16) char smem[SPLIT ? (sizeof(R) << TB) : (sizeof(T) << TB)];
  __shared__ double red[(1 << (TB - kRegBits)) / 64];
  if constexpr (GROUP) {
    const int64_t off = offs[blockIdx.y];
    if (off < 0) return;
    psi += (uint64_t)blockIdx.y * S;
    pass = reinterpret_cast<const FPass*>(reinterpret_cast<const char*>(pass) + off);
  }
  const uint32_t tile = blockIdx.x;

  const uint32_t t = threadIdx.x;
  // tile base index: zeros inserted at the tile's global bits (ascending)
  uint64_t base = tile;
  for (int i = 0; i < TB; ++i) {
    const int p = pass->sorted_pos[i];
    const uint64_t lo = base & ((1ull << p) - 1ull);
    base = ((base ^ lo) << 1) | lo;
  }
  uint64_t gbit[TB];  // global index bit of each tile-local bit
#pragma unroll
  for (int i = 0; i < TB; ++i) gbit[i] = 1ull << pass->pos[i];

  // layout: tl = thread part of the local index, r0..r3 = local bit of
  // register bits 0..3
  uint32_t reg = pass->load_mask;
  uint32_t tl, r0, r1, r2, r3;
  auto set_layout = [&](uint32_t mask) {
    tl = pdep32(t, FULL & ~mask);
    uint32_t m = mask;
    r0 = m & (~m + 1u); m &= m - 1;
    r1 = m & (~m + 1u); m &= m - 1;
    r2 = m & (~m + 1u); m &= m - 1;
    r3 = m & (~m + 1u);
  };
  auto kl = [&](int k) -> uint32_t {
    return tl | ((k & 1) ? r0 : 0u) | ((k & 2) ? r1 : 0u) | ((k & 4) ? r2 : 0u) | ((k & 8) ? r3 : 0u);
  };
  auto goff = [&](uint32_t l) -> uint64_t {
    uint64_t g = 0;
#pragma unroll
    for (int i = 0; i < TB; ++i)
      if (l & (1u << i)) g |= gbit[i];
    return g;
  };
  // global addresses of the 16 slots: thread part | register-bit parts
  auto gaddr = [&](uint64_t (&ga)[kRegs]) {
    const uint64_t gt = base | goff(tl);
    const uint64_t g0 = goff(r0), g1 = goff(r1), g2 = goff(r2), g3 = goff(r3);
#pragma unroll
    for (int k = 0; k < kRegs; ++k)
      ga[k] = gt | ((k & 1) ? g0 : 0ull) | ((k & 2) ? g1 : 0ull) | ((k & 4) ? g2 : 0ull) | ((k & 8) ? g3 : 0ull);
  };

  T x[kRegs];
  // LDS byte offset of slot k. swz is XOR-linear and a local index is the
  // XOR of its thread part and its register part (disjoint bits), so the
  // offset is swz(thread part) * E (one value per thread) XOR
  // swz(register part of k) * E (workgroup-uniform, scalar arithmetic):
  // one vector instruction per slot.
  constexpr uint32_t E = SPLIT ? sizeof(R) : sizeof(T);
  // (q0..q3: swz(r0..r3) * E; the register part of slot k is the XOR of
  // those of its set bits, a few scalar XORs shared by all slots)
  auto roff = [](int k, uint32_t q0, uint32_t q1, uint32_t q2, uint32_t q3) -> uint32_t {
    return ((k & 1) ? q0 : 0u) ^ ((k & 2) ? q1 : 0u) ^ ((k & 4) ? q2 : 0u) ^ ((k & 8) ? q3 : 0u);
  };
  auto rsw = [](uint32_t r) -> uint32_t { return swz(r) * E; };
  auto pl_st = [&](uint32_t off, R v) { *reinterpret_cast<R*>(smem + off) = v; };
  auto pl_ld = [&](uint32_t off) -> R { return *reinterpret_cast<const R*>(smem + off); };
  // move the tile from the current layout to layout `to` through LDS
  // The per-thread part is made opaque so that the compiler keeps the
  // scaling by E on the scalar side instead of factoring it out of the XOR.
  auto tbase = [&]() -> uint32_t {
    uint32_t b = swz(tl) * E;
    asm("" : "+v"(b));
    return b;
  };
  auto relayout = [&](uint32_t to) {
    const uint32_t fb = tbase();
    const uint32_t f0 = rsw(r0), f1 = rsw(r1), f2 = rsw(r2), f3 = rsw(r3);
    set_layout(to);
    const uint32_t tb = tbase();
    const uint32_t t0 = rsw(r0), t1 = rsw(r1), t2 = rsw(r2), t3 = rsw(r3);
    if constexpr (SPLIT) {
#pragma unroll
      for (int k = 0; k < kRegs; ++k) pl_st(fb ^ roff(k, f0, f1, f2, f3), x[k].x);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRegs; ++k) x[k].x = pl_ld(tb ^ roff(k, t0, t1, t2, t3));
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRegs; ++k) pl_st(fb ^ roff(k, f0, f1, f2, f3), x[k].y);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRegs; ++k) x[k].y = pl_ld(tb ^ roff(k, t0, t1, t2, t3));
      __syncthreads();
    } else {
#pragma unroll
      for (int k = 0; k < kRegs; ++k) *reinterpret_cast<T*>(smem + (fb ^ roff(k, f0, f1, f2, f3))) = x[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRegs; ++k) x[k] = *reinterpret_cast<const T*>(smem + (tb ^ roff(k, t0, t1, t2, t3)));
      __syncthreads();
    }
  };

  set_layout(reg);
  {
    uint64_t ga[kRegs];
    gaddr(ga);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) x[k] = __builtin_nontemporal_load(psi + ga[k]);
  }
  const int rn = pass->renorm;
  if (rn & FP_RENORM_APPLY) {
    const A sc = (A)nb.ctl[1];
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
      x[k].x = (R)(sc * (A)x[k].x);
      x[k].y = (R)(sc * (A)x[k].y);
    }
  }
  if (rn & FP_RENORM_IN) {
    const double v = block_sum(norm2_slots(x), red);
    if (t == 0) nb.part_in[tile] = v;
    __syncthreads();
  }

  const int nst = pass->nstages;
  for (int s = 0; s < nst; ++s) {
    const FStage st = pass->stages[s];
    if (st.regmask != reg) {
      relayout(st.regmask);
      reg = st.regmask;
    }
    for (int gi = st.g0; gi < st.g1; ++gi) {
      const FGate* g = &pass->gates[gi];
      const int kind = g->kind;
      const uint32_t cm = g->cmask;
      if (kind == FG_DENSE) {
        // register-resident controls become a slot mask, the others a
        // per-thread predicate. Uncontrolled gates take a path with no
        // per-thread branch, so the slots are updated in place (a lane-
        // divergent skip makes the compiler copy all 16 slots first).
        const uint32_t cthr = cm & ~reg;
        const uint32_t kc = ((cm & r0) ? 1u : 0u) | ((cm & r1) ? 2u : 0u) | ((cm & r2) ? 4u : 0u) |
                            ((cm & r3) ? 8u : 0u);
        const M8<A> m{(A)g->m[0], (A)g->m[1], (A)g->m[2], (A)g->m[3],
                      (A)g->m[4], (A)g->m[5], (A)g->m[6], (A)g->m[7]};
        const uint32_t tb = 1u << g->a;
        const int dm = g->sub;
        if (cm == 0) {
          if (tb == r0) reg_dense_dm<0, true, T, R, A>(x, m, 0u, dm);
          else if (tb == r1) reg_dense_dm<1, true, T, R, A>(x, m, 0u, dm);
          else if (tb == r2) reg_dense_dm<2, true, T, R, A>(x, m, 0u, dm);
          else reg_dense_dm<3, true, T, R, A>(x, m, 0u, dm);
        } else if (dm == DM_EXCHANGE) {
          // controlled exchange (CNOT, Toffoli): per-lane selects, no branch
          const bool act = (tl & cthr) == cthr;
          if (tb == r0) reg_xchg<0, T>(x, kc, act);
          else if (tb == r1) reg_xchg<1, T>(x, kc, act);
          else if (tb == r2) reg_xchg<2, T>(x, kc, act);
          else reg_xchg<3, T>(x, kc, act);
        } else {
          if ((tl & cthr) != cthr) continue;
          if (tb == r0) reg_dense_any<0, T, R, A>(x, m, kc, dm);
          else if (tb == r1) reg_dense_any<1, T, R, A>(x, m, kc, dm);
          else if (tb == r2) reg_dense_any<2, T, R, A>(x, m, kc, dm);
          else reg_dense_any<3, T, R, A>(x, m, kc, dm);
        }
      } else if (kind == FG_PHASE) {
        // controls and target split like a dense gate's: thread-index bits
        // decide once per thread, register bits per slot at compile time
        const uint32_t tb = 1u << g->a;
        const uint32_t cthr = cm & ~reg;
        const uint32_t kc = ((cm & r0) ? 1u : 0u) | ((cm & r1) ? 2u : 0u) | ((cm & r2) ? 4u : 0u) |
                            ((cm & r3) ? 8u : 0u);
        const A d0r = (A)g->m[0], d0i = (A)g->m[1], d1r = (A)g->m[6], d1i = (A)g->m[7];
        const bool skip0 = g->sub != 0;
        if (cm == 0 && !skip0) {
          // uncontrolled, both diagonal entries act: no per-thread branch
          if (tb & reg) {
            const int rb = (tb == r0) ? 0 : (tb == r1) ? 1 : (tb == r2) ? 2 : 3;
            reg_phase_any<true, T, R, A>(x, rb, 0u, d0r, d0i, d1r, d1i, false);
          } else {
            const bool hi = (tl & tb) != 0;
            const A dr = hi ? d1r : d0r, di = hi ? d1i : d0i;
#pragma unroll
            for (int k = 0; k < kRegs; ++k) x[k] = cscale<T, R, A>(dr, di, x[k]);
          }
          continue;
        }
        if ((tl & cthr) != cthr) continue;
        if (tb & reg) {
          const int rb = (tb == r0) ? 0 : (tb == r1) ? 1 : (tb == r2) ? 2 : 3;
          if (kc == 0) reg_phase_any<true, T, R, A>(x, rb, 0u, d0r, d0i, d1r, d1i, skip0);
          else reg_phase_any<false, T, R, A>(x, rb, kc, d0r, d0i, d1r, d1i, skip0);
        } else {
          const bool hi = (tl & tb) != 0;
          if (!hi && skip0) continue;
          const A dr = hi ? d1r : d0r, di = hi ? d1i : d0i;
          if (kc == 0) {
#pragma unroll
            for (int k = 0; k < kRegs; ++k) x[k] = cscale<T, R, A>(dr, di, x[k]);
          } else {
#pragma unroll
            for (int k = 0; k < kRegs; ++k)
              if ((k & kc) == kc) x[k] = cscale<T, R, A>(dr, di, x[k]);
          }
        }
      } else if (kind == FG_SWAP) {
        // exchange local bits a and b (times ph) where the controls are set:
        // every thread writes its slots to permuted LDS addresses
        const uint32_t ba = 1u << g->a, bb = 1u << g->b;
        const A pr = (A)g->m[0], pi = (A)g->m[1];
        // a moving slot is stored at its partner's offset: swz(l ^ ba ^ bb)
        // = swz(l) ^ swz(ba ^ bb)
        const uint32_t sb = tbase(), dsw = swz(ba ^ bb) * E;
        const uint32_t q0 = rsw(r0), q1 = rsw(r1), q2 = rsw(r2), q3 = rsw(r3);
        auto moves = [&](int k) -> bool {
          const uint32_t l = kl(k);
          return ((l & cm) == cm) && (((l & ba) != 0) != ((l & bb) != 0));
        };
        if constexpr (SPLIT) {
          // real plane, then imaginary plane
#pragma unroll
          for (int k = 0; k < kRegs; ++k) {
            const bool mv = moves(k);
            if (mv) x[k] = cscale<T, R, A>(pr, pi, x[k]);
            pl_st(sb ^ roff(k, q0, q1, q2, q3) ^ (mv ? dsw : 0u), x[k].x);
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < kRegs; ++k) x[k].x = pl_ld(sb ^ roff(k, q0, q1, q2, q3));
          __syncthreads();
#pragma unroll
          for (int k = 0; k < kRegs; ++k) pl_st(sb ^ roff(k, q0, q1, q2, q3) ^ (moves(k) ? dsw : 0u), x[k].y);
          __syncthreads();
#pragma unroll
          for (int k = 0; k < kRegs; ++k) x[k].y = pl_ld(sb ^ roff(k, q0, q1, q2, q3));
          __syncthreads();
        } else {
#pragma unroll
          for (int k = 0; k < kRegs; ++k) {
            const bool mv = moves(k);
            *reinterpret_cast<T*>(smem + (sb ^ roff(k, q0, q1, q2, q3) ^ (mv ? dsw : 0u))) =
                mv ? cscale<T, R, A>(pr, pi, x[k]) : x[k];
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < kRegs; ++k) x[k] = *reinterpret_cast<const T*>(smem + (sb ^ roff(k, q0, q1, q2, q3)));
          __syncthreads();
        }
      } else {  // FG_BLOCKADE: phase where exactly one bit of bmask is set
        const uint32_t bm = g->bmask;
        const A pr = (A)g->m[0], pi = (A)g->m[1];
#pragma unroll
        for (int k = 0; k < kRegs; ++k)
          if (__popc(kl(k) & bm) == 1) x[k] = cscale<T, R, A>(pr, pi, x[k]);
      }
    }
  }

  if (reg != pass->store_mask) {
    relayout(pass->store_mask);
    reg = pass->store_mask;
  }
  {
    uint64_t ga[kRegs];
    gaddr(ga);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) __builtin_nontemporal_store(x[k], psi + ga[k]);
  }
  if (rn & FP_RENORM_OUT) {
    const double v = block_sum(norm2_slots(x), red);
    if (t == 0) nb.part_out[tile] = v;
  }
}

// |0..0> in every state of a group.
template <typename R>
__global__ __launch_bounds__(256) void k_fill_group(typename FV<R>::T* __restrict__ base, uint64_t total, uint64_t S)
{
  using T = typename FV<R>::T;
  for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < total; j += (uint64_t)gridDim.x * 256) {
    T v;
    v.x = (j & (S - 1)) == 0 ? R(1) : R(0);
    v.y = R(0);
    base[j] = v;
  }
}

// One workgroup: fixed-order sums of the per-tile norms, then the scale the
// next pass applies at load.
__global__ __launch_bounds__(1024) void k_norm_ctl(const double* __restrict__ part_in,
                                                   const double* __restrict__ part_out, uint64_t ntiles, int first,
                                                   double* __restrict__ ctl)
{
  __shared__ double sa[1024], sb[1024];
  const uint32_t t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (uint64_t i = t; i < ntiles; i += 1024) {
    if (first) a += part_in[i];
    b += part_out[i];
  }
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (uint32_t w = 512; w > 0; w >>= 1) {
    if (t < w) {
      sa[t] += sa[t + w];
      sb[t] += sb[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double target = first ? sa[0] : ctl[0];
    ctl[0] = target;
    ctl[1] = (target > 0.0 && sb[0] > 0.0) ? sqrt(target / sb[0]) : 1.0;
  }
}

void launch_tile(int prec, bool f32, void* psi, const FPass* dpass, uint64_t ntiles, const TileNorm& nb,
                 hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_tile<double, double, kTileBitsC128, false>), dim3((unsigned)ntiles),
                       dim3(1 << (kTileBitsC128 - kRegBits)), 0, st, (FV<double>::T*)psi, dpass, nb, 0,
                       (const int64_t*)nullptr);
  else if (f32)
    hipLaunchKernelGGL((k_tile<float, float, kTileBitsC64, false>), dim3((unsigned)ntiles),
                       dim3(1 << (kTileBitsC64 - kRegBits)), 0, st, (FV<float>::T*)psi, dpass, nb, 0,
                       (const int64_t*)nullptr);
  else
    hipLaunchKernelGGL((k_tile<float, double, kTileBitsC64, false>), dim3((unsigned)ntiles),
                       dim3(1 << (kTileBitsC64 - kRegBits)), 0, st, (FV<float>::T*)psi, dpass, nb, 0,
                       (const int64_t*)nullptr);
}

void launch_tile_group(int prec, void* base, uint64_t S, unsigned count, const char* passes, const int64_t* offs,
                       hipStream_t st)
{
  // complex64 groups compute in FP64: no per-instance renormalisation needed
  const TileNorm none{nullptr, nullptr, nullptr};
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_tile<double, double, kTileBitsC128, true>), dim3((unsigned)(S >> kTileBitsC128), count),
                       dim3(1 << (kTileBitsC128 - kRegBits)), 0, st, (FV<double>::T*)base,
                       reinterpret_cast<const FPass*>(passes), none, S, offs);
  else
    hipLaunchKernelGGL((k_tile<float, double, kTileBitsC64, true>), dim3((unsigned)(S >> kTileBitsC64), count),
                       dim3(1 << (kTileBitsC64 - kRegBits)), 0, st, (FV<float>::T*)base,
                       reinterpret_cast<const FPass*>(passes), none, S, offs);
}

void launch_fill_group(int prec, void* base, uint64_t S, uint64_t count, hipStream_t st)
{
  const uint64_t total = S * count;
  const uint64_t want = (total + 255) / 256;
  const unsigned grid = (unsigned)(want < 65536 ? want : 65536);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_fill_group<double>), dim3(grid), dim3(256), 0, st, (FV<double>::T*)base, total, S);
  else
    hipLaunchKernelGGL((k_fill_group<float>), dim3(grid), dim3(256), 0, st, (FV<float>::T*)base, total, S);
}

void launch_norm_ctl(const TileNorm& nb, uint64_t ntiles, bool first, hipStream_t st)
{
  hipLaunchKernelGGL(k_norm_ctl, dim3(1), dim3(1024), 0, st, nb.part_in, nb.part_out, ntiles, first ? 1 : 0, nb.ctl);
}

}  // namespace qsv
