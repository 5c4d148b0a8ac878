// qsv_fuse.hip — fused tiled passes: many gates per read/write of the state.
//
// A pass owns a set of TB "tile" bits (12 for complex128, 13 for complex64,
// 64 KiB of LDS either way). The 2^(n-TB) tiles are independent: a workgroup
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

struct M8 {
  double m0, m1, m2, m3, m4, m5, m6, m7;
};

template <typename T, typename R>
__device__ __forceinline__ void cmad_pair(const M8& m, T& a, T& b)
{
  const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y0, y1;
  y0.x = (R)(m.m0 * ax - m.m1 * ay + m.m2 * bx - m.m3 * by);
  y0.y = (R)(m.m0 * ay + m.m1 * ax + m.m2 * by + m.m3 * bx);
  y1.x = (R)(m.m4 * ax - m.m5 * ay + m.m6 * bx - m.m7 * by);
  y1.y = (R)(m.m4 * ay + m.m5 * ax + m.m6 * by + m.m7 * bx);
  a = y0;
  b = y1;
}

template <typename T, typename R>
__device__ __forceinline__ T cscale(double re, double im, T a)
{
  const double ax = a.x, ay = a.y;
  T y;
  y.x = (R)(re * ax - im * ay);
  y.y = (R)(re * ay + im * ax);
  return y;
}

// Dense 2x2 on register bit RBIT; slot k acts when (k & kc) == kc.
template <int RBIT, typename T, typename R>
__device__ __forceinline__ void reg_dense(T (&x)[kRegs], const M8& m, uint32_t kc)
{
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (!(k & (1 << RBIT))) {
      if ((k & kc) == kc) cmad_pair<T, R>(m, x[k], x[k | (1 << RBIT)]);
    }
  }
}

// 64 KiB of LDS per workgroup leaves room for two workgroups per CU.
template <typename R, int TB>
__global__ __launch_bounds__(1 << (TB - kRegBits)) void k_tile(typename FV<R>::T* __restrict__ psi,
                                                               const FPass* __restrict__ pass)
{
  using T = typename FV<R>::T;
  constexpr uint32_t FULL = (1u << TB) - 1u;
  __shared__ T lds[1 << TB];

  const uint32_t t = threadIdx.x;
  // tile base index: zeros inserted at the tile's global bits (ascending)
  uint64_t base = blockIdx.x;
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
  set_layout(reg);
  {
    uint64_t ga[kRegs];
    gaddr(ga);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) x[k] = __builtin_nontemporal_load(psi + ga[k]);
  }

  const int nst = pass->nstages;
  for (int s = 0; s < nst; ++s) {
    const FStage st = pass->stages[s];
    if (st.regmask != reg) {
#pragma unroll
      for (int k = 0; k < kRegs; ++k) lds[swz(kl(k))] = x[k];
      __syncthreads();
      reg = st.regmask;
      set_layout(reg);
#pragma unroll
      for (int k = 0; k < kRegs; ++k) x[k] = lds[swz(kl(k))];
      __syncthreads();
    }
    for (int gi = st.g0; gi < st.g1; ++gi) {
      const FGate* g = &pass->gates[gi];
      const int kind = g->kind;
      const uint32_t cm = g->cmask;
      if (kind == FG_DENSE) {
        // register-resident controls become a slot mask, the others a
        // per-thread predicate
        const uint32_t cthr = cm & ~reg;
        if ((tl & cthr) != cthr) continue;
        const uint32_t kc = ((cm & r0) ? 1u : 0u) | ((cm & r1) ? 2u : 0u) | ((cm & r2) ? 4u : 0u) |
                            ((cm & r3) ? 8u : 0u);
        const M8 m{g->m[0], g->m[1], g->m[2], g->m[3], g->m[4], g->m[5], g->m[6], g->m[7]};
        const uint32_t tb = 1u << g->a;
        if (tb == r0) reg_dense<0, T, R>(x, m, kc);
        else if (tb == r1) reg_dense<1, T, R>(x, m, kc);
        else if (tb == r2) reg_dense<2, T, R>(x, m, kc);
        else reg_dense<3, T, R>(x, m, kc);
      } else if (kind == FG_PHASE) {
        const uint32_t tb = 1u << g->a;
        const double d0r = g->m[0], d0i = g->m[1], d1r = g->m[6], d1i = g->m[7];
        const bool skip0 = g->phase_only != 0;
#pragma unroll
        for (int k = 0; k < kRegs; ++k) {
          const uint32_t l = kl(k);
          if ((l & cm) == cm) {
            if (l & tb) x[k] = cscale<T, R>(d1r, d1i, x[k]);
            else if (!skip0) x[k] = cscale<T, R>(d0r, d0i, x[k]);
          }
        }
      } else if (kind == FG_SWAP) {
        // exchange local bits a and b (times ph) where the controls are set:
        // every thread writes its slots to permuted LDS addresses
        const uint32_t ba = 1u << g->a, bb = 1u << g->b;
        const double pr = g->m[0], pi = g->m[1];
#pragma unroll
        for (int k = 0; k < kRegs; ++k) {
          const uint32_t l = kl(k);
          const bool mv = ((l & cm) == cm) && (((l & ba) != 0) != ((l & bb) != 0));
          lds[swz(mv ? (l ^ ba ^ bb) : l)] = mv ? cscale<T, R>(pr, pi, x[k]) : x[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kRegs; ++k) x[k] = lds[swz(kl(k))];
        __syncthreads();
      } else {  // FG_BLOCKADE: phase where exactly one bit of bmask is set
        const uint32_t bm = g->bmask;
        const double pr = g->m[0], pi = g->m[1];
#pragma unroll
        for (int k = 0; k < kRegs; ++k)
          if (__popc(kl(k) & bm) == 1) x[k] = cscale<T, R>(pr, pi, x[k]);
      }
    }
  }

  if (reg != pass->store_mask) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) lds[swz(kl(k))] = x[k];
    __syncthreads();
    reg = pass->store_mask;
    set_layout(reg);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) x[k] = lds[swz(kl(k))];
  }
  {
    uint64_t ga[kRegs];
    gaddr(ga);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) __builtin_nontemporal_store(x[k], psi + ga[k]);
  }
}

void launch_tile(int prec, void* psi, const FPass* dpass, uint64_t ntiles, hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_tile<double, kTileBitsC128>), dim3((unsigned)ntiles), dim3(1 << (kTileBitsC128 - kRegBits)),
                       0, st, (FV<double>::T*)psi, dpass);
  else
    hipLaunchKernelGGL((k_tile<float, kTileBitsC64>), dim3((unsigned)ntiles), dim3(1 << (kTileBitsC64 - kRegBits)),
                       0, st, (FV<float>::T*)psi, dpass);
}

}  // namespace qsv
