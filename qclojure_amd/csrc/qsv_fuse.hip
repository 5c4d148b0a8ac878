// qsv_fuse.hip — fused tiled passes: many gates per read/write of the state.
//
// A pass owns a set of TB "tile" bits (12 for complex128, 13 for complex64,
// 64 KiB of LDS either way). The 2^(n-TB) tiles are independent: a workgroup
// of 256 threads loads one tile (each thread holds 2^RB = 2^(TB-8)
// amplitudes in registers), applies the pass's gates and stores it back, so
// the pass moves the state through HBM once however many gates it holds.
//
// Inside a tile, a layout (a register mask of RB tile bits) decides which
// bits a thread holds in registers: local index l = pdep(k, regmask) |
// pdep(t, ~regmask) for register slot k of thread t. A dense 2x2 or a swap
// needs its target bits in the register mask; controls, diagonal phases and
// blockade phases act on any bits (per-slot predicates). Switching layouts
// between stages is one LDS transpose. Tile bits 0..L-1 are index bits
// 0..L-1, so the lanes of a wave address whole 128 B runs of global memory.
#include "qsv_internal.h"

namespace qsv {

template <typename R> struct FV;
template <> struct FV<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct FV<float> { typedef float T __attribute__((ext_vector_type(2))); };

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

template <typename T, typename R>
__device__ __forceinline__ void cmad_pair(const double* m, T& a, T& b)
{
  const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y0, y1;
  y0.x = (R)(m[0] * ax - m[1] * ay + m[2] * bx - m[3] * by);
  y0.y = (R)(m[0] * ay + m[1] * ax + m[2] * by + m[3] * bx);
  y1.x = (R)(m[4] * ax - m[5] * ay + m[6] * bx - m[7] * by);
  y1.y = (R)(m[4] * ay + m[5] * ax + m[6] * by + m[7] * bx);
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
template <int RBIT, int NREG, typename T, typename R>
__device__ __forceinline__ void reg_dense(T* x, const double* m, uint32_t kc)
{
#pragma unroll
  for (int k = 0; k < NREG; ++k) {
    if (!(k & (1 << RBIT))) {
      if ((k & kc) == kc) cmad_pair<T, R>(m, x[k], x[k | (1 << RBIT)]);
    }
  }
}

// Exchange of register bits RA, RB_ (times ph); slot pairs act when (k & kc) == kc.
template <int RA, int RB_, int NREG, typename T, typename R>
__device__ __forceinline__ void reg_swap(T* x, double pr, double pi, uint32_t kc)
{
#pragma unroll
  for (int k = 0; k < NREG; ++k) {
    if ((k & (1 << RA)) && !(k & (1 << RB_))) {
      if ((k & kc) == kc) {
        const int k2 = (k ^ (1 << RA)) | (1 << RB_);
        const T a = x[k], b = x[k2];
        x[k] = cscale<T, R>(pr, pi, b);
        x[k2] = cscale<T, R>(pr, pi, a);
      }
    }
  }
}

template <int NREG, typename T, typename R>
__device__ __forceinline__ void dispatch_dense(int rb, T* x, const double* m, uint32_t kc)
{
  switch (rb) {
  case 0: reg_dense<0, NREG, T, R>(x, m, kc); break;
  case 1: reg_dense<1, NREG, T, R>(x, m, kc); break;
  case 2: reg_dense<2, NREG, T, R>(x, m, kc); break;
  case 3: reg_dense<3, NREG, T, R>(x, m, kc); break;
  default:
    if constexpr (NREG > 16) reg_dense<4, NREG, T, R>(x, m, kc);
    break;
  }
}

template <int NREG, typename T, typename R>
__device__ __forceinline__ void dispatch_swap(int ra, int rb, T* x, double pr, double pi, uint32_t kc)
{
  if (ra < rb) { const int t = ra; ra = rb; rb = t; }  // reg_swap visits bit ra = 1, rb = 0; symmetric
#define QSV_SW(A, B) case (A * 8 + B): reg_swap<A, B, NREG, T, R>(x, pr, pi, kc); break;
  switch (ra * 8 + rb) {
    QSV_SW(1, 0) QSV_SW(2, 0) QSV_SW(2, 1) QSV_SW(3, 0) QSV_SW(3, 1) QSV_SW(3, 2)
  default:
    if constexpr (NREG > 16) {
      switch (ra * 8 + rb) {
        QSV_SW(4, 0) QSV_SW(4, 1) QSV_SW(4, 2) QSV_SW(4, 3)
      default: break;
      }
    }
    break;
  }
#undef QSV_SW
}

template <typename R, int TB>
__global__ __launch_bounds__(256) void k_tile(typename FV<R>::T* __restrict__ psi, const FPass* __restrict__ pass)
{
  using T = typename FV<R>::T;
  constexpr int RB = TB - 8;
  constexpr int NREG = 1 << RB;
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

  // per-layout offsets: tl = thread part of the local index, rl[r] = local
  // bit of register bit r
  uint32_t reg = pass->load_mask;
  uint32_t tl, rl[RB];
  auto set_layout = [&](uint32_t mask) {
    tl = pdep32(t, FULL & ~mask);
    uint32_t m = mask;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      rl[r] = m & (~m + 1u);
      m &= m - 1;
    }
  };
  auto kl = [&](int k) -> uint32_t {
    uint32_t v = tl;
#pragma unroll
    for (int r = 0; r < RB; ++r)
      if (k & (1 << r)) v |= rl[r];
    return v;
  };
  auto goff = [&](uint32_t l) -> uint64_t {
    uint64_t g = 0;
#pragma unroll
    for (int i = 0; i < TB; ++i)
      if (l & (1u << i)) g |= gbit[i];
    return g;
  };

  T x[NREG];
  set_layout(reg);
#pragma unroll
  for (int k = 0; k < NREG; ++k) x[k] = __builtin_nontemporal_load(psi + (base | goff(kl(k))));

  const int nst = pass->nstages;
  for (int s = 0; s < nst; ++s) {
    const FStage st = pass->stages[s];
    if (st.regmask != reg) {
#pragma unroll
      for (int k = 0; k < NREG; ++k) lds[swz(kl(k))] = x[k];
      __syncthreads();
      reg = st.regmask;
      set_layout(reg);
#pragma unroll
      for (int k = 0; k < NREG; ++k) x[k] = lds[swz(kl(k))];
      __syncthreads();
    }
    for (int gi = st.g0; gi < st.g1; ++gi) {
      const FGate* g = &pass->gates[gi];
      const int kind = g->kind;
      const uint32_t cm = g->cmask;
      if (kind == FG_DENSE || kind == FG_SWAP) {
        // register-resident controls become a slot mask, the others a
        // per-thread predicate
        if ((tl & (cm & ~reg)) != (cm & ~reg)) continue;
        uint32_t kc = 0;
#pragma unroll
        for (int r = 0; r < RB; ++r)
          if (cm & rl[r]) kc |= 1u << r;
        if (kind == FG_DENSE) {
          double mm[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) mm[i] = g->m[i];
          dispatch_dense<NREG, T, R>(__popc(reg & ((1u << g->a) - 1u)), x, mm, kc);
        } else {
          dispatch_swap<NREG, T, R>(__popc(reg & ((1u << g->a) - 1u)), __popc(reg & ((1u << g->b) - 1u)), x,
                                    g->m[0], g->m[1], kc);
        }
      } else if (kind == FG_PHASE) {
        const uint32_t tb = 1u << g->a;
        const double d0r = g->m[0], d0i = g->m[1], d1r = g->m[6], d1i = g->m[7];
        const bool skip0 = g->phase_only != 0;
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
          const uint32_t l = kl(k);
          if ((l & cm) == cm) {
            if (l & tb) x[k] = cscale<T, R>(d1r, d1i, x[k]);
            else if (!skip0) x[k] = cscale<T, R>(d0r, d0i, x[k]);
          }
        }
      } else {  // FG_BLOCKADE: phase where exactly one bit of bmask is set
        const uint32_t bm = g->bmask;
        const double pr = g->m[0], pi = g->m[1];
#pragma unroll
        for (int k = 0; k < NREG; ++k)
          if (__popc(kl(k) & bm) == 1) x[k] = cscale<T, R>(pr, pi, x[k]);
      }
    }
  }

  if (reg != pass->store_mask) {
#pragma unroll
    for (int k = 0; k < NREG; ++k) lds[swz(kl(k))] = x[k];
    __syncthreads();
    reg = pass->store_mask;
    set_layout(reg);
#pragma unroll
    for (int k = 0; k < NREG; ++k) x[k] = lds[swz(kl(k))];
  }
#pragma unroll
  for (int k = 0; k < NREG; ++k) __builtin_nontemporal_store(x[k], psi + (base | goff(kl(k))));
}

void launch_tile(int prec, void* psi, const FPass* dpass, uint64_t ntiles, hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_tile<double, kTileBitsC128>), dim3((unsigned)ntiles), dim3(256), 0, st,
                       (FV<double>::T*)psi, dpass);
  else
    hipLaunchKernelGGL((k_tile<float, kTileBitsC64>), dim3((unsigned)ntiles), dim3(256), 0, st,
                       (FV<float>::T*)psi, dpass);
}

}  // namespace qsv
