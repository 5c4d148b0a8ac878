// qsv_pass.hip — fused tiled passes as host-compiled micro-op programs.
//
// A pass owns TB = 12 tile bits of the index. Each workgroup of 256 threads
// loads one tile of 2^12 amplitudes (16 per thread, in registers), runs the
// pass's micro-ops and stores the tile back, so the state crosses HBM once per
// pass however many gates the pass holds.
//
// The layout of a tile in registers is a register mask `reg` of 4 tile-local
// bits: thread t, slot k holds local index pdep(t, ~reg) | pdep(k, reg). The
// host resolves every gate against the layout it will run in (qsv_api.hip,
// compile_pass), so a micro-op carries the register bit of its target, the
// slot mask of its register-bit controls and the thread-id mask of its
// thread-bit controls: the kernel decodes nothing, takes only uniform
// branches, and a lane whose controls are not set applies the identity by
// selects instead of a divergent skip.
//
// Changing the layout is one LDS round trip (a RELAYOUT op). When the two
// wave-id bits of the thread index map to the same tile bits before and after,
// every wave keeps its own amplitudes and the round trip needs no workgroup
// barrier; the host marks those. complex128 tiles cross LDS one plane (real,
// then imaginary) at a time, so a workgroup needs 32 KiB of LDS.
//
// Arithmetic type A: FP64 for complex128; FP32 or FP64 for complex64. With FP32
// the pass also reports per-tile norms and applies the previous pass's
// renormalisation scale at load (TileNorm in qsv_internal.h).
#include "qsv_internal.h"

namespace qsv {

namespace {

template <typename R> struct PV;
template <> struct PV<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct PV<float> { typedef float T __attribute__((ext_vector_type(2))); };
// One plane (real or imaginary parts) of a thread's 16 slots: a single vector
// value indexed by constants only, so it never leaves registers.
template <typename R> struct SV;
template <> struct SV<double> { typedef double T __attribute__((ext_vector_type(16))); };
template <> struct SV<float> { typedef float T __attribute__((ext_vector_type(16))); };

constexpr int kSlots = 16;
constexpr int kThreads = 256;

// f(k) for k = K .. kSlots-1 with k a compile-time constant: the slot
// arrays must only ever be indexed by constants, or they leave registers.
template <int K = 0, typename F>
__device__ __forceinline__ void each_slot(F&& f)
{
  if constexpr (K < kSlots) {
    f(std::integral_constant<int, K>{});
    each_slot<K + 1>(f);
  }
}

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
// rewrites the low 4 bits, XOR-linear in the index.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ (((l >> 4) ^ (l >> 8)) & 15u); }

template <typename A>
struct Mat {
  A m0, m1, m2, m3, m4, m5, m6, m7;
};

// slot K of (xr, xi) *= (re, im)
template <int K, typename A, typename V>
__device__ __forceinline__ void cmul_slot(A re, A im, V& xr, V& xi)
{
  const A ax = xr[K], ay = xi[K];
  xr[K] = re * ax - im * ay;
  xi[K] = re * ay + im * ax;
}

// (a, b) = M (a, b) for the structure DM.
template <int DM, typename A, typename R>
__device__ __forceinline__ void pair_op(const Mat<A>& m, R& ar, R& ai, R& br, R& bi)
{
  const A ax = ar, ay = ai, bx = br, by = bi;
  if constexpr (DM == PM_EXCHANGE) {
    ar = (R)bx; ai = (R)by;
    br = (R)ax; bi = (R)ay;
  } else if constexpr (DM == PM_REAL) {
    ar = (R)(m.m0 * ax + m.m2 * bx);
    ai = (R)(m.m0 * ay + m.m2 * by);
    br = (R)(m.m4 * ax + m.m6 * bx);
    bi = (R)(m.m4 * ay + m.m6 * by);
  } else if constexpr (DM == PM_RX) {
    // real diagonal, imaginary off-diagonal: m01 = i m3, m10 = i m5
    ar = (R)(m.m0 * ax - m.m3 * by);
    ai = (R)(m.m0 * ay + m.m3 * bx);
    br = (R)(m.m6 * bx - m.m5 * ay);
    bi = (R)(m.m6 * by + m.m5 * ax);
  } else {
    ar = (R)(m.m0 * ax - m.m1 * ay + m.m2 * bx - m.m3 * by);
    ai = (R)(m.m0 * ay + m.m1 * ax + m.m2 * by + m.m3 * bx);
    br = (R)(m.m4 * ax - m.m5 * ay + m.m6 * bx - m.m7 * by);
    bi = (R)(m.m4 * ay + m.m5 * ax + m.m6 * by + m.m7 * bx);
  }
}

// Dense 2x2 on register bit RB. KC: only slot pairs whose register-bit
// controls kc are set (a uniform test per pair); without KC every pair acts
// and no per-pair branch is generated (a skipped pair would make the compiler
// copy its inputs into the outputs' registers). With LS, lanes where `act` is
// false keep their values (selects).
template <int RB, int DM, bool LS, bool KC, typename A, typename R>
__device__ __forceinline__ void op_dense(typename SV<R>::T& xr, typename SV<R>::T& xi, const Mat<A>& m, uint32_t kc, bool act)
{
  each_slot([&](auto kk) __attribute__((always_inline)) {
    constexpr int k = decltype(kk)::value;
    constexpr int d = 1 << RB;
    if constexpr (!(k & d)) {
      if (KC && (k & kc) != kc) return;
      R ar = xr[k], ai = xi[k], br = xr[k | d], bi = xi[k | d];
      pair_op<DM, A, R>(m, ar, ai, br, bi);
      if constexpr (LS) {
        ar = act ? ar : xr[k];
        ai = act ? ai : xi[k];
        br = act ? br : xr[k | d];
        bi = act ? bi : xi[k | d];
      }
      xr[k] = ar;
      xi[k] = ai;
      xr[k | d] = br;
      xi[k | d] = bi;
    }
  });
}

template <int RB, bool LS, bool KC, typename A, typename R>
__device__ __forceinline__ void dense_dm(typename SV<R>::T& xr, typename SV<R>::T& xi, int dm, const Mat<A>& m, uint32_t kc,
                                         bool act)
{
  if (dm == PM_GENERAL) op_dense<RB, PM_GENERAL, LS, KC, A, R>(xr, xi, m, kc, act);
  else if (dm == PM_REAL) op_dense<RB, PM_REAL, LS, KC, A, R>(xr, xi, m, kc, act);
  else if (dm == PM_RX) op_dense<RB, PM_RX, LS, KC, A, R>(xr, xi, m, kc, act);
  else op_dense<RB, PM_EXCHANGE, LS, KC, A, R>(xr, xi, m, kc, act);
}

template <int RB, typename A, typename R>
__device__ __forceinline__ void dense_rb(typename SV<R>::T& xr, typename SV<R>::T& xi, int dm, bool ls, const Mat<A>& m,
                                         uint32_t kc, bool act)
{
  if (kc == 0) {
    if (ls) dense_dm<RB, true, false, A, R>(xr, xi, dm, m, kc, act);
    else dense_dm<RB, false, false, A, R>(xr, xi, dm, m, kc, act);
  } else {
    if (ls) dense_dm<RB, true, true, A, R>(xr, xi, dm, m, kc, act);
    else dense_dm<RB, false, true, A, R>(xr, xi, dm, m, kc, act);
  }
}

// Diagonal diag(d0, d1) on register bit RB (ONLY1: d0 == 1, slots with the bit
// clear are left alone); KC: slots need their register-bit controls kc.
template <int RB, bool ONLY1, bool KC, typename A, typename R>
__device__ __forceinline__ void op_phase_reg(typename SV<R>::T& xr, typename SV<R>::T& xi, uint32_t kc, A d0r, A d0i, A d1r,
                                             A d1i)
{
  each_slot([&](auto kk) __attribute__((always_inline)) {
    constexpr int k = decltype(kk)::value;
    if (KC && (k & kc) != kc) return;
    if constexpr ((k >> RB) & 1) cmul_slot<k>(d1r, d1i, xr, xi);
    else if constexpr (!ONLY1) cmul_slot<k>(d0r, d0i, xr, xi);
  });
}

template <bool ONLY1, bool KC, typename A, typename R>
__device__ __forceinline__ void phase_rb(typename SV<R>::T& xr, typename SV<R>::T& xi, int rb, uint32_t kc, A d0r, A d0i,
                                         A d1r, A d1i)
{
  if (rb == 0) op_phase_reg<0, ONLY1, KC, A, R>(xr, xi, kc, d0r, d0i, d1r, d1i);
  else if (rb == 1) op_phase_reg<1, ONLY1, KC, A, R>(xr, xi, kc, d0r, d0i, d1r, d1i);
  else if (rb == 2) op_phase_reg<2, ONLY1, KC, A, R>(xr, xi, kc, d0r, d0i, d1r, d1i);
  else op_phase_reg<3, ONLY1, KC, A, R>(xr, xi, kc, d0r, d0i, d1r, d1i);
}

// Fixed-order sum of v over the workgroup (valid in thread 0); red[] must
// hold one double per wave.
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

template <typename V>
__device__ __forceinline__ double norm2_slots(const V& xr, const V& xi)
{
  double s = 0.0;
  each_slot([&](auto kk) __attribute__((always_inline)) {
    constexpr int k = decltype(kk)::value;
    const double re = xr[k], im = xi[k];
    s += re * re + im * im;
  });
  return s;
}

// Workgroup-scope LDS ordering, or wave-scope when only the wave's own
// addresses are involved.
__device__ __forceinline__ void lds_sync(bool wave_local)
{
  if (wave_local) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// A wave-uniform value pinned to scalar registers: a per-lane choice between
// two uniform loads would otherwise become a per-lane (vector) load.
__device__ __forceinline__ double uniform(double v)
{
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float uniform(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// Coefficient i of an op in the arithmetic type.
template <typename A> __device__ __forceinline__ A coef(const POp& op, int i);
template <> __device__ __forceinline__ double coef<double>(const POp& op, int i) { return op.m[i]; }
template <> __device__ __forceinline__ float coef<float>(const POp& op, int i) { return op.f[i]; }

}  // namespace

// One tile of one pass per workgroup. GROUP: many states of S amplitudes back
// to back; instance blockIdx.y runs the program at byte offset offs[y] of
// `prog` (negative: no pass this round); group passes do not renormalise.
template <typename R, typename A, bool GROUP>
__global__ __launch_bounds__(kThreads, sizeof(A) == 4 ? 5 : sizeof(R) == 4 ? 4 : 3) void k_pass(typename PV<R>::T* __restrict__ psi,
                                                                           const PProg* __restrict__ prog,
                                                                           TileNorm nb, uint64_t S,
                                                                           const int64_t* __restrict__ offs)
{
  using T = typename PV<R>::T;
  constexpr int TB = kPassTileBits;
  constexpr uint32_t FULL = (1u << TB) - 1u;
  // complex128 tiles cross LDS one FP64 plane at a time; complex64 tiles as
  // (re, im) pairs
  constexpr bool SPLIT = sizeof(T) == 16;
  constexpr uint32_t E = SPLIT ? sizeof(R) : sizeof(T);
  // the tile's LDS; also the per-wave partial sums of the norm reductions,
  // which run when no relayout is in flight
  __shared__ __align__(16) char smem[E << TB];
  double* red = reinterpret_cast<double*>(smem);
  if constexpr (GROUP) {
    const int64_t off = offs[blockIdx.y];
    if (off < 0) return;
    psi += (uint64_t)blockIdx.y * S;
    prog = reinterpret_cast<const PProg*>(reinterpret_cast<const char*>(prog) + off);
  }
  // The dispatcher hands consecutive workgroups to the 8 XCDs in turn: map
  // them so that each XCD streams one contiguous eighth of the tiles (its own
  // L2 and translation caches then see neighbouring addresses)
  const uint32_t g = gridDim.x, b = blockIdx.x;
  const uint32_t tile = (g & 7u) ? b : (b & 7u) * (g >> 3) + (b >> 3);
  const uint32_t t = threadIdx.x;
  const PHead& h = prog->h;
  const PLayout* lays = reinterpret_cast<const PLayout*>(reinterpret_cast<const char*>(prog) + h.lay_off);

  // tile base: zeros inserted at the tile's global bits (ascending)
  uint64_t base = tile;
  for (int i = 0; i < TB; ++i) {
    const int p = h.sorted_pos[i];
    const uint64_t lo = base & ((1ull << p) - 1ull);
    base = ((base ^ lo) << 1) | lo;
  }

  static_assert(E == 8, "the host resolves LDS offsets for 8-byte elements");
  uint32_t reg = h.load_mask;   // register mask of the current layout
  int cur = h.load_lay;         // its descriptor
  // per-thread parts of a layout's offsets (the slot parts are uniform)
  auto lds_base = [&](const PLayout& L) __attribute__((always_inline)) -> uint32_t {
    uint32_t b = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) b ^= ((t >> j) & 1u) ? L.lt[j] : 0u;
    return b;
  };
  auto gaddr = [&](const PLayout& L, uint64_t (&ga)[kSlots]) __attribute__((always_inline)) {
    uint64_t g = base;
#pragma unroll
    for (int j = 0; j < 8; ++j) g |= ((t >> j) & 1u) ? L.gt[j] : 0ull;
    each_slot([&](auto kk) __attribute__((always_inline)) {
      constexpr int k = decltype(kk)::value;
      ga[k] = g | L.gs[k];
    });
  };
  auto st_r = [&](uint32_t off, R v) __attribute__((always_inline)) { *reinterpret_cast<R*>(smem + off) = v; };
  auto ld_r = [&](uint32_t off) __attribute__((always_inline)) -> R { return *reinterpret_cast<const R*>(smem + off); };
  auto st_c = [&](uint32_t off, R re, R im) __attribute__((always_inline)) {
    T v;
    v.x = re;
    v.y = im;
    *reinterpret_cast<T*>(smem + off) = v;
  };
  auto ld_c = [&](uint32_t off) __attribute__((always_inline)) -> T { return *reinterpret_cast<const T*>(smem + off); };

  typename SV<R>::T xr, xi;
  bool local_pending = false;   // a wave-local LDS round trip may still be read by its wave
  // Move the tile from layout lay[from] to lay[to] through LDS; wave_local:
  // no workgroup barrier.
  auto relayout = [&](int from, int to, bool wave_local) __attribute__((always_inline)) {
    if (!wave_local && local_pending) __syncthreads();
    const PLayout& F = lays[from];
    const PLayout& G = lays[to];
    const uint32_t fb = lds_base(F), bb = lds_base(G);
    if constexpr (SPLIT) {
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_r(fb ^ F.ls[k], xr[k]);
      });
      lds_sync(wave_local);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        xr[k] = ld_r(bb ^ G.ls[k]);
      });
      lds_sync(wave_local);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_r(fb ^ F.ls[k], xi[k]);
      });
      lds_sync(wave_local);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        xi[k] = ld_r(bb ^ G.ls[k]);
      });
      lds_sync(wave_local);
    } else {
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_c(fb ^ F.ls[k], xr[k], xi[k]);
      });
      lds_sync(wave_local);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        { const T v = ld_c(bb ^ G.ls[k]); xr[k] = v.x; xi[k] = v.y; }
      });
      lds_sync(wave_local);
    }
    local_pending = wave_local;
  };
  // PO_SWAP: exchange local bits a = kc, b = aux (times the phase m[0..1])
  // where the local control bits tmask are set: every thread writes its slots
  // to permuted LDS addresses and reads its own back.
  auto swap_op = [&](const POp& op) __attribute__((always_inline)) {
    if (local_pending) __syncthreads();
    const uint32_t tm = op.tmask;
    const uint32_t ba = 1u << op.kc, bb = 1u << op.aux;
    const A pr = coef<A>(op, 0), pi = coef<A>(op, 1);
    const PLayout& L = lays[cur];
    const uint32_t sb = lds_base(L), dsw = swz(ba ^ bb) * E;
    // local index of each slot, from the register mask (swaps are rare)
    const uint32_t tl = pdep32(t, FULL & ~reg);
    uint32_t m = reg;
    const uint32_t q0 = m & (~m + 1u); m &= m - 1;
    const uint32_t q1 = m & (~m + 1u); m &= m - 1;
    const uint32_t q2 = m & (~m + 1u); m &= m - 1;
    const uint32_t q3 = m & (~m + 1u);
    auto moves = [&](int k) __attribute__((always_inline)) -> bool {
      const uint32_t l = tl | ((k & 1) ? q0 : 0u) | ((k & 2) ? q1 : 0u) | ((k & 4) ? q2 : 0u) | ((k & 8) ? q3 : 0u);
      return ((l & tm) == tm) && (((l & ba) != 0) != ((l & bb) != 0));
    };
    each_slot([&](auto kk) __attribute__((always_inline)) {
      constexpr int k = decltype(kk)::value;
      const bool mv = moves(k);
      const A ax = xr[k], ay = xi[k];
      const R yr = (R)(pr * ax - pi * ay), yi = (R)(pr * ay + pi * ax);
      xr[k] = mv ? yr : xr[k];
      xi[k] = mv ? yi : xi[k];
    });
    if constexpr (SPLIT) {
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_r(sb ^ L.ls[k] ^ (moves(k) ? dsw : 0u), xr[k]);
      });
      __syncthreads();
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        xr[k] = ld_r(sb ^ L.ls[k]);
      });
      __syncthreads();
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_r(sb ^ L.ls[k] ^ (moves(k) ? dsw : 0u), xi[k]);
      });
      __syncthreads();
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        xi[k] = ld_r(sb ^ L.ls[k]);
      });
      __syncthreads();
    } else {
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_c(sb ^ L.ls[k] ^ (moves(k) ? dsw : 0u), xr[k], xi[k]);
      });
      __syncthreads();
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        { const T v = ld_c(sb ^ L.ls[k]); xr[k] = v.x; xi[k] = v.y; }
      });
      __syncthreads();
    }
    local_pending = false;
  };

  // PO_BLOCK (FP32 arithmetic, complex64 tiles): Y = U X for the wave's
  // 16 x 64 matrix X (row = slot, column = lane) on the matrix cores, as four
  // column blocks of 16 lanes, 12 v_mfma_f32_16x16x4f32 each (three real
  // products per complex one). X crosses LDS at the current layout's own addresses of
  // the wave's amplitudes, so the round trip is wave-local. The offsets are
  // XOR-linear in the thread and slot bits: lane 16 cb + jc's slot s sits at
  // (wave part ^ jc part ^ cb part) ^ ls[s], and ls[4 a + b] = ls[4 a] ^ ls[b].
  auto block_op = [&](const POp& op) __attribute__((always_inline)) {
    if constexpr (sizeof(A) == 4 && !SPLIT) {
      typedef float f4 __attribute__((ext_vector_type(4)));
      const uint32_t l = t & 63u, kr = l >> 4, jc = l & 15u;
      const float2* tab = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(prog) + op.aux);
      float ur[4], ui[4], us[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const float2 u = tab[kk * 64 + l];
        ur[kk] = u.x;
        ui[kk] = u.y;
        us[kk] = u.x + u.y;
      }
      const PLayout& L = lays[cur];
      const uint32_t fb = lds_base(L);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        st_c(fb ^ L.ls[k], xr[k], xi[k]);
      });
      lds_sync(true);
      uint32_t pj = ((t >> 6) & 1u ? L.lt[6] : 0u) ^ ((t >> 7) & 1u ? L.lt[7] : 0u);
#pragma unroll
      for (int j = 0; j < 4; ++j) pj ^= (jc >> j) & 1u ? L.lt[j] : 0u;
      const uint32_t rin = (kr & 1u ? L.ls[1] : 0u) ^ (kr & 2u ? L.ls[2] : 0u);    // B rows 4 kk + kr
      const uint32_t rout = (kr & 1u ? L.ls[4] : 0u) ^ (kr & 2u ? L.ls[8] : 0u);   // C rows 4 kr + e
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const uint32_t pb = pj ^ (cb & 1 ? L.lt[4] : 0u) ^ (cb & 2 ? L.lt[5] : 0u);
        float br[4], bi[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const T v = ld_c(pb ^ rin ^ L.ls[4 * kk]);
          br[kk] = v.x;
          bi[kk] = v.y;
        }
        // three real products (Gauss): P1 = Ur Xr, P2 = Ui Xi,
        // P3 = (Ur + Ui)(Xr + Xi); Y = (P1 - P2) + i (P3 - P1 - P2)
        f4 p1 = {0.f, 0.f, 0.f, 0.f}, p2 = {0.f, 0.f, 0.f, 0.f}, p3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[kk], br[kk], p1, 0, 0, 0);
          p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[kk], bi[kk], p2, 0, 0, 0);
          p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(us[kk], br[kk] + bi[kk], p3, 0, 0, 0);
        }
        // element e of lane l is row 4 (l / 16) + e, column l % 16 of the block
#pragma unroll
        for (int e = 0; e < 4; ++e) st_c(pb ^ rout ^ L.ls[e], p1[e] - p2[e], p3[e] - p1[e] - p2[e]);
      }
      lds_sync(true);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        const T v = ld_c(fb ^ L.ls[k]);
        xr[k] = v.x;
        xi[k] = v.y;
      });
      local_pending = true;
    }
  };

  const int rn = h.renorm;
  {
    uint64_t ga[kSlots];
    gaddr(lays[cur], ga);
    if (rn & FP_INIT_ZERO) {
      // |0..0>: amplitude 1 at global index 0, nothing read
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        xr[k] = ga[k] == 0 ? (R)1 : (R)0;
        xi[k] = (R)0;
      });
    } else {
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        const T v = __builtin_nontemporal_load(psi + ga[k]);
        xr[k] = v.x;
        xi[k] = v.y;
      });
    }
  }
  if (rn & FP_RENORM_APPLY) {
    const A sc = (A)nb.ctl[1];
    each_slot([&](auto kk) __attribute__((always_inline)) {
      constexpr int k = decltype(kk)::value;
      xr[k] = (R)(sc * (A)xr[k]);
      xi[k] = (R)(sc * (A)xi[k]);
    });
  }
  if (rn & FP_RENORM_IN) {
    const double v = block_sum(norm2_slots(xr, xi), red);
    if (t == 0) nb.part_in[tile] = v;
    __syncthreads();
  }

  const int nops = h.nops;
  const POp* ops = prog->ops;
  for (int i = 0; i < nops; ++i) {
    const POp& op = ops[i];
    const int code = op.code;
    const uint32_t tm = op.tmask, kc = op.kc;
    const bool act = (t & tm) == tm;
    if (code < PO_PHASE) {
      // dense 2x2: code = PO_DENSE + 4 * rb + structure
      const Mat<A> m{coef<A>(op, 0), coef<A>(op, 1), coef<A>(op, 2), coef<A>(op, 3),
                     coef<A>(op, 4), coef<A>(op, 5), coef<A>(op, 6), coef<A>(op, 7)};
      const int rb = code >> 2, dm = code & 3;
      const bool ls = tm != 0;
      if (rb == 0) dense_rb<0, A, R>(xr, xi, dm, ls, m, kc, act);
      else if (rb == 1) dense_rb<1, A, R>(xr, xi, dm, ls, m, kc, act);
      else if (rb == 2) dense_rb<2, A, R>(xr, xi, dm, ls, m, kc, act);
      else dense_rb<3, A, R>(xr, xi, dm, ls, m, kc, act);
    } else if (code < PO_PHASE_T) {
      // diagonal on a register bit: code = PO_PHASE + 4 * only1 + rb
      const A one = (A)1, zero = (A)0;
      const A d0r = act ? coef<A>(op, 0) : one, d0i = act ? coef<A>(op, 1) : zero;
      const A d1r = act ? coef<A>(op, 6) : one, d1i = act ? coef<A>(op, 7) : zero;
      const int rb = (code - PO_PHASE) & 3;
      const bool only1 = ((code - PO_PHASE) & 4) != 0;
      if (kc == 0) {
        if (only1) phase_rb<true, false, A, R>(xr, xi, rb, kc, d0r, d0i, d1r, d1i);
        else phase_rb<false, false, A, R>(xr, xi, rb, kc, d0r, d0i, d1r, d1i);
      } else {
        if (only1) phase_rb<true, true, A, R>(xr, xi, rb, kc, d0r, d0i, d1r, d1i);
        else phase_rb<false, true, A, R>(xr, xi, rb, kc, d0r, d0i, d1r, d1i);
      }
    } else if (code == PO_PHASE_T) {
      // diagonal on a thread bit (aux = its thread-id bit): one factor per lane
      const bool hi = (t & op.aux) != 0;
      const bool on = act && (hi || op.only1 == 0);
      const A d0r = uniform(coef<A>(op, 0)), d0i = uniform(coef<A>(op, 1));
      const A d1r = uniform(coef<A>(op, 6)), d1i = uniform(coef<A>(op, 7));
      const A dr = on ? (hi ? d1r : d0r) : (A)1;
      const A di = on ? (hi ? d1i : d0i) : (A)0;
      if (kc == 0) {
        each_slot([&](auto kk) __attribute__((always_inline)) {
          constexpr int k = decltype(kk)::value;
          cmul_slot<k>(dr, di, xr, xi);
        });
      } else {
        each_slot([&](auto kk) __attribute__((always_inline)) {
          constexpr int k = decltype(kk)::value;
          if ((k & kc) == kc) cmul_slot<k>(dr, di, xr, xi);
        });
      }
    } else if (code == PO_BLOCKADE) {
      // phase where exactly one blockade bit is set: thread part popcount(t &
      // aux), register part popcount(k & kc)
      const uint32_t pt = __popc(t & op.aux);
      const A pr = coef<A>(op, 0), pi = coef<A>(op, 1);
      each_slot([&](auto kk) __attribute__((always_inline)) {
        constexpr int k = decltype(kk)::value;
        const bool on = pt + (uint32_t)__builtin_popcount((uint32_t)k & kc) == 1u;
        cmul_slot<k>(on ? pr : (A)1, on ? pi : (A)0, xr, xi);
      });
    } else if (code == PO_RELAYOUT) {
      relayout((int)op.tmask, (int)op.kc, op.only1 != 0);
      reg = op.aux;
      cur = (int)op.kc;
    } else if (code == PO_BLOCK) {
      block_op(op);
    } else {
      swap_op(op);
    }
  }

  if (cur != h.store_lay) relayout(cur, h.store_lay, h.store_local != 0);
  {
    uint64_t ga[kSlots];
    gaddr(lays[h.store_lay], ga);
    each_slot([&](auto kk) __attribute__((always_inline)) {
      constexpr int k = decltype(kk)::value;
      T v;
      v.x = xr[k];
      v.y = xi[k];
      __builtin_nontemporal_store(v, psi + ga[k]);
    });
  }
  if (rn & FP_RENORM_OUT) {
    // red[] overlays the tile's LDS: a workgroup relayout or swap ends with a
    // barrier after its reads; after a wave-local one, wait for every wave
    if (local_pending) __syncthreads();
    const double v = block_sum(norm2_slots(xr, xi), red);
    if (t == 0) nb.part_out[tile] = v;
  }
}

// |0..0> in every state of a group.
template <typename R>
__global__ __launch_bounds__(256) void k_fill_group(typename PV<R>::T* __restrict__ base, uint64_t total, uint64_t S)
{
  using T = typename PV<R>::T;
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

void launch_pass(int prec, bool f32, void* psi, const PProg* dprog, uint64_t ntiles, const TileNorm& nb,
                 hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_pass<double, double, false>), dim3((unsigned)ntiles), dim3(kThreads), 0, st,
                       (PV<double>::T*)psi, dprog, nb, 0, (const int64_t*)nullptr);
  else if (f32)
    hipLaunchKernelGGL((k_pass<float, float, false>), dim3((unsigned)ntiles), dim3(kThreads), 0, st,
                       (PV<float>::T*)psi, dprog, nb, 0, (const int64_t*)nullptr);
  else
    hipLaunchKernelGGL((k_pass<float, double, false>), dim3((unsigned)ntiles), dim3(kThreads), 0, st,
                       (PV<float>::T*)psi, dprog, nb, 0, (const int64_t*)nullptr);
}

void launch_pass_group(int prec, void* base, uint64_t S, unsigned count, const char* progs, const int64_t* offs,
                       hipStream_t st)
{
  // complex64 groups compute in FP64: no per-instance renormalisation needed
  const TileNorm none{nullptr, nullptr, nullptr};
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_pass<double, double, true>), dim3((unsigned)(S >> kPassTileBits), count), dim3(kThreads),
                       0, st, (PV<double>::T*)base, reinterpret_cast<const PProg*>(progs), none, S, offs);
  else
    hipLaunchKernelGGL((k_pass<float, double, true>), dim3((unsigned)(S >> kPassTileBits), count), dim3(kThreads),
                       0, st, (PV<float>::T*)base, reinterpret_cast<const PProg*>(progs), none, S, offs);
}

void launch_fill_group(int prec, void* base, uint64_t S, uint64_t count, hipStream_t st)
{
  const uint64_t total = S * count;
  const uint64_t want = (total + 255) / 256;
  const unsigned grid = (unsigned)(want < 65536 ? want : 65536);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_fill_group<double>), dim3(grid), dim3(256), 0, st, (PV<double>::T*)base, total, S);
  else
    hipLaunchKernelGGL((k_fill_group<float>), dim3(grid), dim3(256), 0, st, (PV<float>::T*)base, total, S);
}

void launch_norm_ctl(const TileNorm& nb, uint64_t ntiles, bool first, hipStream_t st)
{
  hipLaunchKernelGGL(k_norm_ctl, dim3(1), dim3(1024), 0, st, nb.part_in, nb.part_out, ntiles, first ? 1 : 0, nb.ctl);
}

}  // namespace qsv
