// qsv_gates.hip — per-gate streaming kernels (one launch per gate).
//
// A gate's bits are split at a threshold `thr` (3 for complex128, 4 for
// complex64: the smallest bit whose 2^thr-amplitude runs fill a 128 B line):
//   * high bits (>= thr) are enumerated by zero-bit insertion, so a thread
//     owns the 2 (or 1) amplitudes the gate couples across them and every
//     wave-wide load is a run of whole 128 B lines;
//   * low bits (< thr) stay inside the wavefront: index bit b is lane bit b,
//     a low target exchanges its partner with a lane shuffle
//     (__shfl_xor -> ds_swizzle/DPP), a low control is a per-lane predicate.
// Every amplitude a thread loads it also stores (unchanged where the gate is
// the identity), so HBM sees full-line reads and writes. Loads and stores are
// non-temporal: each byte of the state is touched once per gate.
// All arithmetic is FP64 (complex64 states round once per amplitude per gate).
#include <algorithm>

#include "qsv_internal.h"

namespace qsv {

template <typename R> struct NV;
template <> struct NV<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct NV<float> { typedef float T __attribute__((ext_vector_type(2))); };

template <typename T>
__device__ __forceinline__ T ldnt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void stnt(T* p, T v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ uint64_t expand_ins(uint64_t j, const InsArgs& a)
{
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if (k < a.nins) {
      const uint64_t lo = j & ((1ull << a.ins[k]) - 1ull);
      j = ((j ^ lo) << 1) | lo;
    }
  }
  return j | a.orMask;
}

template <typename T>
__device__ __forceinline__ T shfl_xor2(T v, int mask)
{
  T r;
  r.x = __shfl_xor(v.x, mask, 64);
  r.y = __shfl_xor(v.y, mask, 64);
  return r;
}

template <typename T>
__device__ __forceinline__ T shfl2(T v, int src)
{
  T r;
  r.x = __shfl(v.x, src, 64);
  r.y = __shfl(v.y, src, 64);
  return r;
}

// Workgroup -> chunk of the index space. The dispatcher hands consecutive
// workgroups to the 8 XCDs in turn; with XCD = true each XCD streams one
// contiguous eighth of the chunks instead (a bijection when gridDim.x is a
// multiple of 8, the identity otherwise).
template <bool XCD>
__device__ __forceinline__ uint64_t block_chunk()
{
  const uint32_t g = gridDim.x, b = blockIdx.x;
  if (!XCD || (g & 7u)) return b;
  return (uint64_t)(b & 7u) * (g >> 3) + (b >> 3);
}

// (m_r + i m_i) * (x + i y) + (n_r + i n_i) * (u + i v), FP64, rounded to R
template <typename T, typename R>
__device__ __forceinline__ T cmad2(double mr, double mi, T a, double nr, double ni, T b)
{
  const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
  T y;
  y.x = (R)(mr * ax - mi * ay + nr * bx - ni * by);
  y.y = (R)(mr * ay + mi * ax + nr * by + ni * bx);
  return y;
}

template <typename T, typename R>
__device__ __forceinline__ T cmul1(double dr, double di, T a)
{
  const double ax = a.x, ay = a.y;
  T y;
  y.x = (R)(dr * ax - di * ay);
  y.y = (R)(dr * ay + di * ax);
  return y;
}

// ---------------------------------------------------------------------------
// Dense 2x2 (optionally controlled). TLOW=false: target bit is high (inserted),
// each thread slot holds the pair. TLOW=true: target bit is a lane bit.
// ---------------------------------------------------------------------------
template <typename R, int PPT, bool TLOW, bool XCD>
__global__ __launch_bounds__(kBlock) void k_gate1(typename NV<R>::T* __restrict__ psi, GateArgs g, Mat2<double> M,
                                                   uint64_t c0)
{
  using T = typename NV<R>::T;
  const uint64_t base = (c0 + block_chunk<XCD>()) * (kBlock * PPT) + threadIdx.x;
  const uint64_t count = g.ins.count;
  uint64_t id[PPT];
  if constexpr (!TLOW) {
    T x0[PPT], x1[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint64_t j = base + (uint64_t)k * kBlock;
      id[k] = expand_ins(j, g.ins);
      if (j < count) {
        x0[k] = ldnt(psi + id[k]);
        x1[k] = ldnt(psi + (id[k] | g.tbit));
      }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint64_t j = base + (uint64_t)k * kBlock;
      if (j < count) {
        T y0 = x0[k], y1 = x1[k];
        if ((id[k] & g.lo_pred) == g.lo_pred) {
          y0 = cmad2<T, R>(M.m[0], M.m[1], x0[k], M.m[2], M.m[3], x1[k]);
          y1 = cmad2<T, R>(M.m[4], M.m[5], x0[k], M.m[6], M.m[7], x1[k]);
        }
        stnt(psi + id[k], y0);
        stnt(psi + (id[k] | g.tbit), y1);
      }
    }
  } else {
    T x[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint64_t j = base + (uint64_t)k * kBlock;
      id[k] = expand_ins(j, g.ins);
      x[k] = T{R(0), R(0)};
      if (j < count) x[k] = ldnt(psi + id[k]);
    }
    const int lt = (int)g.lo_t;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const T p = shfl_xor2(x[k], lt);  // all lanes take part in the exchange
      const uint64_t j = base + (uint64_t)k * kBlock;
      if (j < count) {
        T y = x[k];
        if ((id[k] & g.lo_pred) == g.lo_pred) {
          if (id[k] & g.lo_t)
            y = cmad2<T, R>(M.m[6], M.m[7], x[k], M.m[4], M.m[5], p);   // m10 p + m11 x
          else
            y = cmad2<T, R>(M.m[0], M.m[1], x[k], M.m[2], M.m[3], p);   // m00 x + m01 p
        }
        stnt(psi + id[k], y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Diagonal: a_i *= (i & tbit) ? d1 : d0 where (i & lo_pred) == lo_pred; the
// high fixed bits (controls, and the target for phase-only gates) are inserted.
// ---------------------------------------------------------------------------
template <typename R, int EPT>
__global__ __launch_bounds__(kBlock) void k_phase(typename NV<R>::T* __restrict__ psi, GateArgs g, Cplx<double> d0,
                                                   Cplx<double> d1, uint64_t c0)
{
  using T = typename NV<R>::T;
  const uint64_t base = (c0 + block_chunk<true>()) * (kBlock * EPT) + threadIdx.x;
  const uint64_t count = g.ins.count;
  T x[EPT];
  uint64_t id[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint64_t j = base + (uint64_t)k * kBlock;
    id[k] = expand_ins(j, g.ins);
    if (j < count) x[k] = ldnt(psi + id[k]);
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint64_t j = base + (uint64_t)k * kBlock;
    if (j < count) {
      T y = x[k];
      if ((id[k] & g.lo_pred) == g.lo_pred) {
        const Cplx<double> d = (id[k] & g.tbit) ? d1 : d0;
        y = cmul1<T, R>(d.re, d.im, x[k]);
      }
      stnt(psi + id[k], y);
    }
  }
}

// ---------------------------------------------------------------------------
// Swap of two bits A, B (phase ph on the exchanged amplitudes), optionally
// controlled. MODE 0: both high; 1: hA high, lB low; 2: both low.
// ---------------------------------------------------------------------------
template <typename R, int PPT, int MODE, bool XCD>
__global__ __launch_bounds__(kBlock) void k_swap2(typename NV<R>::T* __restrict__ psi, SwapArgs s, Cplx<double> ph,
                                                   uint64_t c0)
{
  using T = typename NV<R>::T;
  const uint64_t base = (c0 + block_chunk<XCD>()) * (kBlock * PPT) + threadIdx.x;
  const uint64_t count = s.ins.count;
  uint64_t id[PPT];
  T a[PPT], b[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint64_t j = base + (uint64_t)k * kBlock;
    id[k] = expand_ins(j, s.ins);
    a[k] = T{R(0), R(0)};
    b[k] = T{R(0), R(0)};
    if (j < count) {
      if constexpr (MODE == 0) {
        a[k] = ldnt(psi + (id[k] | s.hA));
        b[k] = ldnt(psi + (id[k] | s.hB));
      } else if constexpr (MODE == 1) {
        a[k] = ldnt(psi + id[k]);
        b[k] = ldnt(psi + (id[k] | s.hA));
      } else {
        a[k] = ldnt(psi + id[k]);
      }
    }
  }
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint64_t j = base + (uint64_t)k * kBlock;
    if constexpr (MODE == 0) {
      if (j < count) {
        T ya = a[k], yb = b[k];
        if ((id[k] & s.lo_pred) == s.lo_pred) {
          ya = cmul1<T, R>(ph.re, ph.im, b[k]);
          yb = cmul1<T, R>(ph.re, ph.im, a[k]);
        }
        stnt(psi + (id[k] | s.hA), ya);
        stnt(psi + (id[k] | s.hB), yb);
      }
    } else if constexpr (MODE == 1) {
      // a = amplitude with hA = 0, b = with hA = 1; lB is a lane bit.
      const T sa = shfl_xor2(a[k], (int)s.lB);
      const T sb = shfl_xor2(b[k], (int)s.lB);
      if (j < count) {
        T ya = a[k], yb = b[k];
        if ((id[k] & s.lo_pred) == s.lo_pred) {
          if (id[k] & s.lB) ya = cmul1<T, R>(ph.re, ph.im, sb);   // (lB=1,hA=0) <- (lB=0,hA=1)
          else yb = cmul1<T, R>(ph.re, ph.im, sa);                // (lB=0,hA=1) <- (lB=1,hA=0)
        }
        stnt(psi + id[k], ya);
        stnt(psi + (id[k] | s.hA), yb);
      }
    } else {
      const int ba = (lane & (int)s.lA) ? 1 : 0, bb = (lane & (int)s.lB) ? 1 : 0;
      const int src = (ba != bb) ? (lane ^ (int)(s.lA | s.lB)) : lane;
      const T p = shfl2(a[k], src);
      if (j < count) {
        T y = a[k];
        if (ba != bb && (id[k] & s.lo_pred) == s.lo_pred) y = cmul1<T, R>(ph.re, ph.im, p);
        stnt(psi + id[k], y);
      }
    }
  }
}

// Rydberg blockade: phase where exactly one of the selected bits is set.
template <typename R, int EPT>
__global__ __launch_bounds__(kBlock) void k_blockade(typename NV<R>::T* __restrict__ psi, uint64_t S, uint64_t bmask,
                                                      Cplx<double> ph, uint64_t c0)
{
  using T = typename NV<R>::T;
  const uint64_t base = (c0 + block_chunk<true>()) * (kBlock * EPT) + threadIdx.x;
  T x[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint64_t i = base + (uint64_t)k * kBlock;
    if (i < S) x[k] = ldnt(psi + i);
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint64_t i = base + (uint64_t)k * kBlock;
    if (i < S) {
      T y = x[k];
      if (__popcll(i & bmask) == 1) y = cmul1<T, R>(ph.re, ph.im, x[k]);
      stnt(psi + i, y);
    }
  }
}

// ---------------------------------------------------------------------------
static inline uint64_t blocks_for(uint64_t count, int per_thread)
{
  const uint64_t per_block = (uint64_t)kBlock * per_thread;
  return (count + per_block - 1) / per_block;
}

// A launch may hold fewer than 2^32 work-items: states of 34 qubits and more
// take several launches of at most kMaxGrid workgroups, each starting at chunk
// c0 (a multiple of 8, so the XCD-contiguous order stays a bijection).
constexpr uint64_t kMaxGrid = 1ull << 23;

template <typename F>
static inline void launch_chunks(uint64_t blocks, F&& launch)
{
  for (uint64_t c0 = 0; c0 < blocks; c0 += kMaxGrid)
    launch((unsigned)std::min<uint64_t>(kMaxGrid, blocks - c0), c0);
}

// Work per thread, from a 30-qubit in-place probe (GB/s of algorithmic
// traffic): one pair per thread in XCD-contiguous workgroup order streams
// complex64 pairs at 6437-6824 against 5780-6783 with two or four pairs per
// thread (10-12 % more for partners 4 MiB or more apart) and complex128 pairs
// within 1 % of the best shape at every distance. Kernels that touch one
// amplitude per element take one element per thread for complex128 (RZ on a
// 30-qubit state 5.23 ms against 5.53 ms with four) and four for complex64
// (2.66 ms against 2.90 ms with two).
constexpr int kPairsPerThread = 1;
template <typename R> constexpr int kEltsPerThread = sizeof(R) == 8 ? 1 : 4;

void launch_gate1(int prec, void* psi, const GateArgs& g, const double m[8], hipStream_t st)
{
  if (g.ins.count == 0) return;
  Mat2<double> M;
  for (int k = 0; k < 8; ++k) M.m[k] = m[k];
  if (g.lo_t == 0) {
    constexpr int PPT = kPairsPerThread;
    launch_chunks(blocks_for(g.ins.count, PPT), [&](unsigned grid, uint64_t c0) {
      if (prec == QSV_C128)
        hipLaunchKernelGGL((k_gate1<double, PPT, false, true>), dim3(grid), dim3(kBlock), 0, st,
                           (NV<double>::T*)psi, g, M, c0);
      else
        hipLaunchKernelGGL((k_gate1<float, PPT, false, true>), dim3(grid), dim3(kBlock), 0, st,
                           (NV<float>::T*)psi, g, M, c0);
    });
  } else if (prec == QSV_C128) {
    constexpr int EPT = kEltsPerThread<double>;
    launch_chunks(blocks_for(g.ins.count, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_gate1<double, EPT, true, true>), dim3(grid), dim3(kBlock), 0, st, (NV<double>::T*)psi,
                         g, M, c0);
    });
  } else {
    constexpr int EPT = kEltsPerThread<float>;
    launch_chunks(blocks_for(g.ins.count, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_gate1<float, EPT, true, true>), dim3(grid), dim3(kBlock), 0, st, (NV<float>::T*)psi,
                         g, M, c0);
    });
  }
}

void launch_phase(int prec, void* psi, const GateArgs& g, const double d0[2], const double d1[2], hipStream_t st)
{
  if (g.ins.count == 0) return;
  const Cplx<double> e0{d0[0], d0[1]}, e1{d1[0], d1[1]};
  if (prec == QSV_C128) {
    constexpr int EPT = kEltsPerThread<double>;
    launch_chunks(blocks_for(g.ins.count, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_phase<double, EPT>), dim3(grid), dim3(kBlock), 0, st, (NV<double>::T*)psi, g, e0, e1, c0);
    });
  } else {
    constexpr int EPT = kEltsPerThread<float>;
    launch_chunks(blocks_for(g.ins.count, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_phase<float, EPT>), dim3(grid), dim3(kBlock), 0, st, (NV<float>::T*)psi, g, e0, e1, c0);
    });
  }
}

void launch_swap2(int prec, void* psi, const SwapArgs& s, const double ph[2], hipStream_t st)
{
  if (s.ins.count == 0) return;
  const Cplx<double> p{ph[0], ph[1]};
  const int mode = (s.hA && s.hB) ? 0 : (s.hA ? 1 : 2);
#define QSV_SWAP_LAUNCH(R, MODE, PPT)                                                                            \
  launch_chunks(blocks_for(s.ins.count, PPT), [&](unsigned grid, uint64_t c0) {                                 \
    hipLaunchKernelGGL((k_swap2<R, PPT, MODE, true>), dim3(grid), dim3(kBlock), 0, st, (NV<R>::T*)psi, s, p, c0); \
  })
  // modes 0 and 1 move pairs; mode 2 (both bits lane bits) one amplitude per element
  if (prec == QSV_C128) {
    if (mode == 0) QSV_SWAP_LAUNCH(double, 0, kPairsPerThread);
    else if (mode == 1) QSV_SWAP_LAUNCH(double, 1, kPairsPerThread);
    else QSV_SWAP_LAUNCH(double, 2, kEltsPerThread<double>);
  } else {
    if (mode == 0) QSV_SWAP_LAUNCH(float, 0, kPairsPerThread);
    else if (mode == 1) QSV_SWAP_LAUNCH(float, 1, kPairsPerThread);
    else QSV_SWAP_LAUNCH(float, 2, kEltsPerThread<float>);
  }
#undef QSV_SWAP_LAUNCH
}

void launch_blockade(int prec, void* psi, uint64_t S, uint64_t bmask, const double ph[2], hipStream_t st)
{
  const Cplx<double> p{ph[0], ph[1]};
  if (prec == QSV_C128) {
    constexpr int EPT = kEltsPerThread<double>;
    launch_chunks(blocks_for(S, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_blockade<double, EPT>), dim3(grid), dim3(kBlock), 0, st, (NV<double>::T*)psi, S, bmask,
                         p, c0);
    });
  } else {
    constexpr int EPT = kEltsPerThread<float>;
    launch_chunks(blocks_for(S, EPT), [&](unsigned grid, uint64_t c0) {
      hipLaunchKernelGGL((k_blockade<float, EPT>), dim3(grid), dim3(kBlock), 0, st, (NV<float>::T*)psi, S, bmask, p,
                         c0);
    });
  }
}

}  // namespace qsv
