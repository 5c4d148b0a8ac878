// qsv_kernels.hip — state initialisation, conversions, FP64 reductions
// (norm, overlaps, Pauli expectations), sampling and measurement of libqsv.
//
// State layout: 2^n amplitudes, interleaved (re, im); c128 = double2 (16 B),
// c64 = float2 (8 B). Reference qubit q is index bit b = n-1-q.
#include "qsv_internal.h"

namespace qsv {

template <typename R> struct VT;
template <> struct VT<double> { using T = double2; };
template <> struct VT<float> { using T = float2; };

template <typename R>
__global__ __launch_bounds__(kBlock) void k_fill_basis(typename VT<R>::T* __restrict__ psi, uint64_t S, uint64_t idx)
{
  using T = typename VT<R>::T;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    T v;
    v.x = (i == idx) ? R(1) : R(0);
    v.y = R(0);
    psi[i] = v;
  }
}

// Plain streaming copy (non-temporal), 4 amplitudes per thread.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_copy(const typename VT<R>::T* __restrict__ src,
                                                 typename VT<R>::T* __restrict__ dst, uint64_t S)
{
  typedef R V __attribute__((ext_vector_type(2)));
  const V* s = reinterpret_cast<const V*>(src);
  V* d = reinterpret_cast<V*>(dst);
  const uint64_t base = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
  V v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + (uint64_t)k * kBlock;
    if (i < S) v[k] = __builtin_nontemporal_load(s + i);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + (uint64_t)k * kBlock;
    if (i < S) __builtin_nontemporal_store(v[k], d + i);
  }
}

// One pass over two states of any precisions: partials of Re/Im <a|b>,
// sum |a|^2, sum |b|^2 (FP64) and max |a_i - b_i|.
template <typename RA, typename RB>
__global__ __launch_bounds__(kBlock) void k_compare(const typename VT<RA>::T* __restrict__ a,
                                                    const typename VT<RB>::T* __restrict__ b, uint64_t S,
                                                    double* __restrict__ partials)
{
  __shared__ double sh[kBlock / 64];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double re = 0.0, im = 0.0, na = 0.0, nb = 0.0, mx = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
    re += ax * bx + ay * by;
    im += ax * by - ay * bx;
    na += ax * ax + ay * ay;
    nb += bx * bx + by * by;
    const double dx = ax - bx, dy = ay - by;
    mx = fmax(mx, sqrt(dx * dx + dy * dy));
  }
  double vals[4] = {re, im, na, nb};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v = vals[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[(uint64_t)q * kReduceBlocks + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0)
    partials[4ull * kReduceBlocks + blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ __launch_bounds__(kBlock) void k_final_max(const double* __restrict__ partials, double* __restrict__ out)
{
  __shared__ double sh[kBlock / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < kReduceBlocks; i += kBlock) m = fmax(m, partials[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_scale(typename VT<R>::T* __restrict__ psi, uint64_t S, R f)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    auto v = psi[i];
    v.x *= f;
    v.y *= f;
    psi[i] = v;
  }
}

// normalize-state with the norm^2 already on the device (norm2[0]): divide by
// the norm when it exceeds 1e-12, else leave the amplitudes.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_scale_by_norm(typename VT<R>::T* __restrict__ psi, uint64_t S,
                                                          const double* __restrict__ norm2)
{
  const double nrm = sqrt(norm2[0]);
  if (!(nrm > 1e-12)) return;
  const R f = (R)(1.0 / nrm);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    auto v = psi[i];
    v.x *= f;
    v.y *= f;
    psi[i] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_f64_to_c64(const double2* __restrict__ src, float2* __restrict__ dst,
                                                        uint64_t count)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ __launch_bounds__(kBlock) void k_c64_to_f64(const float2* __restrict__ src, double2* __restrict__ dst,
                                                        uint64_t count)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
    const float2 v = src[i];
    dst[i] = make_double2((double)v.x, (double)v.y);
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: fixed grid of kReduceBlocks blocks writes per-block
// partials, then one block per value sums them in a fixed order (the result
// is deterministic run to run).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

// partials[v * kReduceBlocks + b] -> out[v]
__global__ __launch_bounds__(kBlock) void k_final_sum(const double* __restrict__ partials, double* __restrict__ out)
{
  __shared__ double sh[kBlock / 64];
  const double* p = partials + (uint64_t)blockIdx.x * kReduceBlocks;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kReduceBlocks; i += kBlock) acc += p[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_norm2(const typename VT<R>::T* __restrict__ psi, uint64_t S,
                                                   double* __restrict__ partials)
{
  __shared__ double sh[kBlock / 64];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double acc0 = 0.0, acc1 = 0.0;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + stride < S; i += 2 * stride) {
    const auto a = psi[i];
    const auto b = psi[i + stride];
    acc0 += (double)a.x * a.x + (double)a.y * a.y;
    acc1 += (double)b.x * b.x + (double)b.y * b.y;
  }
  if (i < S) {
    const auto a = psi[i];
    acc0 += (double)a.x * a.x + (double)a.y * a.y;
  }
  const double r = block_sum(acc0 + acc1, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_inner(const typename VT<R>::T* __restrict__ a,
                                                   const typename VT<R>::T* __restrict__ b, uint64_t S,
                                                   double* __restrict__ partials)
{
  __shared__ double sh[kBlock / 64];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double re = 0.0, im = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    const auto x = a[i];
    const auto y = b[i];
    // conj(x) * y
    re += (double)x.x * y.x + (double)x.y * y.y;
    im += (double)x.x * y.y - (double)x.y * y.x;
  }
  const double r0 = block_sum(re, sh);
  const double r1 = block_sum(im, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = r0;
    partials[kReduceBlocks + blockIdx.x] = r1;
  }
}

struct PauliArgs {
  int32_t nterms;
  uint64_t xmask[kMaxPauliPerPass];
  uint64_t smask[kMaxPauliPerPass];
};

// For each term t: sum_i (-1)^popc(i & s_t) conj(a_i) a_{i ^ x_t}.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_pauli(const typename VT<R>::T* __restrict__ psi, uint64_t S,
                                                   PauliArgs pa, double* __restrict__ partials)
{
  __shared__ double sh[kBlock / 64];
  double re[kMaxPauliPerPass], im[kMaxPauliPerPass];
#pragma unroll
  for (int t = 0; t < kMaxPauliPerPass; ++t) { re[t] = 0.0; im[t] = 0.0; }
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    const auto a = psi[i];
    const double ar = a.x, ai = a.y;
#pragma unroll
    for (int t = 0; t < kMaxPauliPerPass; ++t) {
      if (t < pa.nterms) {
        const double sgn = (__popcll(i & pa.smask[t]) & 1) ? -1.0 : 1.0;
        if (pa.xmask[t] == 0) {
          re[t] += sgn * (ar * ar + ai * ai);
        } else {
          const auto b = psi[i ^ pa.xmask[t]];
          re[t] += sgn * (ar * (double)b.x + ai * (double)b.y);
          im[t] += sgn * (ar * (double)b.y - ai * (double)b.x);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kMaxPauliPerPass; ++t) {
    if (t < pa.nterms) {
      const double r0 = block_sum(re[t], sh);
      const double r1 = block_sum(im[t], sh);
      if (threadIdx.x == 0) {
        partials[(uint64_t)(2 * t) * kReduceBlocks + blockIdx.x] = r0;
        partials[(uint64_t)(2 * t + 1) * kReduceBlocks + blockIdx.x] = r1;
      }
    }
  }
}

// <psi| O on bit tbit |psi>
template <typename R>
__global__ __launch_bounds__(kBlock) void k_expect1q(const typename VT<R>::T* __restrict__ psi, uint64_t half,
                                                      uint64_t tbit, Mat2<double> M, double* __restrict__ partials)
{
  __shared__ double sh[kBlock / 64];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t lo = tbit - 1;
  double re = 0.0, im = 0.0;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < half; p += stride) {
    const uint64_t i0 = ((p & ~lo) << 1) | (p & lo);
    const auto A = psi[i0];
    const auto B = psi[i0 | tbit];
    const double a0r = A.x, a0i = A.y, a1r = B.x, a1i = B.y;
    const double y0r = M.m[0] * a0r - M.m[1] * a0i + M.m[2] * a1r - M.m[3] * a1i;
    const double y0i = M.m[0] * a0i + M.m[1] * a0r + M.m[2] * a1i + M.m[3] * a1r;
    const double y1r = M.m[4] * a0r - M.m[5] * a0i + M.m[6] * a1r - M.m[7] * a1i;
    const double y1i = M.m[4] * a0i + M.m[5] * a0r + M.m[6] * a1i + M.m[7] * a1r;
    re += a0r * y0r + a0i * y0i + a1r * y1r + a1i * y1i;
    im += a0r * y0i - a0i * y0r + a1r * y1i - a1i * y1r;
  }
  const double r0 = block_sum(re, sh);
  const double r1 = block_sum(im, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = r0;
    partials[kReduceBlocks + blockIdx.x] = r1;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_probs(const typename VT<R>::T* __restrict__ psi, uint64_t offset,
                                                   uint64_t count, double* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < count; k += stride) {
    const auto a = psi[offset + k];
    out[k] = (double)a.x * a.x + (double)a.y * a.y;
  }
}

// ---------------------------------------------------------------------------
// Shot sampling (measure-state, inverse CDF): chunk sums, inclusive scan of
// chunk sums, then one wave per shot: binary search over chunks and a wave
// scan inside the chunk.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const typename VT<R>::T* __restrict__ psi, int chunk_log,
                                                        double* __restrict__ sums)
{
  __shared__ double sh[kBlock / 64];
  const uint64_t c0 = (uint64_t)blockIdx.x << chunk_log, len = 1ull << chunk_log;
  double acc = 0.0;
  for (uint64_t k = threadIdx.x; k < len; k += kBlock) {
    const auto a = psi[c0 + k];
    acc += (double)a.x * a.x + (double)a.y * a.y;
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// Inclusive scan of n doubles with one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_inclusive(double* __restrict__ v, uint64_t n)
{
  __shared__ double tot[1024];
  const uint64_t seg = (n + 1023) / 1024;
  const uint64_t b = threadIdx.x * seg, e = (b + seg < n) ? b + seg : n;
  double run = 0.0;
  for (uint64_t k = b; k < e; ++k) { run += v[k]; v[k] = run; }
  tot[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const double t = (threadIdx.x >= (unsigned)off) ? tot[threadIdx.x - off] : 0.0;
    __syncthreads();
    tot[threadIdx.x] += t;
    __syncthreads();
  }
  const double offset = threadIdx.x ? tot[threadIdx.x - 1] : 0.0;
  for (uint64_t k = b; k < e; ++k) v[k] += offset;
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t k)
{
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_sample(const typename VT<R>::T* __restrict__ psi, uint64_t S,
                                                    int chunk_log, const double* __restrict__ prefix,
                                                    uint64_t nchunks, uint64_t shots, uint64_t seed,
                                                    uint64_t* __restrict__ out)
{
  const int lane = threadIdx.x & 63;
  const uint64_t shot = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (shot >= shots) return;
  const double total = prefix[nchunks - 1];
  const double r = uniform01(seed, shot) * total;
  // first chunk whose inclusive prefix is >= r
  uint64_t lo = 0, hi = nchunks;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (prefix[mid] < r) lo = mid + 1; else hi = mid;
  }
  if (lo >= nchunks) {
    if (lane == 0) out[shot] = S - 1;
    return;
  }
  const uint64_t c = lo;
  double run = c ? prefix[c - 1] : 0.0;
  const uint64_t len = 1ull << chunk_log, c0 = c << chunk_log;
  uint64_t result = c0 + len - 1;
  for (uint64_t piece = 0; piece < len; piece += 64) {
    double v = 0.0;
    if (piece + lane < len) {
      const auto a = psi[c0 + piece + lane];
      v = (double)a.x * a.x + (double)a.y * a.y;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double t = __shfl_up(v, off, 64);
      if (lane >= off) v += t;
    }
    const double cum = run + v;
    const uint64_t hit = __ballot(cum >= r && piece + lane < len);
    if (hit) {
      result = c0 + piece + (uint64_t)(__ffsll((unsigned long long)hit) - 1);
      break;
    }
    run += __shfl(v, 63, 64);
  }
  if (lane == 0) out[shot] = result;
}

// ---------------------------------------------------------------------------
// Partial measurement.
// ---------------------------------------------------------------------------
struct PosArgs {
  int32_t m;
  int32_t pos[24];
};

__device__ __forceinline__ uint64_t gather_bits(uint64_t i, const PosArgs& p)
{
  uint64_t o = 0;
  for (int j = 0; j < p.m; ++j) o |= ((i >> p.pos[j]) & 1ull) << j;
  return o;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_outcome_hist_lds(const typename VT<R>::T* __restrict__ psi, uint64_t S,
                                                              PosArgs p, double* __restrict__ hist)
{
  __shared__ double h[4096];
  const int no = 1 << p.m;
  for (int k = threadIdx.x; k < no; k += kBlock) h[k] = 0.0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    const auto a = psi[i];
    atomicAdd(&h[gather_bits(i, p)], (double)a.x * a.x + (double)a.y * a.y);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < no; k += kBlock)
    if (h[k] != 0.0) atomicAdd(&hist[k], h[k]);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_outcome_hist_global(const typename VT<R>::T* __restrict__ psi,
                                                                 uint64_t S, PosArgs p, double* __restrict__ hist)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    const auto a = psi[i];
    const double v = (double)a.x * a.x + (double)a.y * a.y;
    if (v != 0.0) atomicAdd(&hist[gather_bits(i, p)], v);
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_f64(double* __restrict__ v, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) v[i] = 0.0;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse(typename VT<R>::T* __restrict__ psi, uint64_t S, PosArgs p,
                                                      uint64_t sel, R scale)
{
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S; i += stride) {
    auto a = psi[i];
    if (gather_bits(i, p) == sel) {
      a.x *= scale;
      a.y *= scale;
    } else {
      a.x = R(0);
      a.y = R(0);
    }
    psi[i] = a;
  }
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
static inline unsigned stream_grid(uint64_t S)
{
  uint64_t g = (S + kBlock - 1) / kBlock;
  const uint64_t cap = 256ull * 64;  // 64 workgroups per CU on 256 CUs
  return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

void launch_fill_basis(int prec, void* psi, uint64_t S, uint64_t idx, hipStream_t st)
{
  const unsigned g = stream_grid(S);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_fill_basis<double>), dim3(g), dim3(kBlock), 0, st, (double2*)psi, S, idx);
  else
    hipLaunchKernelGGL((k_fill_basis<float>), dim3(g), dim3(kBlock), 0, st, (float2*)psi, S, idx);
}

void launch_copy(int prec, const void* src, void* dst, uint64_t S, hipStream_t st)
{
  const unsigned g = (unsigned)((S + kBlock * 4 - 1) / (kBlock * 4));
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_copy<double>), dim3(g), dim3(kBlock), 0, st, (const double2*)src, (double2*)dst, S);
  else
    hipLaunchKernelGGL((k_copy<float>), dim3(g), dim3(kBlock), 0, st, (const float2*)src, (float2*)dst, S);
}

void launch_compare(int pa, const void* a, int pb, const void* b, uint64_t S, double* partials, double* out,
                    hipStream_t st)
{
  const dim3 g(kReduceBlocks), blk(kBlock);
  if (pa == QSV_C128 && pb == QSV_C128)
    hipLaunchKernelGGL((k_compare<double, double>), g, blk, 0, st, (const double2*)a, (const double2*)b, S, partials);
  else if (pa == QSV_C128)
    hipLaunchKernelGGL((k_compare<double, float>), g, blk, 0, st, (const double2*)a, (const float2*)b, S, partials);
  else if (pb == QSV_C128)
    hipLaunchKernelGGL((k_compare<float, double>), g, blk, 0, st, (const float2*)a, (const double2*)b, S, partials);
  else
    hipLaunchKernelGGL((k_compare<float, float>), g, blk, 0, st, (const float2*)a, (const float2*)b, S, partials);
  hipLaunchKernelGGL(k_final_sum, dim3(4), blk, 0, st, partials, out);
  hipLaunchKernelGGL(k_final_max, dim3(1), blk, 0, st, partials + 4ull * kReduceBlocks, out + 4);
}

void launch_scale(int prec, void* psi, uint64_t S, double f, hipStream_t st)
{
  const unsigned g = stream_grid(S);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_scale<double>), dim3(g), dim3(kBlock), 0, st, (double2*)psi, S, f);
  else
    hipLaunchKernelGGL((k_scale<float>), dim3(g), dim3(kBlock), 0, st, (float2*)psi, S, (float)f);
}

void launch_renorm(int prec, void* psi, uint64_t S, double* partials, double* out, hipStream_t st)
{
  launch_norm2(prec, psi, S, partials, out, st);
  const unsigned g = stream_grid(S);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_scale_by_norm<double>), dim3(g), dim3(kBlock), 0, st, (double2*)psi, S, out);
  else
    hipLaunchKernelGGL((k_scale_by_norm<float>), dim3(g), dim3(kBlock), 0, st, (float2*)psi, S, out);
}

void launch_f64_to_c64(const double* src, void* dst, uint64_t count, hipStream_t st)
{
  hipLaunchKernelGGL(k_f64_to_c64, dim3(stream_grid(count)), dim3(kBlock), 0, st, (const double2*)src, (float2*)dst,
                     count);
}

void launch_c64_to_f64(const void* src, double* dst, uint64_t count, hipStream_t st)
{
  hipLaunchKernelGGL(k_c64_to_f64, dim3(stream_grid(count)), dim3(kBlock), 0, st, (const float2*)src, (double2*)dst,
                     count);
}

void launch_norm2(int prec, const void* psi, uint64_t S, double* partials, double* out, hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_norm2<double>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const double2*)psi, S, partials);
  else
    hipLaunchKernelGGL((k_norm2<float>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const float2*)psi, S, partials);
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, st, partials, out);
}

void launch_inner(int prec, const void* a, const void* b, uint64_t S, double* partials, double* out, hipStream_t st)
{
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_inner<double>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const double2*)a,
                       (const double2*)b, S, partials);
  else
    hipLaunchKernelGGL((k_inner<float>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const float2*)a,
                       (const float2*)b, S, partials);
  hipLaunchKernelGGL(k_final_sum, dim3(2), dim3(kBlock), 0, st, partials, out);
}

void launch_pauli(int prec, const void* psi, uint64_t S, int nterms, const uint64_t* xmask, const uint64_t* smask,
                  double* partials, double* out, hipStream_t st)
{
  PauliArgs pa{};
  pa.nterms = nterms;
  for (int t = 0; t < nterms; ++t) {
    pa.xmask[t] = xmask[t];
    pa.smask[t] = smask[t];
  }
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_pauli<double>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const double2*)psi, S, pa,
                       partials);
  else
    hipLaunchKernelGGL((k_pauli<float>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const float2*)psi, S, pa,
                       partials);
  hipLaunchKernelGGL(k_final_sum, dim3(2 * nterms), dim3(kBlock), 0, st, partials, out);
}

void launch_expect1q(int prec, const void* psi, uint64_t S, uint64_t tbit, const double m[8], double* partials,
                     double* out, hipStream_t st)
{
  Mat2<double> M;
  for (int k = 0; k < 8; ++k) M.m[k] = m[k];
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_expect1q<double>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const double2*)psi, S / 2,
                       tbit, M, partials);
  else
    hipLaunchKernelGGL((k_expect1q<float>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const float2*)psi, S / 2,
                       tbit, M, partials);
  hipLaunchKernelGGL(k_final_sum, dim3(2), dim3(kBlock), 0, st, partials, out);
}

void launch_probs(int prec, const void* psi, uint64_t offset, uint64_t count, double* out, hipStream_t st)
{
  const unsigned g = stream_grid(count);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_probs<double>), dim3(g), dim3(kBlock), 0, st, (const double2*)psi, offset, count, out);
  else
    hipLaunchKernelGGL((k_probs<float>), dim3(g), dim3(kBlock), 0, st, (const float2*)psi, offset, count, out);
}

void launch_sample(int prec, const void* psi, uint64_t S, double* chunk_prefix, uint64_t shots, uint64_t seed,
                   uint64_t* out, hipStream_t st)
{
  int n = 0;
  while ((1ull << n) < S) ++n;
  const int chunk_log = n < kSampleChunkLog ? n : kSampleChunkLog;
  const uint64_t nchunks = S >> chunk_log;
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_chunk_sums<double>), dim3((unsigned)nchunks), dim3(kBlock), 0, st, (const double2*)psi,
                       chunk_log, chunk_prefix);
  else
    hipLaunchKernelGGL((k_chunk_sums<float>), dim3((unsigned)nchunks), dim3(kBlock), 0, st, (const float2*)psi,
                       chunk_log, chunk_prefix);
  hipLaunchKernelGGL(k_scan_inclusive, dim3(1), dim3(1024), 0, st, chunk_prefix, nchunks);
  if (shots == 0) return;
  const unsigned g = (unsigned)((shots + (kBlock / 64) - 1) / (kBlock / 64));
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_sample<double>), dim3(g), dim3(kBlock), 0, st, (const double2*)psi, S, chunk_log,
                       chunk_prefix, nchunks, shots, seed, out);
  else
    hipLaunchKernelGGL((k_sample<float>), dim3(g), dim3(kBlock), 0, st, (const float2*)psi, S, chunk_log,
                       chunk_prefix, nchunks, shots, seed, out);
}

void launch_outcome_hist(int prec, const void* psi, uint64_t S, int m, const int* pos, double* hist, hipStream_t st)
{
  PosArgs p{};
  p.m = m;
  for (int j = 0; j < m; ++j) p.pos[j] = pos[j];
  const uint64_t no = 1ull << m;
  hipLaunchKernelGGL(k_zero_f64, dim3(stream_grid(no)), dim3(kBlock), 0, st, hist, no);
  if (m <= 12) {
    if (prec == QSV_C128)
      hipLaunchKernelGGL((k_outcome_hist_lds<double>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const double2*)psi,
                         S, p, hist);
    else
      hipLaunchKernelGGL((k_outcome_hist_lds<float>), dim3(kReduceBlocks), dim3(kBlock), 0, st, (const float2*)psi,
                         S, p, hist);
  } else {
    if (prec == QSV_C128)
      hipLaunchKernelGGL((k_outcome_hist_global<double>), dim3(stream_grid(S)), dim3(kBlock), 0, st,
                         (const double2*)psi, S, p, hist);
    else
      hipLaunchKernelGGL((k_outcome_hist_global<float>), dim3(stream_grid(S)), dim3(kBlock), 0, st,
                         (const float2*)psi, S, p, hist);
  }
}

void launch_collapse(int prec, void* psi, uint64_t S, int m, const int* pos, uint64_t sel, double scale,
                     hipStream_t st)
{
  PosArgs p{};
  p.m = m;
  for (int j = 0; j < m; ++j) p.pos[j] = pos[j];
  const unsigned g = stream_grid(S);
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_collapse<double>), dim3(g), dim3(kBlock), 0, st, (double2*)psi, S, p, sel, scale);
  else
    hipLaunchKernelGGL((k_collapse<float>), dim3(g), dim3(kBlock), 0, st, (float2*)psi, S, p, sel, (float)scale);
}

}  // namespace qsv
