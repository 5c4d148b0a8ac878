// qsv_batch.hip — many small circuits in one launch.
//
// One workgroup per circuit: the whole 2^n-amplitude state lives in LDS
// (n <= 13 for complex128, n <= 14 for complex64 with 128 KiB of LDS), the
// circuit's decoded gate records are interpreted one by one (one LDS sweep and
// one barrier per gate), and <H> of a Z-string Hamiltonian is reduced in FP64
// before the workgroup exits. Nothing but the gate records and the energies
// crosses HBM. Workgroups are dispatched largest circuit first.
#include "qsv_internal.h"

namespace qsv {

template <typename R> struct BV;
template <> struct BV<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct BV<float> { typedef float T __attribute__((ext_vector_type(2))); };

__device__ __forceinline__ uint64_t ins0(uint64_t j, int p)
{
  const uint64_t lo = j & ((1ull << p) - 1ull);
  return ((j ^ lo) << 1) | lo;
}

template <typename T, typename R>
__device__ __forceinline__ T bmul(double re, double im, T a)
{
  const double ax = a.x, ay = a.y;
  T y;
  y.x = (R)(re * ax - im * ay);
  y.y = (R)(re * ay + im * ax);
  return y;
}

template <typename R>
__global__ __launch_bounds__(kBatchThreads) void k_lds_batch(const BInst* __restrict__ inst,
                                                              const BRec* __restrict__ gates,
                                                              const uint64_t* __restrict__ tz,
                                                              const double* __restrict__ tc,
                                                              const int32_t* __restrict__ order,
                                                              double* __restrict__ energies,
                                                              typename BV<R>::T* __restrict__ out_states)
{
  using T = typename BV<R>::T;
  extern __shared__ unsigned char smem_raw[];
  T* st = reinterpret_cast<T*>(smem_raw);
  __shared__ double red[kBatchThreads / 64];

  const int i = order[blockIdx.x];
  const BInst in = inst[i];
  const uint64_t S = 1ull << in.n;
  const uint32_t tid = threadIdx.x, nt = blockDim.x;

  for (uint64_t j = tid; j < S; j += nt) st[j] = T{(R)(j == 0 ? 1 : 0), (R)0};
  __syncthreads();

  for (uint64_t gi = in.g0; gi < in.g1; ++gi) {
    const BRec g = gates[gi];
    if (g.kind == FG_DENSE) {
      const uint64_t tb = 1ull << g.a;
      for (uint64_t p = tid; p < S / 2; p += nt) {
        const uint64_t i0 = ins0(p, g.a);
        if ((i0 & g.cmask) != g.cmask) continue;
        const T a = st[i0], b = st[i0 | tb];
        const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
        T y0, y1;
        y0.x = (R)(g.m[0] * ax - g.m[1] * ay + g.m[2] * bx - g.m[3] * by);
        y0.y = (R)(g.m[0] * ay + g.m[1] * ax + g.m[2] * by + g.m[3] * bx);
        y1.x = (R)(g.m[4] * ax - g.m[5] * ay + g.m[6] * bx - g.m[7] * by);
        y1.y = (R)(g.m[4] * ay + g.m[5] * ax + g.m[6] * by + g.m[7] * bx);
        st[i0] = y0;
        st[i0 | tb] = y1;
      }
    } else if (g.kind == FG_PHASE) {
      const uint64_t tb = 1ull << g.a;
      for (uint64_t j = tid; j < S; j += nt) {
        if ((j & g.cmask) != g.cmask) continue;
        if (j & tb) st[j] = bmul<T, R>(g.m[6], g.m[7], st[j]);
        else if (!g.phase_only) st[j] = bmul<T, R>(g.m[0], g.m[1], st[j]);
      }
    } else if (g.kind == FG_SWAP) {
      const int lo = g.a < g.b ? g.a : g.b, hi = g.a < g.b ? g.b : g.a;
      const uint64_t ba = 1ull << g.a, bb = 1ull << g.b;
      for (uint64_t p = tid; p < S / 4; p += nt) {
        const uint64_t base = ins0(ins0(p, lo), hi);
        if ((base & g.cmask) != g.cmask) continue;
        const T a = st[base | ba], b = st[base | bb];
        st[base | ba] = bmul<T, R>(g.m[0], g.m[1], b);
        st[base | bb] = bmul<T, R>(g.m[0], g.m[1], a);
      }
    } else if (g.kind == FG_BLOCKADE) {
      for (uint64_t j = tid; j < S; j += nt)
        if (__popcll(j & g.bmask) == 1) st[j] = bmul<T, R>(g.m[0], g.m[1], st[j]);
    } else {  // FG_RENORM: normalize-state after a Kraus operator, FP64 norm
      double acc = 0.0;
      for (uint64_t j = tid; j < S; j += nt) {
        const T a = st[j];
        acc += (double)a.x * a.x + (double)a.y * a.y;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
      if ((tid & 63) == 0) red[tid >> 6] = acc;
      __syncthreads();
      double tot = 0.0;
      for (uint32_t w = 0; w < nt / 64; ++w) tot += red[w];   // same order in every thread
      const double nrm = sqrt(tot);
      if (nrm > 1e-12) {
        const double inv = 1.0 / nrm;
        for (uint64_t j = tid; j < S; j += nt) st[j] = bmul<T, R>(inv, 0.0, st[j]);
      }
    }
    __syncthreads();
  }

  if (out_states && in.out_offset != ~0ull)
    for (uint64_t j = tid; j < S; j += nt) out_states[in.out_offset + j] = st[j];
  if (!energies) return;
  // <H> = sum_x |psi_x|^2 h(x), h(x) = sum_t c_t (-1)^popc(x & z_t), FP64
  double acc = 0.0;
  for (uint64_t j = tid; j < S; j += nt) {
    const T a = st[j];
    const double p = (double)a.x * a.x + (double)a.y * a.y;
    double h = 0.0;
    for (uint64_t t = in.t0; t < in.t1; ++t) h += (__popcll(j & tz[t]) & 1) ? -tc[t] : tc[t];
    acc += p * h;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    double e = 0.0;
    for (uint32_t w = 0; w < nt / 64; ++w) e += red[w];
    energies[i] = e;
  }
}

// Shots for many small states in an arena: one workgroup per state. Chunk
// sums (FP64) and their inclusive scan live in LDS; each shot is a binary
// search over chunks and a scan inside its chunk (inverse CDF, as measure-state).

__device__ __forceinline__ double uni01(uint64_t seed, uint64_t k)
{
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ __launch_bounds__(256) void k_batch_sample(const typename BV<R>::T* __restrict__ arena,
                                                      const SampleJob* __restrict__ jobs,
                                                      uint64_t* __restrict__ outcomes,
                                                      double* __restrict__ norms)
{
  __shared__ double cum[256];
  const SampleJob jb = jobs[blockIdx.x];
  const auto* psi = arena + jb.state_offset;
  const uint64_t S = 1ull << jb.n;
  const uint32_t nch = S < 256 ? (uint32_t)S : 256u;
  const uint64_t C = S / nch;
  const uint32_t tid = threadIdx.x;
  double v = 0.0;
  if (tid < nch)
    for (uint64_t k = 0; k < C; ++k) {
      const auto a = psi[(uint64_t)tid * C + k];
      v += (double)a.x * a.x + (double)a.y * a.y;
    }
  cum[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const double t = (tid >= (uint32_t)off) ? cum[tid - off] : 0.0;
    __syncthreads();
    cum[tid] += t;
    __syncthreads();
  }
  const double total = cum[nch - 1];
  if (norms && tid == 0) norms[blockIdx.x] = total;
  for (uint64_t s = tid; s < jb.shots; s += 256) {
    const double r = uni01(jb.seed, s) * total;
    uint32_t lo = 0, hi = nch;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cum[mid] < r) lo = mid + 1; else hi = mid;
    }
    uint64_t out = S - 1;
    if (lo < nch) {
      double run = lo ? cum[lo - 1] : 0.0;
      out = (uint64_t)lo * C + C - 1;
      for (uint64_t k = 0; k < C; ++k) {
        const auto a = psi[(uint64_t)lo * C + k];
        run += (double)a.x * a.x + (double)a.y * a.y;
        if (run >= r) { out = (uint64_t)lo * C + k; break; }
      }
    }
    outcomes[jb.out_offset + s] = out;
  }
}

void launch_batch_sample(int prec, const void* arena, const void* jobs, unsigned njobs, uint64_t* outcomes,
                         double* norms, hipStream_t st)
{
  if (njobs == 0) return;
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_batch_sample<double>), dim3(njobs), dim3(256), 0, st, (const BV<double>::T*)arena,
                       (const SampleJob*)jobs, outcomes, norms);
  else
    hipLaunchKernelGGL((k_batch_sample<float>), dim3(njobs), dim3(256), 0, st, (const BV<float>::T*)arena,
                       (const SampleJob*)jobs, outcomes, norms);
}

constexpr uint32_t kZTerms = 256;   // Z strings staged in LDS per round
constexpr int kZAmps = 4;           // amplitudes per thread per sweep over the staged strings

// This thread's share of sum_j |a_j|^2 * sum_t c_t (-1)^popcount(j & z_t) over
// the amplitudes j = first, first + stride, ... < S. The strings are read from
// LDS (staged 256 at a time by the whole workgroup, so every thread of the
// block must call this with the same t0, t1), each string against four
// amplitudes per read.
template <typename R>
__device__ __forceinline__ double zstring_share(const typename BV<R>::T* __restrict__ psi, uint64_t S, uint64_t first,
                                                uint64_t stride, const uint64_t* __restrict__ tz,
                                                const double* __restrict__ tc, uint64_t t0, uint64_t t1)
{
  __shared__ uint64_t sz[kZTerms];
  __shared__ double sc[kZTerms];
  double acc = 0.0;
  for (uint64_t c0 = t0; c0 < t1; c0 += kZTerms) {
    const uint32_t nc = (uint32_t)(t1 - c0 < kZTerms ? t1 - c0 : kZTerms);
    __syncthreads();   // the previous round's strings are no longer read
    for (uint32_t k = threadIdx.x; k < nc; k += blockDim.x) {
      sz[k] = tz[c0 + k];
      sc[k] = tc[c0 + k];
    }
    __syncthreads();
    for (uint64_t j = first; j < S; j += kZAmps * stride) {
      uint64_t jj[kZAmps];
      double p[kZAmps], h[kZAmps];
#pragma unroll
      for (int a = 0; a < kZAmps; ++a) {
        jj[a] = j + (uint64_t)a * stride;
        p[a] = 0.0;
        h[a] = 0.0;
        if (jj[a] < S) {
          const auto v = psi[jj[a]];
          p[a] = (double)v.x * v.x + (double)v.y * v.y;
        }
      }
      if (S <= (1ull << 32)) {
        // indices fit 32 bits: one AND and one popcount per string; the
        // parity flips the sign bit of c
#pragma unroll 4
        for (uint32_t k = 0; k < nc; ++k) {
          const uint32_t z = (uint32_t)sz[k];
          const double c = sc[k];
          const int chi = __double2hiint(c), clo = __double2loint(c);
#pragma unroll
          for (int a = 0; a < kZAmps; ++a)
            h[a] += __hiloint2double(chi ^ (int)(__popc((uint32_t)jj[a] & z) << 31), clo);
        }
      } else {
        for (uint32_t k = 0; k < nc; ++k) {
          const uint64_t z = sz[k];
          const double c = sc[k];
#pragma unroll
          for (int a = 0; a < kZAmps; ++a) h[a] += (__popcll(jj[a] & z) & 1) ? -c : c;
        }
      }
#pragma unroll
      for (int a = 0; a < kZAmps; ++a) acc += p[a] * h[a];
    }
  }
  return acc;
}

// <H> of a Z-string Hamiltonian on a state in global memory: per-block FP64
// partials, then a fixed-order final sum into out[0].
template <typename R>
__global__ __launch_bounds__(256) void k_zexpect(const typename BV<R>::T* __restrict__ psi, uint64_t S,
                                                 const uint64_t* __restrict__ tz, const double* __restrict__ tc,
                                                 uint64_t nterms, double* __restrict__ partials)
{
  __shared__ double red[4];
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  double acc = zstring_share<R>(psi, S, (uint64_t)blockIdx.x * 256 + threadIdx.x, stride, tz, tc, 0, nterms);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_sum_to(const double* __restrict__ partials, uint32_t n,
                                                double* __restrict__ out)
{
  __shared__ double red[4];
  double acc = 0.0;
  for (uint32_t k = threadIdx.x; k < n; k += 256) acc += partials[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// <H> = sum_k c_k <Z-string_k> of `count` states of S amplitudes back to
// back: block (x, y) sums its share of state y into partials[y * gridDim.x + x].
template <typename R>
__global__ __launch_bounds__(256) void k_zexpect_group(const typename BV<R>::T* __restrict__ base, uint64_t S,
                                                       const uint64_t* __restrict__ tz,
                                                       const double* __restrict__ tc,
                                                       const uint64_t* __restrict__ trange,
                                                       double* __restrict__ partials)
{
  __shared__ double red[4];
  const uint64_t y = blockIdx.y;
  const auto* psi = base + y * S;
  const uint64_t t0 = trange[2 * y], t1 = trange[2 * y + 1];
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  double acc = zstring_share<R>(psi, S, (uint64_t)blockIdx.x * 256 + threadIdx.x, stride, tz, tc, t0, t1);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// out[dest[y]] = fixed-order sum of the nb partials of state y.
__global__ __launch_bounds__(64) void k_sum_group(const double* __restrict__ partials, uint32_t nb,
                                                  const uint64_t* __restrict__ dest, double* __restrict__ out)
{
  const uint64_t y = blockIdx.x;
  double acc = 0.0;
  for (uint32_t k = threadIdx.x; k < nb; k += 64) acc += partials[y * nb + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) out[dest[y]] = acc;
}

void launch_zexpect_group(int prec, const void* base, uint64_t S, unsigned count, const uint64_t* tz,
                          const double* tc, const uint64_t* trange, double* partials, unsigned blocks_per_state,
                          const uint64_t* dest, double* out, hipStream_t st)
{
  if (count == 0) return;
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_zexpect_group<double>), dim3(blocks_per_state, count), dim3(256), 0, st,
                       (const BV<double>::T*)base, S, tz, tc, trange, partials);
  else
    hipLaunchKernelGGL((k_zexpect_group<float>), dim3(blocks_per_state, count), dim3(256), 0, st,
                       (const BV<float>::T*)base, S, tz, tc, trange, partials);
  hipLaunchKernelGGL(k_sum_group, dim3(count), dim3(64), 0, st, partials, blocks_per_state, dest, out);
}

int setup_lds_batch(int prec, size_t lds_bytes)
{
  hipError_t e;
  if (prec == QSV_C128)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lds_batch<double>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  else
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lds_batch<float>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  return e == hipSuccess ? 0 : -1;
}

void launch_lds_batch(int prec, unsigned nblocks, size_t lds_bytes, const BInst* inst, const BRec* gates,
                      const uint64_t* tz, const double* tc, const int32_t* order, double* energies,
                      void* out_states, hipStream_t st)
{
  if (nblocks == 0) return;
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_lds_batch<double>), dim3(nblocks), dim3(kBatchThreads), lds_bytes, st, inst, gates, tz,
                       tc, order, energies, (BV<double>::T*)out_states);
  else
    hipLaunchKernelGGL((k_lds_batch<float>), dim3(nblocks), dim3(kBatchThreads), lds_bytes, st, inst, gates, tz,
                       tc, order, energies, (BV<float>::T*)out_states);
}

void launch_zexpect(int prec, const void* psi, uint64_t S, const uint64_t* tz, const double* tc, uint64_t nterms,
                    double* partials, double* out, hipStream_t st)
{
  uint64_t blocks = (S + 256 * 8 - 1) / (256 * 8);
  if (blocks > (uint64_t)kReduceBlocks) blocks = kReduceBlocks;
  if (blocks < 1) blocks = 1;
  if (prec == QSV_C128)
    hipLaunchKernelGGL((k_zexpect<double>), dim3((unsigned)blocks), dim3(256), 0, st,
                       (const BV<double>::T*)psi, S, tz, tc, nterms, partials);
  else
    hipLaunchKernelGGL((k_zexpect<float>), dim3((unsigned)blocks), dim3(256), 0, st,
                       (const BV<float>::T*)psi, S, tz, tc, nterms, partials);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, st, partials, (uint32_t)blocks, out);
}

}  // namespace qsv
