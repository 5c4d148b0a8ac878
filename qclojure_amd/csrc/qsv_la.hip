// qsv_la.hip — complex128 linear algebra on the device: the MatrixAlgebra /
// eigen-hermitian contract of src/org/soulspace/qclojure/domain/math/protocols.clj:81-375
// as kernels over row-major interleaved (re, im) matrices in HBM.
//
// * elementwise ops, transposes, Kronecker and outer products: one pass over
//   the output, HBM-bound;
// * matrix-multiply: FP64 MFMA (v_mfma_f64_16x16x4f64) on 64x64 output tiles
//   staged through LDS as split real/imaginary planes, four real products per
//   complex one;
// * matrix-vector product: one wave per row, HBM-bound;
// * reductions (inner product, norm, trace, predicate deviations): fixed grid,
//   FP64 partials, fixed-order final sum, so results repeat run to run;
// * solve / inverse: Gauss-Jordan with partial pivoting (the pivot rule and the
//   1e-14 singularity threshold of fastmath/complex_linear_algebra.clj:174-324),
//   four small launches per column, a device flag instead of a host sync;
// * eigen-hermitian: parallel cyclic Jacobi (round-robin pairing, n/2 disjoint
//   complex rotations per round), convergence checked once per sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "qsv_internal.h"
#include "../../include/qsv.h"

// Products are rounded and then added, as the JVM computes them (no fused
// multiply-add), so elementwise results equal the reference's bit for bit.
#pragma clang fp contract(off)

struct qsv_mat {
  int64_t rows = 0, cols = 0;
  double2* data = nullptr;
  bool owns = true;
  int device = 0;
  hipStream_t dep = nullptr;   // a viewed state's stream: finish its work before touching the data
};

namespace {

using qsv::api_fail;

constexpr int kLaBlock = 256;
constexpr int kLaReduceBlocks = 1024;

std::mutex g_la_mu;
hipStream_t g_la_stream = nullptr;
double* g_la_partials = nullptr;   // kLaReduceBlocks * 2
double* g_la_out = nullptr;        // 8 doubles
int* g_la_flag = nullptr;          // Gauss-Jordan singular flag + pivot index
double2* g_la_piv = nullptr;       // Gauss-Jordan pivot value of the current column

#define LACHK(x)                                                                              \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return api_fail(QSV_E_DEVICE, (std::string(#x) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

int la_setup()
{
  int rc = qsv::api_ensure_init();
  if (rc) return rc;
  if (g_la_stream) return QSV_OK;
  LACHK(hipStreamCreateWithFlags(&g_la_stream, hipStreamNonBlocking));
  LACHK(hipMalloc(&g_la_partials, sizeof(double) * kLaReduceBlocks * 2));
  LACHK(hipMalloc(&g_la_out, sizeof(double) * 8));
  LACHK(hipMalloc(&g_la_flag, sizeof(int) * 4));
  LACHK(hipMalloc(&g_la_piv, sizeof(double2)));
  return QSV_OK;
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double cabs_(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
// fastmath's complex division: (a * conj(b)) / |b|^2
__device__ inline double2 cdiv(double2 a, double2 b)
{
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// ---- elementwise --------------------------------------------------------------------------
__global__ void k_axpby(const double2* __restrict__ A, const double2* __restrict__ B, double2* __restrict__ C,
                        uint64_t n, double2 al, double2 be, int hasB)
{
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    double2 v = cmul(al, A[i]);
    if (hasB) v = cadd(v, cmul(be, B[i]));
    C[i] = v;
  }
}

__global__ void k_hadamard(const double2* __restrict__ A, const double2* __restrict__ B, double2* __restrict__ C,
                           uint64_t n)
{
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    C[i] = cmul(A[i], B[i]);
}

// 32x32 tiles through LDS (one padding column against bank conflicts).
__global__ void k_transpose(const double2* __restrict__ A, double2* __restrict__ C, int64_t rows, int64_t cols,
                            int conj)
{
  __shared__ double2 tile[32][33];
  const int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = by + r, j = bx + tx;
    if (i < rows && j < cols) tile[r][tx] = A[i * cols + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = bx + r, j = by + tx;   // C is cols x rows
    if (i < cols && j < rows) {
      double2 v = tile[tx][r];
      if (conj) v.y = -v.y;
      C[i * rows + j] = v;
    }
  }
}

// C = A (p x q) kron B (r x s): C[i][j] = A[i / r][j / s] * B[i % r][j % s].
__global__ void k_kron(const double2* __restrict__ A, const double2* __restrict__ B, double2* __restrict__ C,
                       int64_t p, int64_t q, int64_t r, int64_t s)
{
  const int64_t cols = q * s, total = p * r * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, j = e - i * cols;
    C[e] = cmul(A[(i / r) * q + j / s], B[(i % r) * s + j % s]);
  }
}

// C[i][j] = x[i] * conj(y[j])   (fastmath/complex_linear_algebra.clj:433-447)
__global__ void k_outer(const double2* __restrict__ x, const double2* __restrict__ y, double2* __restrict__ C,
                        int64_t m, int64_t n)
{
  const int64_t total = m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    C[e] = cmul(x[i], cconj(y[j]));
  }
}

// x <- (I..G_k..I) x for a 2x2 G on index bit b (a lazily kept Kronecker factor).
__global__ void k_apply2(double2* __restrict__ x, uint64_t half, int b, double2 g00, double2 g01, double2 g10,
                         double2 g11)
{
  const uint64_t lo = (1ull << b) - 1;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < half; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i0 = ((t & ~lo) << 1) | (t & lo), i1 = i0 | (1ull << b);
    const double2 a0 = x[i0], a1 = x[i1];
    x[i0] = cadd(cmul(g00, a0), cmul(g01, a1));
    x[i1] = cadd(cmul(g10, a0), cmul(g11, a1));
  }
}

// ---- GEMM on FP64 MFMA --------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int GBM = 64, GBN = 64, GBK = 16;

// C (M x N) = op(A) (M x K) * B (K x N); op = identity or conjugate transpose
// (then A is stored K x M). Four waves per workgroup, each owning a 32 x 32
// block of C as 2 x 2 MFMA tiles of 16 x 16 with separate real and imaginary
// accumulators. MFMA operand layout (16x16x4, FP64): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; it receives C[4 r + l / 16][l % 16]
// in accumulator element r (rows interleaved across the four lane groups).
template <bool CONJA>
__global__ __launch_bounds__(256, 3) void k_zgemm(const double2* __restrict__ A, const double2* __restrict__ B,
                                               double2* __restrict__ C, int M, int N, int K)
{
  __shared__ double sAr[GBK][GBM + 1], sAi[GBK][GBM + 1];
  __shared__ double sBr[GBK][GBN + 1], sBi[GBK][GBN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int i0 = blockIdx.y * GBM, j0 = blockIdx.x * GBN;
  v4d accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
      acci[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    }
  // The next K chunk is fetched into registers (4 elements of A and 4 of B
  // per thread) while the MFMAs consume the current one from LDS.
  double2 ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (CONJA) { kk = e / GBM; ii = e % GBM; }
      else { ii = e / GBK; kk = e % GBK; }
      const int gi = i0 + ii, gk = k0 + kk;
      ra[q] = make_double2(0.0, 0.0);
      if (gi < M && gk < K) ra[q] = CONJA ? A[(size_t)gk * M + gi] : A[(size_t)gi * K + gk];
      const int kb = e / GBN, jb = e % GBN;
      const int gkb = k0 + kb, gj = j0 + jb;
      rb[q] = make_double2(0.0, 0.0);
      if (gkb < K && gj < N) rb[q] = B[(size_t)gkb * N + gj];
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (CONJA) { kk = e / GBM; ii = e % GBM; }
      else { ii = e / GBK; kk = e % GBK; }
      sAr[kk][ii] = ra[q].x;
      sAi[kk][ii] = CONJA ? -ra[q].y : ra[q].y;
      const int kb = e / GBN, jb = e % GBN;
      sBr[kb][jb] = rb[q].x;
      sBi[kb][jb] = rb[q].y;
    }
  };
  fetch(0);
  stage();
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += GBK) {
    const bool more = k0 + GBK < K;
    if (more) fetch(k0 + GBK);
#pragma unroll
    for (int ks = 0; ks < GBK; ks += 4) {
      const int kk = ks + (lane >> 4), r16 = lane & 15;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int ii = wm * 32 + a * 16 + r16;
        ar[a] = sAr[kk][ii];
        ai[a] = sAi[kk][ii];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int jj = wn * 32 + b * 16 + r16;
        br[b] = sBr[kk][jj];
        bi[b] = sBi[kk][jj];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], accr[a][b], 0, 0, 0);
          accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], accr[a][b], 0, 0, 0);
          acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], acci[a][b], 0, 0, 0);
          acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acci[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
    if (more) {
      stage();
      __syncthreads();
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wm * 32 + a * 16 + 4 * r + (lane >> 4);
        const int gj = j0 + wn * 32 + b * 16 + (lane & 15);
        if (gi < M && gj < N) C[(size_t)gi * N + gj] = make_double2(accr[a][b][r], acci[a][b][r]);
      }
}

// y (m) = A (m x n) x (n): one wave per row, lanes stride the row.
__global__ void k_gemv(const double2* __restrict__ A, const double2* __restrict__ x, double2* __restrict__ y,
                       int64_t m, int64_t n)
{
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= m) return;
  const double2* a = A + row * n;
  double re = 0.0, im = 0.0;
  for (int64_t j = lane; j < n; j += 64) {
    const double2 u = a[j], v = x[j];
    re += u.x * v.x - u.y * v.y;
    im += u.x * v.y + u.y * v.x;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_xor(re, off);
    im += __shfl_xor(im, off);
  }
  if (lane == 0) y[row] = make_double2(re, im);
}

// ---- reductions ---------------------------------------------------------------------------
enum RedKind { R_DOT = 0, R_DOTC = 1, R_NORM2 = 2, R_TRACE = 3, R_DEV_HERM = 4, R_DEV_OFFDIAG = 5, R_DEV_IDENT = 6,
               R_OFFNORM = 7 };

__device__ inline void block_reduce2(double& a, double& b, bool is_max)
{
  __shared__ double sa[kLaBlock / 64], sb[kLaBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oa = __shfl_xor(a, off), ob = __shfl_xor(b, off);
    a = is_max ? fmax(a, oa) : a + oa;
    b = is_max ? fmax(b, ob) : b + ob;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kLaBlock / 64; ++i) {
      a = is_max ? fmax(a, sa[i]) : a + sa[i];
      b = is_max ? fmax(b, sb[i]) : b + sb[i];
    }
  }
}

// Partial (sum or max) over the elements of one matrix or a pair of vectors.
// x/y: operands; rows/cols: shape of x (vectors: count x 1).
__global__ void k_reduce(int kind, const double2* __restrict__ x, const double2* __restrict__ y, int64_t rows,
                         int64_t cols, double* __restrict__ partials)
{
  const bool is_max = kind >= R_DEV_HERM && kind <= R_DEV_IDENT;
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (kind == R_TRACE) {
    const int64_t k = rows < cols ? rows : cols;
    for (int64_t i = t0; i < k; i += stride) {
      const double2 v = x[i * cols + i];
      a += v.x;
      b += v.y;
    }
  } else if (kind <= R_NORM2) {
    const int64_t n = rows * cols;
    for (int64_t i = t0; i < n; i += stride) {
      const double2 u = x[i];
      if (kind == R_NORM2) {
        a += u.x * u.x + u.y * u.y;
      } else {
        const double2 v = y[i];
        const double ur = u.x, ui = kind == R_DOTC ? -u.y : u.y;
        a += ur * v.x - ui * v.y;
        b += ur * v.y + ui * v.x;
      }
    }
  } else {
    // deviations over the leading k x k block (k = rows, cols >= rows checked on the host)
    const int64_t k = kind == R_DEV_IDENT || kind == R_OFFNORM ? rows : (rows < cols ? rows : cols);
    const int64_t n = k * k;
    for (int64_t e = t0; e < n; e += stride) {
      const int64_t i = e / k, j = e - i * k;
      const double2 v = x[i * cols + j];
      if (kind == R_DEV_HERM) {
        const double2 w = x[j * cols + i];
        a = fmax(a, cabs_(make_double2(v.x - w.x, v.y + w.y)));
      } else if (kind == R_DEV_OFFDIAG) {
        if (i != j) a = fmax(a, cabs_(v));
      } else if (kind == R_DEV_IDENT) {
        a = fmax(a, cabs_(make_double2(v.x - (i == j ? 1.0 : 0.0), v.y)));
      } else {   // R_OFFNORM: off-diagonal and total squared Frobenius mass
        const double m2 = cabs2(v);
        if (i != j) a += m2;
        b += m2;
      }
    }
  }
  block_reduce2(a, b, is_max);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

__global__ void k_reduce_final(const double* __restrict__ partials, int nblocks, int is_max, double* __restrict__ out)
{
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
    a = is_max ? fmax(a, partials[2 * i]) : a + partials[2 * i];
    b = is_max ? fmax(b, partials[2 * i + 1]) : b + partials[2 * i + 1];
  }
  block_reduce2(a, b, is_max != 0);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

int reduce(int kind, const double2* x, const double2* y, int64_t rows, int64_t cols, double out[2])
{
  const bool is_max = kind >= R_DEV_HERM && kind <= R_DEV_IDENT;
  hipLaunchKernelGGL(k_reduce, dim3(kLaReduceBlocks), dim3(kLaBlock), 0, g_la_stream, kind, x, y, rows, cols,
                     g_la_partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kLaBlock), 0, g_la_stream, (const double*)g_la_partials,
                     kLaReduceBlocks, is_max ? 1 : 0, g_la_out);
  LACHK(hipGetLastError());
  LACHK(hipMemcpyAsync(out, g_la_out, 2 * sizeof(double), hipMemcpyDeviceToHost, g_la_stream));
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

// ---- Gauss-Jordan with partial pivoting on M = [A | B] (n x w) ------------------------------
// flag[0]: 0, or 1 + the first column whose pivot magnitude fell below 1e-14;
// flag[1]: the pivot row of the current column; piv: its value, saved before
// the swap kernel overwrites the column.
__global__ void k_gj_pivot(const double2* __restrict__ M, int64_t n, int64_t w, int64_t col, int* __restrict__ flag,
                           double2* __restrict__ piv)
{
  __shared__ double sv[kLaBlock];
  __shared__ int64_t si[kLaBlock];
  if (flag[0]) return;
  double best = -1.0;
  int64_t bi = col;
  for (int64_t r = col + threadIdx.x; r < n; r += blockDim.x) {
    const double v = cabs_(M[r * w + col]);
    if (v > best) { best = v; bi = r; }   // rows visited in increasing order per thread
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double ov = sv[threadIdx.x + s];
      const int64_t oi = si[threadIdx.x + s];
      // larger magnitude wins; equal magnitudes keep the lower row (the first maximum)
      if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (sv[0] < 1e-14) flag[0] = (int)(col + 1);
    flag[1] = (int)si[0];
    *piv = M[si[0] * w + col];
  }
}

// Swap rows col and p, then divide the pivot row by the pivot. Each thread
// owns whole columns c of both rows.
__global__ void k_gj_swap_scale(double2* __restrict__ M, int64_t w, int64_t col, const int* __restrict__ flag,
                                const double2* __restrict__ pivp)
{
  if (flag[0]) return;
  const int64_t p = flag[1];
  const double2 piv = *pivp;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < w; c += (int64_t)gridDim.x * blockDim.x) {
    const double2 top = M[col * w + c], other = M[p * w + c];
    if (p != col) M[p * w + c] = top;
    M[col * w + c] = cdiv(other, piv);
  }
}

__global__ void k_gj_factors(const double2* __restrict__ M, int64_t n, int64_t w, int64_t col,
                             double2* __restrict__ f, const int* __restrict__ flag)
{
  if (flag[0]) return;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    f[r] = r == col ? make_double2(0.0, 0.0) : M[r * w + col];
}

__global__ void k_gj_eliminate(double2* __restrict__ M, int64_t n, int64_t w, int64_t col,
                               const double2* __restrict__ f, const int* __restrict__ flag)
{
  if (flag[0]) return;
  const int64_t total = n * w;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / w, c = e - r * w;
    if (r == col) continue;
    if (c == col) M[e] = make_double2(0.0, 0.0);
    else M[e] = csub(M[e], cmul(f[r], M[col * w + c]));
  }
}

__global__ void k_gj_build(const double2* __restrict__ A, const double2* __restrict__ B, double2* __restrict__ M,
                           int64_t n, int64_t k, int identity)
{
  const int64_t w = n + k, total = n * w;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / w, c = e - r * w;
    if (c < n) M[e] = A[r * n + c];
    else if (identity) M[e] = make_double2(r == c - n ? 1.0 : 0.0, 0.0);
    else M[e] = B[r * k + (c - n)];
  }
}

__global__ void k_gj_extract(const double2* __restrict__ M, double2* __restrict__ X, int64_t n, int64_t k)
{
  const int64_t w = n + k, total = n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / k, c = e - r * k;
    X[e] = M[r * w + n + c];
  }
}

// ---- parallel cyclic Jacobi for Hermitian matrices -------------------------------------------
// Rotation for pair (p, q) with b = M[p][q] = |b| e^{i phi}: U = D J with
// D = diag(1, e^{-i phi}) (makes the pair block real) and the symmetric Schur
// rotation J = [[c, s], [-s, c]] (tan of the angle t = sgn(tau)/(|tau| + sqrt(1 + tau^2)),
// tau = (M_qq - M_pp) / (2 |b|)). A <- U^H A U, V <- V U.
struct Rot {
  double c, s, er, ei;   // E = e^{-i phi}
};

__global__ void k_jac_params(const double2* __restrict__ M, int64_t n, const int2* __restrict__ pairs, int npairs,
                             Rot* __restrict__ rot)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const int p = pairs[k].x, q = pairs[k].y;
  Rot R{1.0, 0.0, 1.0, 0.0};
  if (p < n && q < n) {
    // the Hermitian part of the pair's off-diagonal entry (rounding leaves
    // M_qp and conj(M_pq) a few ulps apart)
    const double2 bu = M[(int64_t)p * n + q], bl = M[(int64_t)q * n + p];
    const double2 b = make_double2(0.5 * (bu.x + bl.x), 0.5 * (bu.y - bl.y));
    // |b| and the phase e^{-i phi} with scaling, so that tiny entries (whose
    // squares would be subnormal) still give a phase of modulus exactly 1 up
    // to rounding; entries below 1e-290 are treated as zero.
    const double sc = fmax(fabs(b.x), fabs(b.y));
    if (sc > 1e-290) {
      const double bx = b.x / sc, by = b.y / sc;
      const double r = sqrt(bx * bx + by * by);
      const double ab = sc * r;
      const double app = M[(int64_t)p * n + p].x, aqq = M[(int64_t)q * n + q].x;
      const double tau = (aqq - app) / (2.0 * ab);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      const double c = 1.0 / sqrt(1.0 + t * t);
      R = Rot{c, t * c, bx / r, -by / r};
    }
  }
  rot[k] = R;
}

// Rows: A_p <- c A_p - s conj(E) A_q ; A_q <- s A_p + c conj(E) A_q.
__global__ void k_jac_rows(double2* __restrict__ M, int64_t n, const int2* __restrict__ pairs, int npairs,
                           const Rot* __restrict__ rot)
{
  const int64_t total = (int64_t)npairs * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / n);
    const int64_t j = e - (int64_t)k * n;
    const int p = pairs[k].x, q = pairs[k].y;
    if (p >= n || q >= n) continue;
    const Rot R = rot[k];
    const double2 Ec = make_double2(R.er, -R.ei);
    const double2 ap = M[(int64_t)p * n + j], aq = M[(int64_t)q * n + j];
    const double2 eq = cmul(Ec, aq);
    M[(int64_t)p * n + j] = make_double2(R.c * ap.x - R.s * eq.x, R.c * ap.y - R.s * eq.y);
    M[(int64_t)q * n + j] = make_double2(R.s * ap.x + R.c * eq.x, R.s * ap.y + R.c * eq.y);
  }
}

// Columns of A and V: X_:p <- c X_:p - s E X_:q ; X_:q <- s X_:p + c E X_:q.
// The pair's off-diagonal entries, which the rotation annihilates, are stored
// as exact zeros and its diagonal entries as real numbers (thread (p, k) owns
// row p of the pair's two columns, thread (q, k) row q).
__global__ void k_jac_cols(double2* __restrict__ M, double2* __restrict__ V, int64_t n,
                           const int2* __restrict__ pairs, int npairs, const Rot* __restrict__ rot)
{
  const int64_t total = (int64_t)npairs * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / npairs;
    const int k = (int)(e - i * npairs);
    const int p = pairs[k].x, q = pairs[k].y;
    if (p >= n || q >= n) continue;
    const Rot R = rot[k];
    const double2 E = make_double2(R.er, R.ei);
    double2* X[2] = {M, V};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const double2 xp = X[m][i * n + p], xq = X[m][i * n + q];
      const double2 eq = cmul(E, xq);
      double2 np_ = make_double2(R.c * xp.x - R.s * eq.x, R.c * xp.y - R.s * eq.y);
      double2 nq_ = make_double2(R.s * xp.x + R.c * eq.x, R.s * xp.y + R.c * eq.y);
      if (m == 0 && i == p) { np_.y = 0.0; nq_ = make_double2(0.0, 0.0); }
      if (m == 0 && i == q) { np_ = make_double2(0.0, 0.0); nq_.y = 0.0; }
      X[m][i * n + p] = np_;
      X[m][i * n + q] = nq_;
    }
  }
}

__global__ void k_identity(double2* __restrict__ V, int64_t n)
{
  const int64_t total = n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
}

// The whole Jacobi iteration for n <= kJacSmall in one workgroup, A and V in
// LDS: the same rotations, round-robin schedule and stopping rule as the
// multi-launch path, with barriers in place of launches.
constexpr int kJacSmall = 64;
constexpr int kJacPad = kJacSmall + 1;   // row stride: a column walk crosses LDS banks

__global__ __launch_bounds__(256) void k_jacobi_small(const double2* __restrict__ A, double2* __restrict__ Vout,
                                                      double* __restrict__ w, int n, int* __restrict__ sweeps_out)
{
  __shared__ double2 sM[kJacSmall][kJacPad];
  __shared__ double2 sV[kJacSmall][kJacPad];
  __shared__ Rot rot[kJacSmall / 2];
  __shared__ int2 pr[kJacSmall / 2];
  __shared__ double ra[kLaBlock / 64], rb[kLaBlock / 64];
  __shared__ int stop;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    sM[i][j] = A[e];
    sV[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  const int N = n + (n & 1), npairs = N / 2, nrounds = N - 1;
  double prev = -1.0;
  int sweeps = 0;
  __syncthreads();
  for (; sweeps < 40; ++sweeps) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < n * n; e += blockDim.x) {
      const int i = e / n, j = e - i * n;
      const double m2 = cabs2(sM[i][j]);
      tot += m2;
      if (i != j) off += m2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o);
      tot += __shfl_xor(tot, o);
    }
    if ((tid & 63) == 0) { ra[tid >> 6] = off; rb[tid >> 6] = tot; }
    __syncthreads();
    if (tid == 0) {
      double o = 0.0, t = 0.0;
      for (int i = 0; i < kLaBlock / 64; ++i) { o += ra[i]; t += rb[i]; }
      stop = (!(o > 1e-30 * t) || o == 0.0 || (prev >= 0.0 && !(o < prev))) ? 1 : 0;
      ra[0] = o;
    }
    __syncthreads();
    if (stop) break;
    prev = ra[0];
    __syncthreads();
    for (int r = 0; r < nrounds; ++r) {
      if (tid < npairs) {
        // circle method: position 0 fixed, the others rotated by r
        auto player = [&](int pos) { return pos == 0 ? 0 : ((pos - 1 - r) % (N - 1) + (N - 1)) % (N - 1) + 1; };
        int p = player(tid), q = player(N - 1 - tid);
        if (p > q) { const int t = p; p = q; q = t; }
        pr[tid] = make_int2(p, q);
        Rot R{1.0, 0.0, 1.0, 0.0};
        if (q < n) {
          const double2 bu = sM[p][q], bl = sM[q][p];
          const double2 b = make_double2(0.5 * (bu.x + bl.x), 0.5 * (bu.y - bl.y));
          const double sc = fmax(fabs(b.x), fabs(b.y));
          if (sc > 1e-290) {
            const double bx = b.x / sc, by = b.y / sc;
            const double rr = sqrt(bx * bx + by * by);
            const double ab = sc * rr;
            const double tau = (sM[q][q].x - sM[p][p].x) / (2.0 * ab);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + t * t);
            R = Rot{c, t * c, bx / rr, -by / rr};
          }
        }
        rot[tid] = R;
      }
      __syncthreads();
      for (int e = tid; e < npairs * n; e += blockDim.x) {
        const int k = e / n, j = e - k * n;
        const int p = pr[k].x, q = pr[k].y;
        if (q >= n) continue;
        const Rot R = rot[k];
        const double2 eq = cmul(make_double2(R.er, -R.ei), sM[q][j]);
        const double2 ap = sM[p][j];
        sM[p][j] = make_double2(R.c * ap.x - R.s * eq.x, R.c * ap.y - R.s * eq.y);
        sM[q][j] = make_double2(R.s * ap.x + R.c * eq.x, R.s * ap.y + R.c * eq.y);
      }
      __syncthreads();
      for (int e = tid; e < npairs * n; e += blockDim.x) {
        const int i = e / npairs, k = e - i * npairs;
        const int p = pr[k].x, q = pr[k].y;
        if (q >= n) continue;
        const Rot R = rot[k];
        const double2 E = make_double2(R.er, R.ei);
        {
          const double2 xp = sM[i][p], eq = cmul(E, sM[i][q]);
          double2 np_ = make_double2(R.c * xp.x - R.s * eq.x, R.c * xp.y - R.s * eq.y);
          double2 nq_ = make_double2(R.s * xp.x + R.c * eq.x, R.s * xp.y + R.c * eq.y);
          if (i == p) { np_.y = 0.0; nq_ = make_double2(0.0, 0.0); }
          if (i == q) { np_ = make_double2(0.0, 0.0); nq_.y = 0.0; }
          sM[i][p] = np_;
          sM[i][q] = nq_;
        }
        {
          const double2 xp = sV[i][p], eq = cmul(E, sV[i][q]);
          sV[i][p] = make_double2(R.c * xp.x - R.s * eq.x, R.c * xp.y - R.s * eq.y);
          sV[i][q] = make_double2(R.s * xp.x + R.c * eq.x, R.s * xp.y + R.c * eq.y);
        }
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    Vout[e] = sV[i][j];
  }
  for (int i = tid; i < n; i += blockDim.x) w[i] = sM[i][i].x;
  if (tid == 0) *sweeps_out = sweeps;
}

// eigen-general of a square complex matrix (n <= kJacSmall), one wavefront:
// Householder reduction to upper Hessenberg form H = Q^H A Q, then the
// shifted QR algorithm (implicit single shift, Wilkinson's shift from the
// trailing 2x2, an exceptional shift every tenth iteration, deflation where
// |h(l, l-1)| <= eps (|h(l-1, l-1)| + |h(l, l)|)) to the Schur form T = Z^H A Z,
// then eigenvectors by back-substitution in T (column k of Y solves
// (T - t_kk I) y = 0 with y_k = 1, y_j = 0 below) and X = Z Y, columns of unit
// length. Lane i owns row or column i of each update. w: the eigenvalues
// t_kk as (re, im) pairs; Y: n x n scratch; fail[0] = 1 when QR did not
// converge within 30 iterations per eigenvalue.
__device__ inline double cabs1(double2 a) { return hypot(a.x, a.y); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 csqrt1(double2 a)
{
  const double r = cabs1(a);
  if (r == 0.0) return make_double2(0.0, 0.0);
  const double re = sqrt(0.5 * (r + fabs(a.x)));
  const double im = 0.5 * a.y / re;
  return a.x >= 0.0 ? make_double2(re, im) : make_double2(fabs(im), a.y >= 0.0 ? re : -re);
}

__global__ __launch_bounds__(64) void k_eig_general(const double2* __restrict__ A, double2* __restrict__ Xout,
                                                    double2* __restrict__ Y, double* __restrict__ w, int n,
                                                    int* __restrict__ fail)
{
  __shared__ double2 sH[kJacSmall][kJacPad];
  __shared__ double2 sZ[kJacSmall][kJacPad];
  __shared__ double2 sv[kJacSmall];
  __shared__ double2 sw[kJacSmall];
  const int l = threadIdx.x;
  for (int e = l; e < n * n; e += 64) {
    const int i = e / n, j = e - i * n;
    sH[i][j] = A[e];
    sZ[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // ---- Householder reduction: column k below the subdiagonal -> 0
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;   // length of x = H[k+1 .. n-1][k]
    double nx = 0.0, tail = 0.0;
    for (int i = 0; i < m; ++i) {
      const double a2 = cabs2(sH[k + 1 + i][k]);
      nx += a2;
      if (i) tail += a2;
    }
    if (tail == 0.0) continue;   // already Hessenberg in this column (all lanes agree)
    nx = sqrt(nx);
    const double2 x0 = sH[k + 1][k];
    const double a0 = cabs1(x0);
    const double2 ph = a0 > 0.0 ? cscale(x0, 1.0 / a0) : make_double2(1.0, 0.0);
    const double2 alpha = cscale(ph, -nx);                // x -> alpha e_0
    // v = x - alpha e_0, normalised
    if (l < m) sv[l] = l == 0 ? csub(x0, alpha) : sH[k + 1 + l][k];
    __syncthreads();
    double nv = 0.0;
    for (int i = 0; i < m; ++i) nv += cabs2(sv[i]);
    const double inv = 1.0 / sqrt(nv);
    __syncthreads();
    if (l < m) sv[l] = cscale(sv[l], inv);
    __syncthreads();
    // H <- P H (P = I - 2 v v^H on rows k+1 ..): lane j takes column j
    if (l < n && l >= k) {
      double2 t = make_double2(0.0, 0.0);
      for (int i = 0; i < m; ++i) t = cadd(t, cmul(cconj(sv[i]), sH[k + 1 + i][l]));
      for (int i = 0; i < m; ++i) sH[k + 1 + i][l] = csub(sH[k + 1 + i][l], cscale(cmul(sv[i], t), 2.0));
    }
    __syncthreads();
    // H <- H P and Z <- Z P: lane i takes row i
    if (l < n) {
      double2 t = make_double2(0.0, 0.0), u = make_double2(0.0, 0.0);
      for (int j = 0; j < m; ++j) {
        t = cadd(t, cmul(sH[l][k + 1 + j], sv[j]));
        u = cadd(u, cmul(sZ[l][k + 1 + j], sv[j]));
      }
      for (int j = 0; j < m; ++j) {
        sH[l][k + 1 + j] = csub(sH[l][k + 1 + j], cscale(cmul(t, cconj(sv[j])), 2.0));
        sZ[l][k + 1 + j] = csub(sZ[l][k + 1 + j], cscale(cmul(u, cconj(sv[j])), 2.0));
      }
    }
    __syncthreads();
    if (l < m && l >= 1) sH[k + 1 + l][k] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  // ---- shifted QR on the Hessenberg matrix, deflating from the bottom
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iter = 0;
  bool ok = true;
  while (hi > 0) {
    int lo = hi;
    while (lo > 0) {
      const double s = cabs1(sH[lo - 1][lo - 1]) + cabs1(sH[lo][lo]);
      if (cabs1(sH[lo][lo - 1]) <= eps * (s > 0.0 ? s : 1.0)) break;
      --lo;
    }
    __syncthreads();
    if (lo > 0 && l == 0) sH[lo][lo - 1] = make_double2(0.0, 0.0);
    __syncthreads();
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 30) {
      ok = false;
      break;
    }
    // Wilkinson's shift: the eigenvalue of the trailing 2x2 nearer d
    const double2 a = sH[hi - 1][hi - 1], b = sH[hi - 1][hi], c = sH[hi][hi - 1], d = sH[hi][hi];
    double2 mu;
    if (iter % 10 == 0) {
      mu = cadd(d, make_double2(cabs1(c), 0.0));   // exceptional shift
    } else {
      const double2 hd = cscale(csub(a, d), 0.5);
      const double2 disc = csqrt1(cadd(cmul(hd, hd), cmul(b, c)));
      const double2 m1 = cadd(cscale(cadd(a, d), 0.5), disc), m2 = csub(cscale(cadd(a, d), 0.5), disc);
      mu = cabs2(csub(m1, d)) <= cabs2(csub(m2, d)) ? m1 : m2;
    }
    // implicit single-shift QR sweep over rows lo .. hi
    double2 x = csub(sH[lo][lo], mu), y = sH[lo + 1][lo];
    for (int k = lo; k < hi; ++k) {
      if (k > lo) {
        x = sH[k][k - 1];
        y = sH[k + 1][k - 1];
      }
      // G = [c s; -conj(s) c] with G (x, y) = (r, 0), c real
      const double ax = cabs1(x), r = hypot(ax, cabs1(y));
      double cr;
      double2 sg;
      if (r == 0.0) {
        cr = 1.0;
        sg = make_double2(0.0, 0.0);
      } else if (ax == 0.0) {
        cr = 0.0;
        sg = cscale(cconj(y), 1.0 / cabs1(y));
      } else {
        cr = ax / r;
        sg = cscale(cmul(cscale(x, 1.0 / ax), cconj(y)), 1.0 / r);
      }
      __syncthreads();
      // rows k, k+1 of H (columns max(lo, k-1) .. n-1): lane j takes column j
      const int j0 = k > lo ? k - 1 : lo;
      if (l < n && l >= j0) {
        const double2 p = sH[k][l], q = sH[k + 1][l];
        sH[k][l] = cadd(cscale(p, cr), cmul(sg, q));
        sH[k + 1][l] = csub(cscale(q, cr), cmul(cconj(sg), p));
      }
      __syncthreads();
      // columns k, k+1 of H (rows 0 .. min(k+2, hi)) and of Z (all rows): G^H on the right
      const int i1 = k + 2 < hi ? k + 2 : hi;
      if (l < n) {
        if (l <= i1) {
          const double2 p = sH[l][k], q = sH[l][k + 1];
          sH[l][k] = cadd(cscale(p, cr), cmul(q, cconj(sg)));
          sH[l][k + 1] = csub(cscale(q, cr), cmul(p, sg));
        }
        const double2 p = sZ[l][k], q = sZ[l][k + 1];
        sZ[l][k] = cadd(cscale(p, cr), cmul(q, cconj(sg)));
        sZ[l][k + 1] = csub(cscale(q, cr), cmul(p, sg));
      }
      __syncthreads();
    }
    // the bulge of the chase is removed: keep the subdiagonal band exact
    if (l < n && l >= lo + 2 && l <= hi) sH[l][l - 2] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  if (l == 0) fail[0] = ok ? 0 : 1;
  if (l < n) sw[l] = sH[l][l];
  __syncthreads();
  // ---- eigenvectors of T (lane k: column k of Y, by back-substitution)
  double tn = 0.0;
  for (int i = 0; i < n; ++i) tn = fmax(tn, cabs1(sH[i][i]));
  const double small = eps * (tn > 0.0 ? tn : 1.0);
  if (l < n) {
    const int k = l;
    const double2 lam = sw[k];
    for (int i = n - 1; i > k; --i) Y[(int64_t)i * n + k] = make_double2(0.0, 0.0);
    Y[(int64_t)k * n + k] = make_double2(1.0, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double2 t = make_double2(0.0, 0.0);
      for (int j = i + 1; j <= k; ++j) t = cadd(t, cmul(sH[i][j], Y[(int64_t)j * n + k]));
      double2 den = csub(sH[i][i], lam);
      if (cabs1(den) < small) den = make_double2(small, 0.0);
      Y[(int64_t)i * n + k] = cdiv(make_double2(-t.x, -t.y), den);
    }
    // X = Z Y, column k, unit length
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      double2 t = make_double2(0.0, 0.0);
      for (int j = 0; j <= k; ++j) t = cadd(t, cmul(sZ[i][j], Y[(int64_t)j * n + k]));
      Xout[(int64_t)i * n + k] = t;
      nrm += cabs2(t);
    }
    const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 1.0;
    for (int i = 0; i < n; ++i) Xout[(int64_t)i * n + k] = cscale(Xout[(int64_t)i * n + k], inv);
    w[2 * k] = lam.x;
    w[2 * k + 1] = lam.y;
  }
}

__global__ void k_diag_real(const double2* __restrict__ M, int64_t n, double* __restrict__ w)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = M[i * n + i].x;
}

unsigned grid_for(uint64_t n)
{
  const uint64_t b = (n + kLaBlock - 1) / kLaBlock;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 65536));
}

int check_mat(const qsv_mat* m, const char* what)
{
  if (!m || !m->data) return api_fail(QSV_E_INVALID, (std::string("null matrix: ") + what).c_str());
  // a view of an engine state: its own stream may still be working on it
  if (m->dep && hipStreamSynchronize(m->dep) != hipSuccess)
    return api_fail(QSV_E_DEVICE, "waiting for the viewed state's stream failed");
  return QSV_OK;
}

int alloc(double2** p, uint64_t count)
{
  if (hipMalloc(p, std::max<uint64_t>(count, 1) * sizeof(double2)) != hipSuccess)
    return api_fail(QSV_E_NOMEM, ("hipMalloc of " + std::to_string(count * 16) + " B failed").c_str());
  return QSV_OK;
}

}  // namespace

// ---- LU, QR and Cholesky: one workgroup each ----------------------------------------------
// The factorizations are column-sequential, so each runs as a single
// workgroup that loops over the columns with barriers between the steps
// (every intermediate in HBM/L2); meant for the facade's matrix sizes.
constexpr int kDecBlock = 1024;

// Sum of v over the workgroup, in a fixed order; every thread gets it.
__device__ double2 wg_sum(double2 v, double2* red)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o);
    v.y += __shfl_xor(v.y, o);
  }
  __syncthreads();   // red may still be read from the previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double2 s = make_double2(0.0, 0.0);
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = cadd(s, red[w]);
  return s;
}

// Cholesky A = L Lᴴ as fastmath/complex_linear_algebra.clj:1119-1166 defines
// it: L[i][j] = (A[i][j] - Σ_{k<j} L[i][k] conj(L[j][k])) / L[j][j] below the
// diagonal, L[j][j] = sqrt(A[j][j] - Σ_k |L[j][k]|^2), which must be real
// (|imag| <= eps, else status 1) and at least eps (else status 2); status[1]
// = the failing index. Column by column: the first failing diagonal is the
// one the reference's row-by-row order reaches first.
__global__ __launch_bounds__(kDecBlock) void k_cholesky(const double2* __restrict__ A, double2* __restrict__ L,
                                                        int n, double eps, int* __restrict__ status)
{
  __shared__ double2 red[kDecBlock / 64];
  __shared__ double2 djj;
  __shared__ int bad;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int64_t e = t; e < (int64_t)n * n; e += nt) L[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    double2 s = make_double2(0.0, 0.0);
    for (int k = t; k < j; k += nt) s.x += cabs2(L[(int64_t)j * n + k]);
    s = wg_sum(s, red);
    if (t == 0) {
      const double2 diff = csub(A[(int64_t)j * n + j], s);
      int b = 0;
      if (fabs(diff.y) > eps) b = 1;
      else if (diff.x < eps) b = 2;
      if (b) {
        status[0] = b;
        status[1] = j;
      } else {
        djj = make_double2(sqrt(diff.x), 0.0);
        L[(int64_t)j * n + j] = djj;
      }
      bad = b;
    }
    __syncthreads();
    if (bad) return;
    for (int i = j + 1 + t; i < n; i += nt) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < j; ++k) acc = cadd(acc, cmul(L[(int64_t)i * n + k], cconj(L[(int64_t)j * n + k])));
      L[(int64_t)i * n + j] = cdiv(csub(A[(int64_t)i * n + j], acc), djj);
    }
    __syncthreads();
  }
}

// LU with partial pivoting on M (n x n, overwritten: multipliers below the
// diagonal, U on and above it): the pivot of column k is the first row of
// largest |M[i][k]|, i >= k; perm[k] = the row swapped into row k. A pivot
// below 1e-11 (the singularity threshold of Apache Commons Math's
// LUDecomposition, which the reference calls) sets status[0] = 1, status[1] = k.
__global__ __launch_bounds__(kDecBlock) void k_lu(double2* __restrict__ M, int n, int* __restrict__ perm,
                                                  int* __restrict__ status)
{
  __shared__ double bv[kDecBlock / 64];
  __shared__ int bi[kDecBlock / 64];
  __shared__ int prow;
  __shared__ int bad;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < n; ++k) {
    double v = -1.0;
    int idx = n;
    for (int i = k + t; i < n; i += nt) {
      const double a = cabs_(M[(int64_t)i * n + k]);
      if (a > v) { v = a; idx = i; }   // rows ascending per thread: the first of equal values stays
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o);
      const int oi = __shfl_xor(idx, o);
      if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if ((t & 63) == 0) { bv[t >> 6] = v; bi[t >> 6] = idx; }
    __syncthreads();
    if (t == 0) {
      double bvv = -1.0;
      int bii = n;
      for (int w = 0; w < nt / 64; ++w)
        if (bv[w] > bvv || (bv[w] == bvv && bi[w] < bii)) { bvv = bv[w]; bii = bi[w]; }
      bad = !(bvv >= 1e-11);
      if (bad) { status[0] = 1; status[1] = k; }
      prow = bii;
      perm[k] = bii;
    }
    __syncthreads();
    if (bad) return;
    const int p = prow;
    if (p != k)
      for (int j = t; j < n; j += nt) {
        const double2 a = M[(int64_t)k * n + j];
        M[(int64_t)k * n + j] = M[(int64_t)p * n + j];
        M[(int64_t)p * n + j] = a;
      }
    __syncthreads();
    const double2 piv = M[(int64_t)k * n + k];
    for (int i = k + 1 + t; i < n; i += nt) M[(int64_t)i * n + k] = cdiv(M[(int64_t)i * n + k], piv);
    __syncthreads();
    const int64_t w = n - k - 1;
    for (int64_t e = t; e < w * w; e += nt) {
      const int64_t i = k + 1 + e / w, j = k + 1 + e % w;
      M[i * n + j] = csub(M[i * n + j], cmul(M[i * n + k], M[(int64_t)k * n + j]));
    }
    __syncthreads();
  }
}

// L (unit lower) and U (upper) out of the overwritten LU matrix.
__global__ void k_lu_split(const double2* __restrict__ M, double2* __restrict__ L, double2* __restrict__ U, int64_t n)
{
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    const double2 z = make_double2(0.0, 0.0);
    L[e] = i > j ? M[e] : (i == j ? make_double2(1.0, 0.0) : z);
    U[e] = i <= j ? M[e] : z;
  }
}

// QR by modified Gram-Schmidt as fastmath/complex_linear_algebra.clj:1048-1102:
// for each column j, v = A[:, j]; for i < j: R[i][j] = <q_i, v> (conjugating
// q_i), v -= R[i][j] q_i; R[j][j] = ||v||, q_j = v / ||v||, or the unit vector
// e_j (zero when j >= m) when ||v|| < 1e-12. Q is m x n, R is n x n.
__global__ __launch_bounds__(kDecBlock) void k_qr_mgs(const double2* __restrict__ A, double2* __restrict__ Q,
                                                      double2* __restrict__ R, int m, int n)
{
  __shared__ double2 red[kDecBlock / 64];
  const int t = threadIdx.x, nt = blockDim.x;
  for (int64_t e = t; e < (int64_t)n * n; e += nt) R[e] = make_double2(0.0, 0.0);
  for (int j = 0; j < n; ++j) {
    for (int r = t; r < m; r += nt) Q[(int64_t)r * n + j] = A[(int64_t)r * n + j];
    __syncthreads();
    for (int i = 0; i < j; ++i) {
      double2 s = make_double2(0.0, 0.0);
      for (int r = t; r < m; r += nt) s = cadd(s, cmul(cconj(Q[(int64_t)r * n + i]), Q[(int64_t)r * n + j]));
      s = wg_sum(s, red);
      if (t == 0) R[(int64_t)i * n + j] = s;
      for (int r = t; r < m; r += nt) Q[(int64_t)r * n + j] = csub(Q[(int64_t)r * n + j], cmul(s, Q[(int64_t)r * n + i]));
      __syncthreads();
    }
    double2 s = make_double2(0.0, 0.0);
    for (int r = t; r < m; r += nt) s.x += cabs2(Q[(int64_t)r * n + j]);
    s = wg_sum(s, red);
    const double nrm = sqrt(s.x);
    if (t == 0) R[(int64_t)j * n + j] = make_double2(nrm, 0.0);
    for (int r = t; r < m; r += nt) {
      const int64_t e = (int64_t)r * n + j;
      Q[e] = nrm < 1e-12 ? make_double2(r == j ? 1.0 : 0.0, 0.0) : cdiv(Q[e], make_double2(nrm, 0.0));
    }
    __syncthreads();
  }
}

extern "C" {

int qsv_mat_create(int64_t rows, int64_t cols, qsv_mat** out)
{
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  *out = nullptr;
  if (rows < 1 || cols < 1) return api_fail(QSV_E_INVALID, "matrix dimensions must be positive");
  std::lock_guard<std::mutex> lk(g_la_mu);
  int rc = la_setup();
  if (rc) return rc;
  auto* m = new qsv_mat();
  m->rows = rows;
  m->cols = cols;
  if ((rc = alloc(&m->data, (uint64_t)rows * cols))) {
    delete m;
    return rc;
  }
  if (hipMemsetAsync(m->data, 0, (size_t)rows * cols * sizeof(double2), g_la_stream) != hipSuccess ||
      hipStreamSynchronize(g_la_stream) != hipSuccess) {
    hipFree(m->data);
    delete m;
    return api_fail(QSV_E_DEVICE, "clearing a new matrix failed");
  }
  *out = m;
  return QSV_OK;
}

int qsv_mat_destroy(qsv_mat* m)
{
  if (!m) return QSV_OK;
  if (m->owns && m->data) hipFree(m->data);
  delete m;
  return QSV_OK;
}

int qsv_mat_shape(const qsv_mat* m, int64_t* rows, int64_t* cols)
{
  if (!m || !rows || !cols) return api_fail(QSV_E_INVALID, "null argument");
  *rows = m->rows;
  *cols = m->cols;
  return QSV_OK;
}

int qsv_mat_upload(qsv_mat* m, const double* host)
{
  int rc = check_mat(m, "upload");
  if (rc) return rc;
  if (!host) return api_fail(QSV_E_INVALID, "null host buffer");
  std::lock_guard<std::mutex> lk(g_la_mu);
  LACHK(hipMemcpyAsync(m->data, host, (size_t)m->rows * m->cols * sizeof(double2), hipMemcpyHostToDevice,
                       g_la_stream));
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_mat_download(const qsv_mat* m, double* host)
{
  int rc = check_mat(m, "download");
  if (rc) return rc;
  if (!host) return api_fail(QSV_E_INVALID, "null host buffer");
  std::lock_guard<std::mutex> lk(g_la_mu);
  LACHK(hipMemcpyAsync(host, m->data, (size_t)m->rows * m->cols * sizeof(double2), hipMemcpyDeviceToHost,
                       g_la_stream));
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_mat_copy(const qsv_mat* src, qsv_mat* dst)
{
  int rc = check_mat(src, "copy source");
  if (rc || (rc = check_mat(dst, "copy destination"))) return rc;
  if (src->rows * src->cols != dst->rows * dst->cols) return api_fail(QSV_E_INVALID, "copy: element counts differ");
  std::lock_guard<std::mutex> lk(g_la_mu);
  LACHK(hipMemcpyAsync(dst->data, src->data, (size_t)src->rows * src->cols * sizeof(double2),
                       hipMemcpyDeviceToDevice, g_la_stream));
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_mat_view_state(qsv_state* s, qsv_mat** out)
{
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  void* amps = nullptr;
  hipStream_t st = nullptr;
  int n = 0, prec = 0;
  int rc = qsv::api_state_raw(s, &amps, &st, &n, &prec);
  if (rc) return rc;
  if (prec != QSV_C128) return api_fail(QSV_E_UNSUPPORTED, "matrix views need a complex128 state");
  std::lock_guard<std::mutex> lk(g_la_mu);
  if ((rc = la_setup())) return rc;
  LACHK(hipStreamSynchronize(st));
  auto* m = new qsv_mat();
  m->rows = (int64_t)1 << n;
  m->cols = 1;
  m->data = (double2*)amps;
  m->owns = false;
  m->dep = st;
  *out = m;
  return QSV_OK;
}

int qsv_la_axpby(double ar, double ai, const qsv_mat* A, double br, double bi, const qsv_mat* B, qsv_mat* C)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(C, "C")) || (B && (rc = check_mat(B, "B")))) return rc;
  if (B && (B->rows != A->rows || B->cols != A->cols))
    return api_fail(QSV_E_INVALID, "axpby: A and B shapes differ");
  if (C->rows != A->rows || C->cols != A->cols) return api_fail(QSV_E_INVALID, "axpby: C shape differs");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const uint64_t n = (uint64_t)A->rows * A->cols;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kLaBlock), 0, g_la_stream, (const double2*)A->data,
                     (const double2*)(B ? B->data : A->data), C->data, n, make_double2(ar, ai), make_double2(br, bi),
                     B ? 1 : 0);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_hadamard(const qsv_mat* A, const qsv_mat* B, qsv_mat* C)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(B, "B")) || (rc = check_mat(C, "C"))) return rc;
  if (B->rows != A->rows || B->cols != A->cols || C->rows != A->rows || C->cols != A->cols)
    return api_fail(QSV_E_INVALID, "hadamard: shapes differ");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const uint64_t n = (uint64_t)A->rows * A->cols;
  hipLaunchKernelGGL(k_hadamard, dim3(grid_for(n)), dim3(kLaBlock), 0, g_la_stream, (const double2*)A->data,
                     (const double2*)B->data, C->data, n);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_gemm(int conj_transpose_a, const qsv_mat* A, const qsv_mat* B, qsv_mat* C)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(B, "B")) || (rc = check_mat(C, "C"))) return rc;
  const int64_t M = conj_transpose_a ? A->cols : A->rows, K = conj_transpose_a ? A->rows : A->cols;
  if (B->rows != K) return api_fail(QSV_E_INVALID, "matrix-multiply: inner dimensions differ");
  if (C->rows != M || C->cols != B->cols) return api_fail(QSV_E_INVALID, "matrix-multiply: C shape differs");
  if (M > (1 << 30) || K > (1 << 30) || B->cols > (1 << 30) || (M + GBM - 1) / GBM > 65535)
    return api_fail(QSV_E_INVALID, "matrix-multiply: dimension too large");
  if (C->data == A->data || C->data == B->data) return api_fail(QSV_E_INVALID, "matrix-multiply: C aliases A or B");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const dim3 grid((unsigned)((B->cols + GBN - 1) / GBN), (unsigned)((M + GBM - 1) / GBM));
  if (conj_transpose_a)
    hipLaunchKernelGGL(k_zgemm<true>, grid, dim3(256), 0, g_la_stream, (const double2*)A->data,
                       (const double2*)B->data, C->data, (int)M, (int)B->cols, (int)K);
  else
    hipLaunchKernelGGL(k_zgemm<false>, grid, dim3(256), 0, g_la_stream, (const double2*)A->data,
                       (const double2*)B->data, C->data, (int)M, (int)B->cols, (int)K);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_gemv(const qsv_mat* A, const qsv_mat* x, qsv_mat* y)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(x, "x")) || (rc = check_mat(y, "y"))) return rc;
  if (x->rows * x->cols != A->cols) return api_fail(QSV_E_INVALID, "matrix-vector-product: length differs");
  if (y->rows * y->cols != A->rows) return api_fail(QSV_E_INVALID, "matrix-vector-product: output length differs");
  if (y->data == x->data) return api_fail(QSV_E_INVALID, "matrix-vector-product: y aliases x");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const unsigned rows_per_block = kLaBlock / 64;
  const unsigned grid = (unsigned)((A->rows + rows_per_block - 1) / rows_per_block);
  hipLaunchKernelGGL(k_gemv, dim3(grid), dim3(kLaBlock), 0, g_la_stream, (const double2*)A->data,
                     (const double2*)x->data, y->data, A->rows, A->cols);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_dot(int conj_x, const qsv_mat* x, const qsv_mat* y, double out[2])
{
  int rc = check_mat(x, "x");
  if (rc || (rc = check_mat(y, "y"))) return rc;
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  if (x->rows * x->cols != y->rows * y->cols) return api_fail(QSV_E_INVALID, "inner-product: lengths differ");
  std::lock_guard<std::mutex> lk(g_la_mu);
  return reduce(conj_x ? R_DOTC : R_DOT, x->data, y->data, x->rows * x->cols, 1, out);
}

int qsv_la_norm2(const qsv_mat* x, double* out)
{
  int rc = check_mat(x, "x");
  if (rc) return rc;
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  std::lock_guard<std::mutex> lk(g_la_mu);
  double r[2];
  if ((rc = reduce(R_NORM2, x->data, nullptr, x->rows * x->cols, 1, r))) return rc;
  *out = std::sqrt(r[0]);
  return QSV_OK;
}

int qsv_la_outer(const qsv_mat* x, const qsv_mat* y, qsv_mat* C)
{
  int rc = check_mat(x, "x");
  if (rc || (rc = check_mat(y, "y")) || (rc = check_mat(C, "C"))) return rc;
  const int64_t m = x->rows * x->cols, n = y->rows * y->cols;
  if (C->rows != m || C->cols != n) return api_fail(QSV_E_INVALID, "outer-product: C shape differs");
  std::lock_guard<std::mutex> lk(g_la_mu);
  hipLaunchKernelGGL(k_outer, dim3(grid_for((uint64_t)m * n)), dim3(kLaBlock), 0, g_la_stream,
                     (const double2*)x->data, (const double2*)y->data, C->data, m, n);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_kron(const qsv_mat* A, const qsv_mat* B, qsv_mat* C)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(B, "B")) || (rc = check_mat(C, "C"))) return rc;
  if (C->rows != A->rows * B->rows || C->cols != A->cols * B->cols)
    return api_fail(QSV_E_INVALID, "kronecker-product: C shape differs");
  std::lock_guard<std::mutex> lk(g_la_mu);
  hipLaunchKernelGGL(k_kron, dim3(grid_for((uint64_t)C->rows * C->cols)), dim3(kLaBlock), 0, g_la_stream,
                     (const double2*)A->data, (const double2*)B->data, C->data, A->rows, A->cols, B->rows, B->cols);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_transpose(int conj, const qsv_mat* A, qsv_mat* C)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(C, "C"))) return rc;
  if (C->rows != A->cols || C->cols != A->rows) return api_fail(QSV_E_INVALID, "transpose: C shape differs");
  if (C->data == A->data) return api_fail(QSV_E_INVALID, "transpose: C aliases A");
  if ((A->cols + 31) / 32 > 0x7fffffff || (A->rows + 31) / 32 > 65535)
    return api_fail(QSV_E_INVALID, "transpose: too many rows");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const dim3 grid((unsigned)((A->cols + 31) / 32), (unsigned)((A->rows + 31) / 32));
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, g_la_stream, (const double2*)A->data, C->data, A->rows,
                     A->cols, conj);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_trace(const qsv_mat* A, double out[2])
{
  int rc = check_mat(A, "A");
  if (rc) return rc;
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  if (A->rows != A->cols) return api_fail(QSV_E_INVALID, "Matrix must be square for trace");
  std::lock_guard<std::mutex> lk(g_la_mu);
  return reduce(R_TRACE, A->data, nullptr, A->rows, A->cols, out);
}

int qsv_la_deviation(int kind, const qsv_mat* A, double* out)
{
  int rc = check_mat(A, "A");
  if (rc) return rc;
  if (!out) return api_fail(QSV_E_INVALID, "null output");
  int rk;
  switch (kind) {
    case QSV_LA_DEV_HERMITIAN: rk = R_DEV_HERM; break;
    case QSV_LA_DEV_OFFDIAG:
      rk = R_DEV_OFFDIAG;
      if (A->cols < A->rows) return api_fail(QSV_E_INVALID, "diagonal?: fewer columns than rows");
      break;
    case QSV_LA_DEV_IDENTITY:
      rk = R_DEV_IDENT;
      if (A->cols != A->rows) return api_fail(QSV_E_INVALID, "identity deviation needs a square matrix");
      break;
    default: return api_fail(QSV_E_INVALID, "unknown deviation kind");
  }
  std::lock_guard<std::mutex> lk(g_la_mu);
  double r[2];
  if ((rc = reduce(rk, A->data, nullptr, A->rows, A->cols, r))) return rc;
  *out = r[0];
  return QSV_OK;
}

int qsv_la_solve(const qsv_mat* A, const qsv_mat* B, qsv_mat* X)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(X, "X")) || (B && (rc = check_mat(B, "B")))) return rc;
  const int64_t n = A->rows;
  if (A->cols != n) return api_fail(QSV_E_INVALID, B ? "Matrix must be square for solving"
                                                     : "Matrix must be square for inverse");
  const int64_t k = B ? B->cols : n;
  if (B && B->rows != n) return api_fail(QSV_E_INVALID, "Incompatible dimensions");
  if (X->rows != n || X->cols != k) return api_fail(QSV_E_INVALID, "solve: X shape differs");
  std::lock_guard<std::mutex> lk(g_la_mu);
  const int64_t w = n + k;
  double2* M = nullptr;
  double2* f = nullptr;
  if ((rc = alloc(&M, (uint64_t)n * w))) return rc;
  if ((rc = alloc(&f, (uint64_t)n))) {
    hipFree(M);
    return rc;
  }
  auto done = [&](int code) {
    hipFree(M);
    hipFree(f);
    return code;
  };
  if (hipMemsetAsync(g_la_flag, 0, 4 * sizeof(int), g_la_stream) != hipSuccess)
    return done(api_fail(QSV_E_DEVICE, "clearing the pivot flag failed"));
  hipLaunchKernelGGL(k_gj_build, dim3(grid_for((uint64_t)n * w)), dim3(kLaBlock), 0, g_la_stream,
                     (const double2*)A->data, (const double2*)(B ? B->data : A->data), M, n, k, B ? 0 : 1);
  for (int64_t col = 0; col < n; ++col) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(kLaBlock), 0, g_la_stream, (const double2*)M, n, w, col, g_la_flag,
                       g_la_piv);
    hipLaunchKernelGGL(k_gj_swap_scale, dim3(grid_for((uint64_t)w)), dim3(kLaBlock), 0, g_la_stream, M, w, col,
                       (const int*)g_la_flag, (const double2*)g_la_piv);
    hipLaunchKernelGGL(k_gj_factors, dim3(grid_for((uint64_t)n)), dim3(kLaBlock), 0, g_la_stream,
                       (const double2*)M, n, w, col, f, (const int*)g_la_flag);
    hipLaunchKernelGGL(k_gj_eliminate, dim3(grid_for((uint64_t)n * w)), dim3(kLaBlock), 0, g_la_stream, M, n, w,
                       col, (const double2*)f, (const int*)g_la_flag);
  }
  hipLaunchKernelGGL(k_gj_extract, dim3(grid_for((uint64_t)n * k)), dim3(kLaBlock), 0, g_la_stream,
                     (const double2*)M, X->data, n, k);
  if (hipGetLastError() != hipSuccess) return done(api_fail(QSV_E_DEVICE, "Gauss-Jordan launch failed"));
  int flag[2] = {0, 0};
  if (hipMemcpyAsync(flag, g_la_flag, 2 * sizeof(int), hipMemcpyDeviceToHost, g_la_stream) != hipSuccess ||
      hipStreamSynchronize(g_la_stream) != hipSuccess)
    return done(api_fail(QSV_E_DEVICE, "Gauss-Jordan failed"));
  if (flag[0]) {
    const std::string msg = "Matrix is singular (pivot magnitude below 1e-14 in column " +
                            std::to_string(flag[0] - 1) + ")";
    return done(api_fail(QSV_E_INVALID, msg.c_str()));
  }
  return done(QSV_OK);
}

int qsv_la_eig(const qsv_mat* A, double* eigenvalues, qsv_mat* V)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(V, "V"))) return rc;
  if (!eigenvalues) return api_fail(QSV_E_INVALID, "null eigenvalue buffer");
  const int64_t n = A->rows;
  if (A->cols != n) return api_fail(QSV_E_INVALID, "Matrix must be square for eigendecomposition");
  if (V->rows != n || V->cols != n) return api_fail(QSV_E_INVALID, "eigen-general: V shape differs");
  if (n < 1 || n > kJacSmall) return api_fail(QSV_E_INVALID, "eigen-general: the :hip backend takes n <= 64");
  if (V->data == A->data) return api_fail(QSV_E_INVALID, "eigen-general: V aliases A");
  std::lock_guard<std::mutex> lk(g_la_mu);
  double2* Y = nullptr;
  double* w = nullptr;
  if ((rc = alloc(&Y, (uint64_t)n * n))) return rc;
  if (hipMalloc(&w, (size_t)2 * n * sizeof(double) + sizeof(int)) != hipSuccess) {
    hipFree(Y);
    return api_fail(QSV_E_NOMEM, "eigen-general scratch allocation failed");
  }
  int* dfail = reinterpret_cast<int*>(w + 2 * n);
  int failed = 1;
  hipLaunchKernelGGL(k_eig_general, dim3(1), dim3(64), 0, g_la_stream, (const double2*)A->data, V->data, Y, w, (int)n,
                     dfail);
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(eigenvalues, w, (size_t)2 * n * sizeof(double), hipMemcpyDeviceToHost, g_la_stream) ==
                      hipSuccess &&
                  hipMemcpyAsync(&failed, dfail, sizeof(int), hipMemcpyDeviceToHost, g_la_stream) == hipSuccess &&
                  hipStreamSynchronize(g_la_stream) == hipSuccess;
  hipFree(Y);
  hipFree(w);
  if (!ok) return api_fail(QSV_E_DEVICE, "eigen-general failed on the device");
  if (failed) return api_fail(QSV_E_DEVICE, "eigen-general: the QR iteration did not converge");
  return QSV_OK;
}

int qsv_la_eigh(const qsv_mat* A, double* eigenvalues, qsv_mat* V, int* sweeps_out)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(V, "V"))) return rc;
  if (!eigenvalues) return api_fail(QSV_E_INVALID, "null eigenvalue buffer");
  const int64_t n = A->rows;
  if (A->cols != n) return api_fail(QSV_E_INVALID, "Matrix must be square for eigendecomposition");
  if (V->rows != n || V->cols != n) return api_fail(QSV_E_INVALID, "eigen-hermitian: V shape differs");
  if (n > 16384) return api_fail(QSV_E_INVALID, "eigen-hermitian: matrix too large");
  if (V->data == A->data) return api_fail(QSV_E_INVALID, "eigen-hermitian: V aliases A");
  std::lock_guard<std::mutex> lk(g_la_mu);
  if (n <= kJacSmall) {
    double* w = nullptr;
    if (hipMalloc(&w, (size_t)n * sizeof(double) + sizeof(int)) != hipSuccess)
      return api_fail(QSV_E_NOMEM, "eigen-hermitian scratch allocation failed");
    int* dsw = reinterpret_cast<int*>(w + n);
    int sweeps = 0;
    hipLaunchKernelGGL(k_jacobi_small, dim3(1), dim3(kLaBlock), 0, g_la_stream, (const double2*)A->data, V->data, w,
                       (int)n, dsw);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(eigenvalues, w, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, g_la_stream) ==
                        hipSuccess &&
                    hipMemcpyAsync(&sweeps, dsw, sizeof(int), hipMemcpyDeviceToHost, g_la_stream) == hipSuccess &&
                    hipStreamSynchronize(g_la_stream) == hipSuccess;
    hipFree(w);
    if (!ok) return api_fail(QSV_E_DEVICE, "eigen-hermitian (single workgroup) failed");
    if (sweeps >= 40) return api_fail(QSV_E_DEVICE, "eigen-hermitian: Jacobi did not converge");
    if (sweeps_out) *sweeps_out = sweeps;
    return QSV_OK;
  }
  // round-robin schedule: N (even) players, N - 1 rounds of N / 2 disjoint pairs
  const int64_t N = n + (n & 1);
  const int npairs = (int)(N / 2), nrounds = (int)std::max<int64_t>(1, N - 1);
  std::vector<int2> sched((size_t)npairs * nrounds);
  std::vector<int> ring(N);
  for (int64_t i = 0; i < N; ++i) ring[i] = (int)i;
  for (int r = 0; r < nrounds; ++r) {
    for (int k = 0; k < npairs; ++k) {
      int p = ring[k], q = ring[N - 1 - k];
      if (p > q) std::swap(p, q);
      sched[(size_t)r * npairs + k] = make_int2(p, q);
    }
    // keep ring[0], rotate the rest by one
    const int last = ring[N - 1];
    for (int64_t i = N - 1; i > 1; --i) ring[i] = ring[i - 1];
    if (N > 1) ring[1] = last;
  }
  double2* M = nullptr;
  int2* dpairs = nullptr;
  Rot* rot = nullptr;
  double* w = nullptr;
  if ((rc = alloc(&M, (uint64_t)n * n))) return rc;
  auto done = [&](int code) {
    hipFree(M);
    if (dpairs) hipFree(dpairs);
    if (rot) hipFree(rot);
    if (w) hipFree(w);
    return code;
  };
  if (hipMalloc(&dpairs, sched.size() * sizeof(int2)) != hipSuccess ||
      hipMalloc(&rot, (size_t)npairs * sizeof(Rot)) != hipSuccess || hipMalloc(&w, (size_t)n * sizeof(double)) != hipSuccess)
    return done(api_fail(QSV_E_NOMEM, "eigen-hermitian scratch allocation failed"));
  if (hipMemcpyAsync(dpairs, sched.data(), sched.size() * sizeof(int2), hipMemcpyHostToDevice, g_la_stream) != hipSuccess ||
      hipMemcpyAsync(M, A->data, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToDevice, g_la_stream) != hipSuccess)
    return done(api_fail(QSV_E_DEVICE, "eigen-hermitian setup copy failed"));
  hipLaunchKernelGGL(k_identity, dim3(grid_for((uint64_t)n * n)), dim3(kLaBlock), 0, g_la_stream, V->data, n);
  int sweeps = 0;
  const int kMaxSweeps = 40;
  double r[2], prev_off = -1.0;
  if (n > 1) {
    for (; sweeps < kMaxSweeps; ++sweeps) {
      if ((rc = reduce(R_OFFNORM, M, nullptr, n, n, r))) return done(rc);
      // stop when the off-diagonal mass is negligible (1e-15 of the norm) or a sweep no longer reduces it
      if (!(r[0] > 1e-30 * r[1]) || r[0] == 0.0 || (prev_off >= 0.0 && !(r[0] < prev_off))) break;
      prev_off = r[0];
      for (int rr = 0; rr < nrounds; ++rr) {
        const int2* pr = dpairs + (size_t)rr * npairs;
        hipLaunchKernelGGL(k_jac_params, dim3((npairs + kLaBlock - 1) / kLaBlock), dim3(kLaBlock), 0, g_la_stream,
                           (const double2*)M, n, pr, npairs, rot);
        hipLaunchKernelGGL(k_jac_rows, dim3(grid_for((uint64_t)npairs * n)), dim3(kLaBlock), 0, g_la_stream, M, n,
                           pr, npairs, (const Rot*)rot);
        hipLaunchKernelGGL(k_jac_cols, dim3(grid_for((uint64_t)npairs * n)), dim3(kLaBlock), 0, g_la_stream, M,
                           V->data, n, pr, npairs, (const Rot*)rot);
      }
      if (hipGetLastError() != hipSuccess) return done(api_fail(QSV_E_DEVICE, "Jacobi launch failed"));
    }
    if (sweeps == kMaxSweeps) return done(api_fail(QSV_E_DEVICE, "eigen-hermitian: Jacobi did not converge"));
  }
  hipLaunchKernelGGL(k_diag_real, dim3(grid_for((uint64_t)n)), dim3(kLaBlock), 0, g_la_stream, (const double2*)M,
                     n, w);
  if (hipMemcpyAsync(eigenvalues, w, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, g_la_stream) != hipSuccess ||
      hipStreamSynchronize(g_la_stream) != hipSuccess)
    return done(api_fail(QSV_E_DEVICE, "eigen-hermitian readback failed"));
  if (sweeps_out) *sweeps_out = sweeps;
  return done(QSV_OK);
}

int qsv_la_kron_apply(const double* const* factors, int nfactors, qsv_mat* x)
{
  int rc = check_mat(x, "x");
  if (rc) return rc;
  if (nfactors < 1 || nfactors > 40 || !factors) return api_fail(QSV_E_INVALID, "kron-apply: bad factor list");
  const uint64_t S = (uint64_t)x->rows * x->cols;
  if (S != (1ull << nfactors)) return api_fail(QSV_E_INVALID, "kron-apply: vector length is not 2^factors");
  std::lock_guard<std::mutex> lk(g_la_mu);
  for (int k = 0; k < nfactors; ++k) {
    const double* g = factors[k];
    if (!g) continue;   // identity factor
    const int b = nfactors - 1 - k;   // factor k acts on reference qubit k = index bit n-1-k
    hipLaunchKernelGGL(k_apply2, dim3(grid_for(S / 2)), dim3(kLaBlock), 0, g_la_stream, x->data, S / 2, b,
                       make_double2(g[0], g[1]), make_double2(g[2], g[3]), make_double2(g[4], g[5]),
                       make_double2(g[6], g[7]));
  }
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

int qsv_la_cholesky(const qsv_mat* A, double eps, qsv_mat* L)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(L, "L"))) return rc;
  const int64_t n = A->rows;
  if (A->cols != n) return api_fail(QSV_E_INVALID, "Cholesky decomposition requires square matrix");
  if (L->rows != n || L->cols != n) return api_fail(QSV_E_INVALID, "cholesky-decomposition: L shape differs");
  if (L->data == A->data) return api_fail(QSV_E_INVALID, "cholesky-decomposition: L aliases A");
  if (n > 8192) return api_fail(QSV_E_INVALID, "cholesky-decomposition: matrix too large");
  std::lock_guard<std::mutex> lk(g_la_mu);
  int st[2] = {0, 0};
  LACHK(hipMemsetAsync(g_la_flag, 0, 2 * sizeof(int), g_la_stream));
  hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(kDecBlock), 0, g_la_stream, (const double2*)A->data, L->data, (int)n,
                     eps, g_la_flag);
  LACHK(hipGetLastError());
  LACHK(hipMemcpyAsync(st, g_la_flag, 2 * sizeof(int), hipMemcpyDeviceToHost, g_la_stream));
  LACHK(hipStreamSynchronize(g_la_stream));
  if (st[0] == 1)
    return api_fail(QSV_E_INVALID, ("Hermitian matrix diagonal must be real (i = " + std::to_string(st[1]) + ")").c_str());
  if (st[0] == 2)
    return api_fail(QSV_E_INVALID, ("Matrix is not positive definite (i = " + std::to_string(st[1]) + ")").c_str());
  return QSV_OK;
}

int qsv_la_lu(const qsv_mat* A, qsv_mat* L, qsv_mat* U, int64_t* perm)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(L, "L")) || (rc = check_mat(U, "U"))) return rc;
  const int64_t n = A->rows;
  if (A->cols != n) return api_fail(QSV_E_INVALID, "LU decomposition requires square matrix");
  if (L->rows != n || L->cols != n || U->rows != n || U->cols != n)
    return api_fail(QSV_E_INVALID, "lu-decomposition: L or U shape differs");
  if (!perm) return api_fail(QSV_E_INVALID, "null permutation buffer");
  if (n > 8192) return api_fail(QSV_E_INVALID, "lu-decomposition: matrix too large");
  std::lock_guard<std::mutex> lk(g_la_mu);
  double2* M = nullptr;
  int* dperm = nullptr;
  if (hipMalloc(&M, (size_t)n * n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&dperm, (size_t)std::max<int64_t>(n, 1) * sizeof(int)) != hipSuccess) {
    hipFree(M);
    return api_fail(QSV_E_NOMEM, "lu-decomposition scratch allocation failed");
  }
  std::vector<int> pv((size_t)n);
  int st[2] = {0, 0};
  bool ok = hipMemcpyAsync(M, A->data, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToDevice, g_la_stream) ==
                hipSuccess &&
            hipMemsetAsync(g_la_flag, 0, 2 * sizeof(int), g_la_stream) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_lu, dim3(1), dim3(kDecBlock), 0, g_la_stream, M, (int)n, dperm, g_la_flag);
    ok = hipGetLastError() == hipSuccess;
  }
  if (ok) {
    hipLaunchKernelGGL(k_lu_split, dim3(grid_for((uint64_t)n * n)), dim3(kLaBlock), 0, g_la_stream,
                       (const double2*)M, L->data, U->data, n);
    ok = hipGetLastError() == hipSuccess;
  }
  ok = ok && hipMemcpyAsync(st, g_la_flag, 2 * sizeof(int), hipMemcpyDeviceToHost, g_la_stream) == hipSuccess &&
       (n == 0 ||
        hipMemcpyAsync(pv.data(), dperm, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, g_la_stream) == hipSuccess);
  ok = hipStreamSynchronize(g_la_stream) == hipSuccess && ok;
  hipFree(M);
  hipFree(dperm);
  if (!ok) return api_fail(QSV_E_DEVICE, "lu-decomposition failed");
  if (st[0]) return api_fail(QSV_E_INVALID, ("Matrix is singular (column " + std::to_string(st[1]) + ")").c_str());
  // row k of P·A is row perm[k] of A: apply the swaps in order to the identity order
  std::vector<int64_t> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  for (int64_t k = 0; k < n; ++k) std::swap(order[k], order[pv[k]]);
  for (int64_t k = 0; k < n; ++k) perm[k] = order[k];
  return QSV_OK;
}

int qsv_la_qr(const qsv_mat* A, qsv_mat* Q, qsv_mat* R)
{
  int rc = check_mat(A, "A");
  if (rc || (rc = check_mat(Q, "Q")) || (rc = check_mat(R, "R"))) return rc;
  const int64_t m = A->rows, n = A->cols;
  if (Q->rows != m || Q->cols != n || R->rows != n || R->cols != n)
    return api_fail(QSV_E_INVALID, "qr-decomposition: Q or R shape differs");
  if (Q->data == A->data || R->data == A->data) return api_fail(QSV_E_INVALID, "qr-decomposition: output aliases A");
  if (m > (1 << 20) || n > 8192) return api_fail(QSV_E_INVALID, "qr-decomposition: matrix too large");
  std::lock_guard<std::mutex> lk(g_la_mu);
  hipLaunchKernelGGL(k_qr_mgs, dim3(1), dim3(kDecBlock), 0, g_la_stream, (const double2*)A->data, Q->data, R->data,
                     (int)m, (int)n);
  LACHK(hipGetLastError());
  LACHK(hipStreamSynchronize(g_la_stream));
  return QSV_OK;
}

}  // extern "C"
