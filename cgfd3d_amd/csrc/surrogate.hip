// Surrogate network kernels: BF16 MFMA GEMMs with FP32 accumulation and the
// small elementwise/reduction kernels that training needs (loss gradient,
// column sums, Adam).  Every layer of the network (the MLP and the 1-D
// transposed convolutions with kernel = stride, whose channels-last layout
// makes them plain GEMMs) runs through cgfds_gemm_bf16.
//
// Storage: bf16 values are raw 16-bit words (upper half of an IEEE binary32,
// round-to-nearest-even on conversion); matrices are row-major with explicit
// leading dimensions in elements.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "cgfdh.h"
#include "cgfds.h"

int cgfdh_internal_fail(int code, const char* msg);

namespace cgfds {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static int sfail(int code, const char* what)
{
  return cgfdh_internal_fail(code, what);
}

__device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// bf16 value > 0 (the ReLU mask of a stored activation)
__device__ __forceinline__ bool bf_pos(uint16_t h) { return h != 0 && !(h & 0x8000u); }

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  long lda;
  const uint16_t* B;
  long ldb;
  const float* bias;
  int epi;                // EPI_* below
  const uint16_t* mask;   // EPI_MASK: bf16 [M][ldm]
  long ldm;
  float* C32;
  long ldc32;
  uint16_t* C16;
  long ldc16;
  int kchunk;             // K range per blockIdx.z, multiple of 32
  float* work;            // split-K partials [gridDim.z][M][N] (then no epilogue)
};

enum { EPI_NONE = 0, EPI_RELU = 1, EPI_MASK = 2 };

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;   // LDS rows padded by 16 B

// C[M][N] = sum_k opA[m][k] * opB[k][n].  opA = A ([M][K]) or, TA, A^T with A
// stored [K][M]; opB = B ([K][N]) or, TB, B^T with B stored [N][K].
// 256 threads = 4 waves in a 2x2 arrangement, each wave a 32x32 sub-tile of
// 2x2 MFMA 16x16x32 tiles.  Tiles go through LDS as [row][k] with k
// contiguous, double buffered with one barrier per K step.
template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g)
{
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN][LDK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * g.kchunk;
  const int ke = min(g.K, kb + g.kchunk);
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;

  // global -> register staging: each thread moves 8 bf16 (16 B) of A and of B
  const int a_r = TA ? (t >> 3) : (t >> 2);          // k row (TA) or m row
  const int a_c = TA ? (t & 7) * 8 : (t & 3) * 8;    // m col (TA) or k col
  const int b_r = TB ? (t >> 2) : (t >> 3);          // n row (TB) or k row
  const int b_c = TB ? (t & 3) * 8 : (t & 7) * 8;    // k col (TB) or n col
  uint4 ra, rb;
  auto gload = [&](int k0) {
    const uint4 z = {0u, 0u, 0u, 0u};
    if (TA) {
      const int k = k0 + a_r, m = m0 + a_c;
      ra = (k < ke && m < g.M) ? *(const uint4*)(g.A + (long)k * g.lda + m) : z;
    } else {
      const int m = m0 + a_r, k = k0 + a_c;
      ra = (k < ke && m < g.M) ? *(const uint4*)(g.A + (long)m * g.lda + k) : z;
    }
    if (TB) {
      const int n = n0 + b_r, k = k0 + b_c;
      rb = (k < ke && n < g.N) ? *(const uint4*)(g.B + (long)n * g.ldb + k) : z;
    } else {
      const int k = k0 + b_r, n = n0 + b_c;
      rb = (k < ke && n < g.N) ? *(const uint4*)(g.B + (long)k * g.ldb + n) : z;
    }
  };
  auto lstore = [&](int buf) {
    if (TA) {
      const uint16_t* e = (const uint16_t*)&ra;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[buf][a_c + q][a_r] = e[q];
    } else {
      *(uint4*)&As[buf][a_r][a_c] = ra;
    }
    if (TB) {
      *(uint4*)&Bs[buf][b_r][b_c] = rb;
    } else {
      const uint16_t* e = (const uint16_t*)&rb;
#pragma unroll
      for (int q = 0; q < 8; ++q) Bs[buf][b_c + q][b_r] = e[q];
    }
  };

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};

  // lane l feeds row/col l%16 of its 16x16 tile with k = 8*(l/16) .. +7
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  int buf = 0;
  if (kb < ke) {
    gload(kb);
    lstore(0);
  }
  __syncthreads();
  for (int k0 = kb; k0 < ke; k0 += BK) {
    const bool more = k0 + BK < ke;
    if (more) gload(k0 + BK);
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8*)&As[buf][wm + 16 * i + lr][lk];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8*)&Bs[buf][wn + 16 * j + lr][lk];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // D layout of a 16x16 MFMA tile: lane l holds rows 4*(l/16)+e, column l%16
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm + 16 * i + 4 * (lane >> 4) + e;
        const int c = n0 + wn + 16 * j + lr;
        if (r >= g.M || c >= g.N) continue;
        float v = acc[i][j][e];
        if (g.work) {
          g.work[((long)blockIdx.z * g.M + r) * g.N + c] = v;
          continue;
        }
        if (g.bias) v = v + g.bias[c];
        if (g.epi == EPI_RELU) v = v > 0.f ? v : 0.f;
        if (g.epi == EPI_MASK) v = bf_pos(g.mask[(long)r * g.ldm + c]) ? v : 0.f;
        if (g.C32) g.C32[(long)r * g.ldc32 + c] = v;
        if (g.C16) g.C16[(long)r * g.ldc16 + c] = f2bf(v);
      }
}

// out[r][c] = sum over s of work[s][r][c], s in increasing order (deterministic)
__global__ void splitk_reduce_kernel(int S, int M, int N, const float* work, float* out, long ldo)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const long r = i / N, c = i % N;
  float s = 0.f;
  for (int z = 0; z < S; ++z) s = s + work[(long)z * M * N + i];
  out[r * ldo + c] = s;
}

// partial column sums of a bf16 matrix: block (x, y) sums rows of split y for
// 64 columns; 4 row lanes per column are combined in a fixed order
__global__ __launch_bounds__(256) void colsum_bf16_kernel(int M, int N, const uint16_t* G,
                                                         long ldg, int S, float* work)
{
  __shared__ float part[4][64];
  const int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cx;
  const long r0 = (long)M * blockIdx.y / S, r1 = (long)M * (blockIdx.y + 1) / S;
  float s = 0.f;
  if (c < N)
    for (long r = r0 + ry; r < r1; r += 4) s = s + bf2f(G[r * ldg + c]);
  part[ry][cx] = s;
  __syncthreads();
  if (ry == 0 && c < N)
    work[(long)blockIdx.y * N + c] = ((part[0][cx] + part[1][cx]) + part[2][cx]) + part[3][cx];
}

// per-sample relative squared misfit and its gradient, one block per sample:
// loss[b] = |Y_b - T_b|^2 / (tn2[b] + eps);  G_b = bf16(2 (Y_b - T_b) / (nsamp (tn2[b] + eps)))
__global__ __launch_bounds__(256) void loss_grad_kernel(long n, const float* Y, const float* T,
                                                       const float* tn2, float eps, int nsamp,
                                                       uint16_t* G, float* loss)
{
  __shared__ float red[256];
  const int b = blockIdx.x;
  const float den = tn2[b] + eps;
  const float scale = 2.0f / ((float)nsamp * den);
  const long o = (long)b * n;
  float s = 0.f;
  for (long i = threadIdx.x; i < n; i += 256) {
    const float d = Y[o + i] - T[o + i];
    s = s + d * d;
    if (G) G[o + i] = f2bf(d * scale);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[b] = red[0] / den;
}

// Adam on FP32 master values; w16 (optional) receives the bf16 copy
__global__ void adam_kernel(long n, float* w, float* m, float* v, const float* g, uint16_t* w16,
                            float lr, float b1, float b2, float c1, float c2, float bc1, float bc2,
                            float eps)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + c1 * gi;
  const float vi = b2 * v[i] + c2 * (gi * gi);
  const float mh = mi / bc1, vh = vi / bc2;
  const float wi = w[i] - lr * (mh / (sqrtf(vh) + eps));
  m[i] = mi;
  v[i] = vi;
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

// moment-tensor contraction of basis waveforms, one block per sample:
// y[b][t][c] = sum_k m[b][k] g[b][t][k][c] (k in increasing order)
constexpr int MT_MAX = 8;

__global__ __launch_bounds__(256) void mt_contract_kernel(int nt, int nc, int nk, const float* m,
                                                         const float* g, float* y)
{
  const int b = blockIdx.x;
  float mk[MT_MAX];
  for (int k = 0; k < nk; ++k) mk[k] = m[b * nk + k];
  const long n = (long)nt * nc;
  const float* gb = g + (long)b * n * nk;
  for (long e = threadIdx.x; e < n; e += 256) {
    const long t = e / nc, c = e - t * nc;
    const float* q = gb + t * nk * nc + c;
    float v = mk[0] * q[0];
    for (int k = 1; k < nk; ++k) v = v + mk[k] * q[(long)k * nc];
    y[(long)b * n + e] = v;
  }
}

// the contraction fused with the per-sample relative misfit and its gradient
// with respect to the basis waveforms: dg[b][t][k][c] = bf16(m[b][k] dy[b][t][c]),
// dy = 2 (y - T) / (nsamp (tn2[b] + eps))
__global__ __launch_bounds__(256) void mt_loss_grad_kernel(int nt, int nc, int nk, const float* m,
                                                          const float* g, const float* T,
                                                          const float* tn2, float eps, int nsamp,
                                                          uint16_t* dg, float* loss)
{
  __shared__ float red[256];
  const int b = blockIdx.x;
  float mk[MT_MAX];
  for (int k = 0; k < nk; ++k) mk[k] = m[b * nk + k];
  const float den = tn2[b] + eps;
  const float scale = 2.0f / ((float)nsamp * den);
  const long n = (long)nt * nc;
  const float* gb = g + (long)b * n * nk;
  uint16_t* db = dg ? dg + (long)b * n * nk : nullptr;
  float s = 0.f;
  for (long e = threadIdx.x; e < n; e += 256) {
    const long t = e / nc, c = e - t * nc;
    const long o = t * nk * nc + c;
    float v = mk[0] * gb[o];
    for (int k = 1; k < nk; ++k) v = v + mk[k] * gb[o + (long)k * nc];
    const float d = v - T[(long)b * n + e];
    s = s + d * d;
    if (db) {
      const float dy = d * scale;
      for (int k = 0; k < nk; ++k) db[o + (long)k * nc] = f2bf(mk[k] * dy);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[b] = red[0] / den;
}

__global__ void to_bf16_kernel(long n, const float* src, uint16_t* dst)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

static bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

static int launch_check(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return 0;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: launch failed: %s", what, hipGetErrorString(e));
  return sfail(CGFDH_EDEVICE, buf);
}

}  // namespace cgfds

using namespace cgfds;

extern "C" int cgfds_gemm_bf16(int M, int N, int K, const void* A, long lda, int transA,
                               const void* B, long ldb, int transB, const float* bias,
                               int epilogue, const void* mask, long ldm, float* C32, long ldc32,
                               void* C16, long ldc16, int splitk, float* work, void* stream)
{
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || (!C32 && !C16))
    return sfail(CGFDH_EINVAL, "cgfds_gemm_bf16: empty shape or null operand");
  if (M % 8 || N % 8 || K % 8 || lda % 8 || ldb % 8 || !al16(A) || !al16(B))
    return sfail(CGFDH_EINVAL,
                 "cgfds_gemm_bf16: M, N, K and the A/B leading dimensions must be multiples "
                 "of 8 and A, B 16-byte aligned");
  if (lda < (transA ? M : K) || ldb < (transB ? K : N) || (C32 && ldc32 < N) ||
      (C16 && ldc16 < N))
    return sfail(CGFDH_EINVAL, "cgfds_gemm_bf16: leading dimension below the row length");
  if (epilogue < EPI_NONE || epilogue > EPI_MASK || (epilogue == EPI_MASK && (!mask || ldm < N)))
    return sfail(CGFDH_EINVAL, "cgfds_gemm_bf16: bad epilogue or missing mask");
  if (splitk < 1 || splitk > 1024)
    return sfail(CGFDH_EINVAL, "cgfds_gemm_bf16: splitk outside 1..1024");
  int kchunk = ((K + splitk - 1) / splitk + BK - 1) / BK * BK;
  int S = (K + kchunk - 1) / kchunk;
  if (S > 1 && (!work || !C32 || C16 || bias || epilogue != EPI_NONE))
    return sfail(CGFDH_EINVAL,
                 "cgfds_gemm_bf16: split-K needs a work buffer, FP32 output and no epilogue");
  GemmArgs g{M, N, K, (const uint16_t*)A, lda, (const uint16_t*)B, ldb, bias, epilogue,
             (const uint16_t*)mask, ldm, C32, ldc32, (uint16_t*)C16, ldc16, kchunk,
             S > 1 ? work : nullptr};
  hipStream_t st = (hipStream_t)stream;
  // row tiles on x (up to 2^31 - 1 blocks), column tiles on y
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, S);
  if (grid.y > 65535)
    return sfail(CGFDH_EINVAL, "cgfds_gemm_bf16: N above 65535*64 columns");
  if (!transA && !transB) hipLaunchKernelGGL((gemm_bf16_kernel<0, 0>), grid, dim3(256), 0, st, g);
  if (!transA && transB) hipLaunchKernelGGL((gemm_bf16_kernel<0, 1>), grid, dim3(256), 0, st, g);
  if (transA && !transB) hipLaunchKernelGGL((gemm_bf16_kernel<1, 0>), grid, dim3(256), 0, st, g);
  if (transA && transB) hipLaunchKernelGGL((gemm_bf16_kernel<1, 1>), grid, dim3(256), 0, st, g);
  if (int rc = launch_check("cgfds_gemm_bf16")) return rc;
  if (S > 1) {
    long n = (long)M * N;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       S, M, N, (const float*)work, C32, ldc32);
    return launch_check("cgfds_gemm_bf16 (split-K reduce)");
  }
  return 0;
}

extern "C" int cgfds_colsum_bf16(int M, int N, const void* G, long ldg, float* out, int splits,
                                 float* work, void* stream)
{
  if (M <= 0 || N <= 0 || !G || !out || !work || ldg < N || splits < 1 || splits > 65535)
    return sfail(CGFDH_EINVAL, "cgfds_colsum_bf16: bad argument");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(colsum_bf16_kernel, dim3((N + 63) / 64, splits), dim3(256), 0, st, M, N,
                     (const uint16_t*)G, ldg, splits, work);
  if (int rc = launch_check("cgfds_colsum_bf16")) return rc;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, st, splits, 1, N,
                     (const float*)work, out, (long)N);
  return launch_check("cgfds_colsum_bf16 (reduce)");
}

extern "C" int cgfds_loss_grad(int nsamp, long n, const float* Y, const float* T,
                               const float* tnorm2, float eps, void* G16, float* loss,
                               void* stream)
{
  if (nsamp <= 0 || n <= 0 || !Y || !T || !tnorm2 || !loss)
    return sfail(CGFDH_EINVAL, "cgfds_loss_grad: bad argument");
  hipLaunchKernelGGL(loss_grad_kernel, dim3(nsamp), dim3(256), 0, (hipStream_t)stream, n, Y, T,
                     tnorm2, eps, nsamp, (uint16_t*)G16, loss);
  return launch_check("cgfds_loss_grad");
}

extern "C" int cgfds_adam(long n, float* w, float* m, float* v, const float* g, void* w16,
                          float lr, float b1, float b2, float c1, float c2, float bc1, float bc2,
                          float eps, void* stream)
{
  if (n <= 0 || !w || !m || !v || !g) return sfail(CGFDH_EINVAL, "cgfds_adam: bad argument");
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, w, m, v, g, (uint16_t*)w16, lr, b1, b2, c1, c2, bc1,
                     bc2, eps);
  return launch_check("cgfds_adam");
}

extern "C" int cgfds_mt_contract(int nsamp, int nt, int nc, int nk, const float* m,
                                 const float* g, float* y, void* stream)
{
  if (nsamp <= 0 || nt <= 0 || nc <= 0 || nk < 1 || nk > MT_MAX || !m || !g || !y)
    return sfail(CGFDH_EINVAL, "cgfds_mt_contract: bad argument (nk must be 1..8)");
  hipLaunchKernelGGL(mt_contract_kernel, dim3(nsamp), dim3(256), 0, (hipStream_t)stream, nt, nc,
                     nk, m, g, y);
  return launch_check("cgfds_mt_contract");
}

extern "C" int cgfds_mt_loss_grad(int nsamp, int nt, int nc, int nk, const float* m,
                                  const float* g, const float* T, const float* tnorm2, float eps,
                                  void* dg16, float* loss, void* stream)
{
  if (nsamp <= 0 || nt <= 0 || nc <= 0 || nk < 1 || nk > MT_MAX || !m || !g || !T || !tnorm2 ||
      !loss)
    return sfail(CGFDH_EINVAL, "cgfds_mt_loss_grad: bad argument (nk must be 1..8)");
  hipLaunchKernelGGL(mt_loss_grad_kernel, dim3(nsamp), dim3(256), 0, (hipStream_t)stream, nt, nc,
                     nk, m, g, T, tnorm2, eps, nsamp, (uint16_t*)dg16, loss);
  return launch_check("cgfds_mt_loss_grad");
}

extern "C" int cgfds_f32_to_bf16(long n, const float* src, void* dst, void* stream)
{
  if (n <= 0 || !src || !dst) return sfail(CGFDH_EINVAL, "cgfds_f32_to_bf16: bad argument");
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, src, (uint16_t*)dst);
  return launch_check("cgfds_f32_to_bf16");
}
