// FP64 residual check ||C - A·B||_F^2 and ||A·B||_F^2 on the GPU.
//
// This is the only FP64 arithmetic of the library: the product A·B of the
// (FP32 or BF16) operands is formed exactly enough in FP64 (products of FP32
// values are exact in FP64; sums carry 53 bits) and compared with an FP32
// result C, so the relative Frobenius error of any precision mode can be
// measured on the device at full block-stream sizes.
//
// 64x64 output tile per 256-lane workgroup, 4x4 outputs per lane, K step 16
// staged through LDS as FP64; the two sums are reduced per workgroup and
// added atomically into out2 (order-dependent only in the last bits).
#include "common.h"

namespace bbts_hip {
namespace {

constexpr int TM = 64, TN = 64, TK = 16;

__device__ inline double ld(const void *p, int64_t idx, int bf) {
  if (bf) {
    uint16_t h = static_cast<const uint16_t *>(p)[idx];
    return (double)__uint_as_float(((uint32_t)h) << 16);
  }
  return (double)static_cast<const float *>(p)[idx];
}

__global__ __launch_bounds__(256) void residual_kernel(int M, int N, int K, const void *A, int64_t lda, int abf,
                                                       const void *B, int64_t ldb, int bbf, const float *C,
                                                       int64_t ldc, double *out2) {
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  __shared__ double red[2][256];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int e = threadIdx.x; e < TM * TK; e += 256) {
      int m = e / TK, k = e % TK;
      int gm = m0 + m, gk = k0 + k;
      As[k][m] = (gm < M && gk < K) ? ld(A, (int64_t)gm * lda + gk, abf) : 0.0;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      int k = e / TN, n = e % TN;
      int gk = k0 + k, gn = n0 + n;
      Bs[k][n] = (gk < K && gn < N) ? ld(B, (int64_t)gk * ldb + gn, bbf) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TK; ++k) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  double se = 0.0, sr = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
      if (gm < M && gn < N) {
        double r = acc[i][j];
        double d = (double)C[(int64_t)gm * ldc + gn] - r;
        se += d * d;
        sr += r * r;
      }
    }
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out2[0], red[0][0]);
    atomicAdd(&out2[1], red[1][0]);
  }
}

__global__ void zero2_kernel(double *p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0.0;
}

}  // namespace
}  // namespace bbts_hip

using namespace bbts_hip;

extern "C" int bbts_hip_residual_f64(bbts_hip_ctx *ctx, int M, int N, int K, const void *A, int64_t lda, int a_bf16,
                                     const void *B, int64_t ldb, int b_bf16, const float *C, int64_t ldc,
                                     double *out2, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  BBTS_REQUIRE(M >= 0 && N >= 0 && K >= 0 && out2, BBTS_E_INVALID, "bad residual arguments");
  BBTS_REQUIRE(M <= 65535 * 64, BBTS_E_UNSUPPORTED, "M too large for the residual grid");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, s, out2);
  BBTS_LAUNCH_CHECK();
  if (M == 0 || N == 0) return BBTS_OK;
  BBTS_REQUIRE(C && (K == 0 || (A && B)), BBTS_E_INVALID, "null operand");
  dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM));
  hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, s, M, N, K, A, lda, a_bf16, B, ldb, b_bf16, C, ldc, out2);
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}
