// FP32 block GEMM on gfx950 MFMA (v_mfma_f32_32x32x2_f32): the drop-in for
// the reference's matrix_mult (dense_matrix_mult.cc:50-86 CPU semantics,
// replacing dense_matrix_gpu_mult.cu's cublasSgemm).
//
// Workgroup: 256 lanes = 4 wave64s in a 2x2 grid; macro tile 128x128, K step
// 16; each wave owns a 64x64 sub-tile (2x2 MFMA tiles, 64 FP32 accumulators
// per lane).  Operands are staged global -> registers -> LDS (k-major, row
// stride 160 floats so the two half-waves of an MFMA operand read land in
// disjoint banks) with two LDS buffers: the next tile's global loads are in
// flight while the current tile feeds the MFMAs; one barrier per K step.
// Tile order is XCD-aware: consecutive tiles of a launch are dispatched
// round-robin over the 8 XCDs, so the linear id is remapped to give each
// XCD a contiguous, M-grouped band of output tiles that shares A rows and B
// columns in that XCD's L2.
#include "common.h"

namespace bbts_hip {

namespace {

typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int BK = 16;
constexpr int kThreads = 256;
constexpr int kGroupM = 8;

__device__ inline void map_tile(int lin, int tiles_m, int tiles_n, int *tm, int *tn) {
  int total = tiles_m * tiles_n;
  if (total % 8 == 0) {
    int per = total / 8;
    lin = (lin % 8) * per + lin / 8;
  }
  int width = kGroupM * tiles_n;
  int group = lin / width;
  int first_m = group * kGroupM;
  int gsize = tiles_m - first_m < kGroupM ? tiles_m - first_m : kGroupM;
  int in_group = lin % width;
  *tm = first_m + in_group % gsize;
  *tn = in_group / gsize;
}

// TB x TB output tile (TB = 128: 64x64 per wave; TB = 64: 32x32 per wave, for blocks too
// small to fill 256 CUs with 128x128 tiles).  LDS rows are TB + 32 floats, which keeps the
// two half-waves of an MFMA operand read in disjoint banks.
// AT: A stored K x M (element (m,k) at A[k*lda + m]); else A[m*lda + k].
// BT: B stored N x K (element (k,n) at B[n*ldb + k]); else B[k*ldb + n].
template <int TB, bool AT, bool BT>
__global__ __launch_bounds__(kThreads, 2) void sgemm_kernel(gemm_args g, int vecA, int vecB, int tiles_m,
                                                            int tiles_n) {
  constexpr int STRIDE = TB + 32;
  constexpr int W = TB / 2;        // per-wave tile
  constexpr int F = W / 32;        // 32x32 MFMA tiles per wave side
  constexpr int PER = TB * BK / 4 / kThreads;  // float4 per lane per operand and stage
  constexpr int RV = TB / 4;       // float4 per MN-contiguous row
  __shared__ float As[2][BK][STRIDE];
  __shared__ float Bs[2][BK][STRIDE];

  if (g.meta_dst != nullptr && blockIdx.x == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
    g.meta_dst[0] = g.w0;
    g.meta_dst[1] = g.w1;
  }

  int tm, tn;
  map_tile(blockIdx.x, tiles_m, tiles_n, &tm, &tn);
  const int bz = blockIdx.z;
  const float *__restrict__ A = g.A + (int64_t)bz * g.sa;
  const float *__restrict__ B = g.B + (int64_t)bz * g.sb;
  float *__restrict__ C = g.C + (int64_t)bz * g.sc;
  const int M = g.M, N = g.N, K = g.K;
  const int m0 = tm * TB, n0 = tn * TB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  float ra[4 * PER], rb[4 * PER];

  auto load_a = [&](int k0) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      int q = tid + kThreads * r;
      if (!AT) {
        int m = q >> 2, kq = (q & 3) * 4;
        int gm = m0 + m, gk = k0 + kq;
        const float *p = A + (int64_t)gm * g.lda + gk;
        if (vecA && gm < M && gk + 3 < K) {
          float4 v = *reinterpret_cast<const float4 *>(p);
          ra[4 * r + 0] = v.x; ra[4 * r + 1] = v.y; ra[4 * r + 2] = v.z; ra[4 * r + 3] = v.w;
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) ra[4 * r + c] = (gm < M && gk + c < K) ? p[c] : 0.0f;
        }
      } else {
        int k = q / RV, mq = (q % RV) * 4;
        int gm = m0 + mq, gk = k0 + k;
        const float *p = A + (int64_t)gk * g.lda + gm;
        if (vecA && gk < K && gm + 3 < M) {
          float4 v = *reinterpret_cast<const float4 *>(p);
          ra[4 * r + 0] = v.x; ra[4 * r + 1] = v.y; ra[4 * r + 2] = v.z; ra[4 * r + 3] = v.w;
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) ra[4 * r + c] = (gk < K && gm + c < M) ? p[c] : 0.0f;
        }
      }
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      int q = tid + kThreads * r;
      if (!BT) {
        int k = q / RV, nq = (q % RV) * 4;
        int gk = k0 + k, gn = n0 + nq;
        const float *p = B + (int64_t)gk * g.ldb + gn;
        if (vecB && gk < K && gn + 3 < N) {
          float4 v = *reinterpret_cast<const float4 *>(p);
          rb[4 * r + 0] = v.x; rb[4 * r + 1] = v.y; rb[4 * r + 2] = v.z; rb[4 * r + 3] = v.w;
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) rb[4 * r + c] = (gk < K && gn + c < N) ? p[c] : 0.0f;
        }
      } else {
        int n = q >> 2, kq = (q & 3) * 4;
        int gn = n0 + n, gk = k0 + kq;
        const float *p = B + (int64_t)gn * g.ldb + gk;
        if (vecB && gn < N && gk + 3 < K) {
          float4 v = *reinterpret_cast<const float4 *>(p);
          rb[4 * r + 0] = v.x; rb[4 * r + 1] = v.y; rb[4 * r + 2] = v.z; rb[4 * r + 3] = v.w;
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) rb[4 * r + c] = (gn < N && gk + c < K) ? p[c] : 0.0f;
        }
      }
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      int q = tid + kThreads * r;
      if (!AT) {
        int m = q >> 2, kq = (q & 3) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) As[buf][kq + c][m] = ra[4 * r + c];
      } else {
        int k = q / RV, mq = (q % RV) * 4;
        *reinterpret_cast<float4 *>(&As[buf][k][mq]) =
            float4{ra[4 * r + 0], ra[4 * r + 1], ra[4 * r + 2], ra[4 * r + 3]};
      }
      if (!BT) {
        int k = q / RV, nq = (q % RV) * 4;
        *reinterpret_cast<float4 *>(&Bs[buf][k][nq]) =
            float4{rb[4 * r + 0], rb[4 * r + 1], rb[4 * r + 2], rb[4 * r + 3]};
      } else {
        int n = q >> 2, kq = (q & 3) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) Bs[buf][kq + c][n] = rb[4 * r + c];
      }
    }
  };

  f16v acc[F][F];
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load_a(0);
    load_b(0);
    store_lds(0);
  }
  __syncthreads();

  const int half = lane >> 5, col = lane & 31;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) {
      load_a((kt + 1) * BK);
      load_b((kt + 1) * BK);
    }
#pragma unroll
    for (int kk = 0; kk < BK / 2; ++kk) {
      const int k = kk * 2 + half;
      float a[F], b[F];
#pragma unroll
      for (int i = 0; i < F; ++i) a[i] = As[buf][k][wm * W + i * 32 + col];
#pragma unroll
      for (int j = 0; j < F; ++j) b[j] = Bs[buf][k][wn * W + j * 32 + col];
#pragma unroll
      for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store_lds(buf ^ 1);
    __syncthreads();
  }

  // Epilogue: accumulator v of a 32x32 tile holds row 8*(v/4) + 4*(lane/32) + v%4, column lane%32.
  const float alpha = g.alpha;
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const int gn = n0 + wn * W + j * 32 + col;
      if (gn >= N) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int gm = m0 + wm * W + i * 32 + (v >> 2) * 8 + half * 4 + (v & 3);
        if (gm < M) C[(int64_t)gm * g.ldc + gn] = alpha * acc[i][j][v];
      }
    }
}

inline bool al16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <int TB>
void launch_tiles(const gemm_args &g, dim3 grid, hipStream_t s, int vecA, int vecB, int tiles_m, int tiles_n) {
  if (!g.ta && !g.tb)
    hipLaunchKernelGGL((sgemm_kernel<TB, false, false>), grid, dim3(kThreads), 0, s, g, vecA, vecB, tiles_m, tiles_n);
  else if (g.ta && !g.tb)
    hipLaunchKernelGGL((sgemm_kernel<TB, true, false>), grid, dim3(kThreads), 0, s, g, vecA, vecB, tiles_m, tiles_n);
  else if (!g.ta && g.tb)
    hipLaunchKernelGGL((sgemm_kernel<TB, false, true>), grid, dim3(kThreads), 0, s, g, vecA, vecB, tiles_m, tiles_n);
  else
    hipLaunchKernelGGL((sgemm_kernel<TB, true, true>), grid, dim3(kThreads), 0, s, g, vecA, vecB, tiles_m, tiles_n);
}

}  // namespace

int launch_sgemm_f32(bbts_hip_ctx *ctx, const gemm_args &g, hipStream_t s) {
  BBTS_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0 && g.batch >= 1, BBTS_E_INVALID, "bad GEMM sizes");
  BBTS_REQUIRE(g.batch <= 65535, BBTS_E_INVALID, "batch %d exceeds 65535", g.batch);
  if (g.M == 0 || g.N == 0) return BBTS_OK;
  // 64x64 tiles when 128x128 tiles would leave CUs idle (fewer tiles than CUs)
  const long long t128 = (long long)((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch;
  const int cus = ctx ? ctx->num_cus : 256;
  const int TB = t128 < (long long)cus ? 64 : 128;
  int tiles_m = (g.M + TB - 1) / TB, tiles_n = (g.N + TB - 1) / TB;
  int vecA = al16(g.A) && (g.lda % 4 == 0) && (g.batch == 1 || g.sa % 4 == 0);
  int vecB = al16(g.B) && (g.ldb % 4 == 0) && (g.batch == 1 || g.sb % 4 == 0);
  dim3 grid((unsigned)(tiles_m * tiles_n), 1, (unsigned)g.batch);
  if (TB == 64)
    launch_tiles<64>(g, grid, s, vecA, vecB, tiles_m, tiles_n);
  else
    launch_tiles<128>(g, grid, s, vecA, vecB, tiles_m, tiles_n);
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}

int launch_gemm_any(bbts_hip_ctx *ctx, const gemm_args &g, int precision, hipStream_t s) {
  switch (precision) {
    case BBTS_PREC_FP32: return launch_sgemm_f32(ctx, g, s);
    case BBTS_PREC_BF16X1: return launch_gemm_split_bf16(ctx, g, 1, 1, 1, s);
    case BBTS_PREC_BF16X3: return launch_gemm_split_bf16(ctx, g, 2, 2, 2, s);
    case BBTS_PREC_BF16X6: return launch_gemm_split_bf16(ctx, g, 3, 3, 3, s);
    case BBTS_PREC_AUTO: return launch_gemm_split_bf16(ctx, g, -1, -1, 3, s);
    case BBTS_PREC_FP8X1: return launch_gemm_fp8(ctx, g, 0, s);
    case BBTS_PREC_FP8X3: return launch_gemm_fp8(ctx, g, 1, s);
    default: set_error("unknown precision mode %d", precision); return BBTS_E_INVALID;
  }
}

}  // namespace bbts_hip

using namespace bbts_hip;

extern "C" {

int bbts_hip_gemm_batched(bbts_hip_ctx *ctx, int M, int N, int K, const float *A, int64_t lda, int64_t sa, int ta,
                          const float *B, int64_t ldb, int64_t sb, int tb, float *C, int64_t ldc, int64_t sc,
                          int batch, float alpha, int precision, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  BBTS_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 1, BBTS_E_INVALID, "bad sizes M=%d N=%d K=%d batch=%d", M, N, K,
               batch);
  BBTS_REQUIRE((M == 0 || N == 0) || (C && (K == 0 || (A && B))), BBTS_E_INVALID, "null operand");
  BBTS_REQUIRE(ldc >= N, BBTS_E_INVALID, "ldc %lld < N %d", (long long)ldc, N);
  BBTS_REQUIRE(ta ? lda >= M : lda >= K, BBTS_E_INVALID, "lda %lld too small", (long long)lda);
  BBTS_REQUIRE(tb ? ldb >= K : ldb >= N, BBTS_E_INVALID, "ldb %lld too small", (long long)ldb);
  gemm_args g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.A = A; g.lda = lda; g.sa = sa; g.ta = ta ? 1 : 0;
  g.B = B; g.ldb = ldb; g.sb = sb; g.tb = tb ? 1 : 0;
  g.C = C; g.ldc = ldc; g.sc = sc;
  g.alpha = alpha;
  return launch_gemm_any(ctx, g, precision, (hipStream_t)stream);
}

int bbts_hip_gemm(bbts_hip_ctx *ctx, int M, int N, int K, const float *A, int64_t lda, int ta, const float *B,
                  int64_t ldb, int tb, float *C, int64_t ldc, float alpha, int precision, bbts_stream_t stream) {
  return bbts_hip_gemm_batched(ctx, M, N, K, A, lda, 0, ta, B, ldb, 0, tb, C, ldc, 0, 1, alpha, precision, stream);
}

// matrix_mult ud (dense_matrix_mult.cc:50-86): I = a.rows, J = b.cols, K = a.cols;
// the reference passes lda = K, ldb = J, which is only consistent for transposes of square blocks.
int bbts_hip_dense_matmul(bbts_hip_ctx *ctx, const bbts_tensor_t *a, const bbts_tensor_t *b, bbts_tensor_t *out,
                          int trans_a, int trans_b, int precision, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  block_rec ra, rb, ro;
  if ((rc = lookup(ctx, a, &ra)) || (rc = lookup(ctx, b, &rb)) || (rc = lookup(ctx, out, &ro))) return rc;
  dense_m_t ma = dense_meta(ra.meta), mb = dense_meta(rb.meta);
  const uint32_t I = ma.num_rows, J = mb.num_cols, K = ma.num_cols;
  BBTS_REQUIRE(bbts_hip_dense_size(ma.num_rows, ma.num_cols) <= ra.bytes &&
                   bbts_hip_dense_size(mb.num_rows, mb.num_cols) <= rb.bytes,
               BBTS_E_SHAPE, "input block smaller than its meta");
  // dense_matrix_mult.cc:70 asserts a.cols == b.rows (debug builds); enforced here
  BBTS_REQUIRE(ma.num_cols == mb.num_rows, BBTS_E_SHAPE, "matrix_mult: a is %ux%u, b is %ux%u", ma.num_rows,
               ma.num_cols, mb.num_rows, mb.num_cols);
  BBTS_REQUIRE(!trans_a || I == K, BBTS_E_UNSUPPORTED,
               "transposed a must be square (the reference passes lda = a.cols)");
  BBTS_REQUIRE(!trans_b || mb.num_rows == J, BBTS_E_UNSUPPORTED,
               "transposed b must be square (the reference passes ldb = b.cols)");
  BBTS_REQUIRE(bbts_hip_dense_size(I, J) <= ro.bytes, BBTS_E_SHAPE, "output block of %zu bytes cannot hold %ux%u",
               ro.bytes, I, J);
  tensor_meta_t mo = ro.meta;
  set_dense_meta(mo, I, J);
  if ((rc = update_meta(ctx, out, mo))) return rc;
  gemm_args g{};
  g.M = (int)I; g.N = (int)J; g.K = (int)K; g.batch = 1;
  g.A = payload(a); g.lda = (int64_t)K; g.ta = trans_a ? 1 : 0;
  g.B = payload(b); g.ldb = (int64_t)J; g.tb = trans_b ? 1 : 0;
  g.C = payload(out); g.ldc = (int64_t)J;
  g.alpha = 1.0f;
  g.meta_dst = meta_words(out); g.w0 = I; g.w1 = J;
  if (I == 0 || J == 0) {
    // nothing to compute; the meta words still have to land on the device
    return launch_add(nullptr, nullptr, nullptr, 0, g.meta_dst, I, J, (hipStream_t)stream);
  }
  return launch_gemm_any(ctx, g, precision, (hipStream_t)stream);
}

}  // extern "C"
