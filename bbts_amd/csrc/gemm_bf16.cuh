// BF16 MFMA tile engine shared by the single-op split-precision GEMM and the
// grouped persistent stream kernel.
//
// One workgroup = 256 lanes = 4 wave64s in a 2x2 grid computes a BM x BN tile
// of C = alpha * A·B (row-major A: M x K, B: K x N, C: M x N, FP32 out).
// Per K stage (BK = 32) both operands are staged into LDS as k-contiguous
// rows of 32 BF16 (64 B), with 16-byte chunks XOR-swizzled by (row>>1)&3 so
// that eight lanes reading one 16-byte chunk each from eight consecutive rows
// hit eight distinct 16-byte bank groups.  B arrives N-contiguous and is
// transposed in registers on the way in (two k rows x eight n per lane,
// packed into 32-bit pairs).  MFMA: v_mfma_f32_32x32x16_bf16; each wave owns a
// (BM/2) x (BN/2) sub-tile of 32x32 fragments.  Two LDS buffers: the next
// stage's global loads are in flight while the current stage is consumed.
//
// Split precision (S slices, T terms): FP32 operands are split on the fly into
// S BF16 slices x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), ...), and
// the products A_s·B_t with s + t < T are accumulated in FP32.  S = T = 1 on
// BF16 operands is the exact BF16 product; S = T = 2 is the 3-product
// "bf16x3" and S = T = 3 the 6-product "bf16x6" scheme of Henry, Tang and
// Heinecke (ARITH 2019).
#pragma once
#include "common.h"

namespace bbts_hip {
namespace bf16gemm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int BK = 32;
constexpr int kThreads = 256;

__device__ __forceinline__ int sw(int row, int k) {
  return row * BK + ((((k >> 3) ^ ((row >> 1) & 3))) << 3) + (k & 7);
}

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Split x into S bf16 slices (largest first).
template <int S>
__device__ __forceinline__ void split(float x, uint16_t *out) {
  float r = x;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    uint16_t h = bf16_rne(r);
    out[s] = h;
    r = r - bf16_to_f(h);
  }
}

template <typename InT>
struct in_traits;
template <>
struct in_traits<uint16_t> {
  static constexpr int kElemsPer16B = 8;
};
template <>
struct in_traits<float> {
  static constexpr int kElemsPer16B = 4;
};

// LDS footprint (bytes) of a config.
template <int BM, int BN, int S>
struct smem_layout {
  static constexpr int a_elems = BM * BK;
  static constexpr int b_elems = BN * BK;
  static constexpr int bytes = 2 * S * (a_elems + b_elems) * 2;
};

struct op_view {
  const void *A;
  const void *B;
  float *C;
  int64_t lda, ldb, ldc;
  int M, N, K;
  float alpha;
  int vecA, vecB;
};

// Compute one BM x BN tile at (m0, n0).  smem must hold smem_layout<BM,BN,S>::bytes.
template <int BM, int BN, typename InT, int S, int T>
__device__ __forceinline__ void tile(const op_view &op, int m0, int n0, uint16_t *smem) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 32, FN = WN / 32;
  constexpr int AE = BM * BK, BE = BN * BK;
  constexpr int EPV = in_traits<InT>::kElemsPer16B;  // elements per 16-byte load
  // A: BM rows x (BK / EPV) vectors per row
  constexpr int A_VECS = BM * (BK / EPV);
  constexpr int A_PER = (A_VECS + kThreads - 1) / kThreads;
  // B: items of 2 k-rows x EPV n-columns
  constexpr int B_ITEMS = (BK / 2) * (BN / EPV);
  constexpr int B_PER = (B_ITEMS + kThreads - 1) / kThreads;

  uint16_t *As = smem;                      // [2][S][AE]
  uint16_t *Bs = smem + 2 * S * AE;         // [2][S][BE]

  const InT *__restrict__ A = static_cast<const InT *>(op.A);
  const InT *__restrict__ B = static_cast<const InT *>(op.B);
  const int M = op.M, N = op.N, K = op.K;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  uint4 ra[A_PER];
  uint4 rb[B_PER][2];

  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      int q = tid + kThreads * r;
      if (q < A_VECS) {
        int row = q / (BK / EPV), c = q % (BK / EPV);
        int gm = m0 + row, gk = k0 + c * EPV;
        const InT *p = A + (int64_t)gm * op.lda + gk;
        if (op.vecA && gm < M && gk + EPV <= K) {
          ra[r] = *reinterpret_cast<const uint4 *>(p);
        } else {
          InT tmp[EPV];
#pragma unroll
          for (int e = 0; e < EPV; ++e) tmp[e] = (gm < M && gk + e < K) ? p[e] : (InT)0;
          __builtin_memcpy(&ra[r], tmp, 16);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      int q = tid + kThreads * r;
      if (q < B_ITEMS) {
        // k-pair fastest across lanes: the transposing 32-bit LDS stores of one
        // instruction then spread over the 16 banks of a row half (4-way, not 16-way)
        int kp = q % (BK / 2), ng = q / (BK / 2);
        int gk = k0 + kp * 2, gn = n0 + ng * EPV;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const InT *p = B + (int64_t)(gk + h) * op.ldb + gn;
          if (op.vecB && gk + h < K && gn + EPV <= N) {
            rb[r][h] = *reinterpret_cast<const uint4 *>(p);
          } else {
            InT tmp[EPV];
#pragma unroll
            for (int e = 0; e < EPV; ++e) tmp[e] = (gk + h < K && gn + e < N) ? p[e] : (InT)0;
            __builtin_memcpy(&rb[r][h], tmp, 16);
          }
        }
      }
    }
  };

  auto store = [&](int buf) {
    uint16_t *as = As + buf * S * AE;
    uint16_t *bs = Bs + buf * S * BE;
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      int q = tid + kThreads * r;
      if (q < A_VECS) {
        int row = q / (BK / EPV), c = q % (BK / EPV);
        if constexpr (EPV == 8) {
          *reinterpret_cast<uint4 *>(&as[sw(row, c * 8)]) = ra[r];
        } else {
          const float *f = reinterpret_cast<const float *>(&ra[r]);
          uint16_t sl[4][S];
#pragma unroll
          for (int e = 0; e < 4; ++e) split<S>(f[e], sl[e]);
#pragma unroll
          for (int s = 0; s < S; ++s) {
            uint2 v;
            v.x = (uint32_t)sl[0][s] | ((uint32_t)sl[1][s] << 16);
            v.y = (uint32_t)sl[2][s] | ((uint32_t)sl[3][s] << 16);
            *reinterpret_cast<uint2 *>(&as[s * AE + sw(row, c * 4)]) = v;
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      int q = tid + kThreads * r;
      if (q < B_ITEMS) {
        int kp = q % (BK / 2), ng = q / (BK / 2);
        int k = kp * 2, n = ng * EPV;
        if constexpr (EPV == 8) {
          const uint16_t *h0 = reinterpret_cast<const uint16_t *>(&rb[r][0]);
          const uint16_t *h1 = reinterpret_cast<const uint16_t *>(&rb[r][1]);
#pragma unroll
          for (int e = 0; e < 8; ++e)
            *reinterpret_cast<uint32_t *>(&bs[sw(n + e, k)]) = (uint32_t)h0[e] | ((uint32_t)h1[e] << 16);
        } else {
          const float *f0 = reinterpret_cast<const float *>(&rb[r][0]);
          const float *f1 = reinterpret_cast<const float *>(&rb[r][1]);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            uint16_t s0[S], s1[S];
            split<S>(f0[e], s0);
            split<S>(f1[e], s1);
#pragma unroll
            for (int s = 0; s < S; ++s)
              *reinterpret_cast<uint32_t *>(&bs[s * BE + sw(n + e, k)]) = (uint32_t)s0[s] | ((uint32_t)s1[s] << 16);
          }
        }
      }
    }
  };

  f16v acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

  const int nk = (K + BK - 1) / BK;
  __syncthreads();  // LDS may still be read by the previous tile of a persistent loop
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();

  const int half = lane >> 5, col = lane & 31;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    const uint16_t *as = As + buf * S * AE;
    const uint16_t *bs = Bs + buf * S * BE;
#pragma unroll
    for (int kk = 0; kk < BK / 16; ++kk) {
      const int k = kk * 16 + half * 8;
      bf16x8 af[S][FM], bfr[S][FN];
#pragma unroll
      for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          uint4 v = *reinterpret_cast<const uint4 *>(&as[s * AE + sw(wm * WM + i * 32 + col, k)]);
          af[s][i] = __builtin_bit_cast(bf16x8, v);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          uint4 v = *reinterpret_cast<const uint4 *>(&bs[s * BE + sw(wn * WN + j * 32 + col, k)]);
          bfr[s][j] = __builtin_bit_cast(bf16x8, v);
        }
      }
      // small terms first, the leading product last
#pragma unroll
      for (int d = T - 1; d >= 0; --d)
#pragma unroll
        for (int sa = 0; sa < S; ++sa) {
          const int sb = d - sa;
          if (sb < 0 || sb >= S) continue;
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[sa][i], bfr[sb][j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  const float alpha = op.alpha;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn * WN + j * 32 + col;
      if (gn >= N) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int gm = m0 + wm * WM + i * 32 + (v >> 2) * 8 + half * 4 + (v & 3);
        if (gm < M) op.C[(int64_t)gm * op.ldc + gn] = alpha * acc[i][j][v];
      }
    }
}

}  // namespace bf16gemm
}  // namespace bbts_hip
