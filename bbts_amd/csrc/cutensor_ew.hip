// Elementwise and mode-reduction operations of the cutensor library on column-major FP32
// blocks (format {u32 rank; i64 dims[4]}, libraries/cutensor/cu.h:60-63):
//
//   ew           out = alpha * f(in) with an output mode order: out mode r is input mode
//                ordering[r] (ew.h:19-41 parse / set_out_meta, 128-186 cpu_op, 189-358
//                gpu_op).  f: 0 sigmoid, 1 exp, 2 square, 3 relu, 4 relu', 5 sqrt,
//                6 add a constant.
//   ewb          out = op(alpha * lhs, beta * rhs); input mode x is output mode ordering[x],
//                and an input that lacks an output mode is broadcast along it (ewb.h:16-53,
//                212-330).  op: 0 add, 1 max, 2 min, 3 mul, 4 sub, 5 div.
//   ewb_castable out = op(lhs, rhs) on blocks of one shape, op: 0 add, 1 max, 2 min, 3 mul
//                (ewb_castable.h:23-47, utils.h:37-67).
//   reduction    out = alpha * the op-fold of every input element that maps to an output
//                element, over the input modes missing from `ordering` in column-major order;
//                out mode r is input mode ordering[r] (reduction.h:13-37, 114-177).
//
// All index arithmetic is 32-bit (blocks here hold fewer than 2^31 elements); coordinates
// come from a multiply-high division by a precomputed magic number.  Extent-1 modes are
// dropped and neighbouring modes that stay neighbours are merged before a kernel is chosen,
// so most calls reduce to a flat pass, a coalesced gather along the contiguous mode, or one
// LDS-tiled transpose.
#include "common.h"

#include <algorithm>

// op(alpha * l, beta * r) rounds both products before the op: this file is compiled with
// -ffp-contract=off (the build's per-source flags), so no product is fused into an FMA.

namespace bbts_hip {
namespace {

constexpr int kT = 256;
// Non-temporal 16-byte store for outputs written once and not read again by the kernel:
// measured on 2048 x 2048 blocks, ew relu +4 %, the LDS-tiled transposes +19-35 %,
// ewb_castable +8 %; the broadcast-operand ewb form ran 10 % slower with it and keeps
// ordinary stores.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(float4 *p, const float4 &v) {
  __builtin_nontemporal_store(*reinterpret_cast<const nt_f4 *>(&v), reinterpret_cast<nt_f4 *>(p));
}

// n / d for 0 <= n < 2^31, 1 <= d < 2^31: q = (umulhi(n, m) + n) >> s with s = ceil(log2 d)
// and m = floor(2^32 (2^s - d) / d) + 1 (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994; exhaustively checked on the host for n < 2^31).
struct fastdiv {
  uint32_t d = 1, m = 1, s = 0;
};
fastdiv make_fastdiv(uint32_t d) {
  fastdiv f;
  f.d = d;
  while ((uint64_t(1) << f.s) < d) ++f.s;
  f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << f.s) - d)) / d + 1);
  return f;
}
__device__ __forceinline__ uint32_t fd_div(const fastdiv &f, uint32_t n) { return (__umulhi(n, f.m) + n) >> f.s; }

// ---------------------------------------------------------------- element functions
// ew.h:140-171 (cpu_op): the function, then cblas_sscal by alpha.  exp is evaluated in
// double and rounded once to FP32, which is the correctly rounded FP32 exp but for rare
// double-rounding ties (the single-precision device exp loses up to 4 ulp at |v| ~ 7, as it
// scales v by log2(e) in FP32).  Sigmoid is the CPU form exp(v) / (1 + exp(v)) with the
// division in double; where exp(v) overflows FP32 it is 1 (the cuTENSOR sigmoid of gpu_op;
// the CPU form would give inf / inf).  relu' is the CPU form v < 0 ? 0 : 1 (gpu_op's
// floor(2 sigmoid(v)) ignores alpha and reaches 2).
__device__ __forceinline__ float exp_f32(float v) { return (float)exp((double)v); }

template <int F>
__device__ __forceinline__ float ew_fn(float v, float c) {
  if (F == 0) {
    const float e = exp_f32(v);
    return e == INFINITY ? 1.0f : (float)((double)e / (1.0 + (double)e));
  }
  if (F == 1) return exp_f32(v);
  if (F == 2) return v * v;
  if (F == 3) return fmaxf(v, 0.0f);
  if (F == 4) return v < 0.0f ? 0.0f : 1.0f;
  if (F == 5) return sqrtf(v);
  return v + c;
}

// ewb.h:246-253: MKL vsAdd / vsFmax / vsFmin / vsMul / vsSub / vsDiv on the scaled operands
template <int OP>
__device__ __forceinline__ float bin_op(float a, float b) {
  if (OP == 0) return a + b;
  if (OP == 1) return fmaxf(a, b);
  if (OP == 2) return fminf(a, b);
  if (OP == 3) return a * b;
  if (OP == 4) return a - b;
  return a / b;
}

// reduction.h:161-166 folds with utils.h:33-36 scalarAdd / scalarMax / scalarMin / scalarMul
template <int OP>
__device__ __forceinline__ float red_op(float a, float b) {
  if (OP == 0) return a + b;
  if (OP == 1) return a > b ? a : b;
  if (OP == 2) return a < b ? a : b;
  return a * b;
}
template <int OP>
__device__ __forceinline__ float red_identity() {
  return OP == 0 ? 0.0f : (OP == 1 ? -INFINITY : (OP == 2 ? INFINITY : 1.0f));
}

// ---------------------------------------------------------------- ew kernels
template <int F>
__global__ __launch_bounds__(kT) void ew_flat_kernel(const float *__restrict__ in, float *__restrict__ out, uint32_t n,
                                                    float alpha, float c) {
  const uint32_t n4 = n / 4;
  const float4 *in4 = reinterpret_cast<const float4 *>(in);
  float4 *out4 = reinterpret_cast<float4 *>(out);
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n4; i += gridDim.x * kT) {
    float4 v = in4[i];
    v.x = ew_fn<F>(v.x, c) * alpha;
    v.y = ew_fn<F>(v.y, c) * alpha;
    v.z = ew_fn<F>(v.z, c) * alpha;
    v.w = ew_fn<F>(v.w, c) * alpha;
    store_nt(&out4[i], v);
  }
  for (uint32_t i = n4 * 4 + blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT)
    out[i] = ew_fn<F>(in[i], c) * alpha;
}

struct gather_t {
  int rank = 0;
  fastdiv od[4];       // output extents
  uint32_t is[4] = {};  // input stride of each output mode
};

__device__ __forceinline__ uint32_t gather_off(const gather_t &g, uint32_t i) {
  uint32_t off = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r < g.rank) {
      const uint32_t q = r + 1 < g.rank ? fd_div(g.od[r], i) : 0u;
      off += (i - q * g.od[r].d) * g.is[r];
      i = q;
    }
  }
  return off;
}

// output-order gather; coalesced when output mode 0 is input mode 0
template <int F>
__global__ __launch_bounds__(kT) void ew_gather_kernel(const float *__restrict__ in, float *__restrict__ out, uint32_t n,
                                                      gather_t g, float alpha, float c) {
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT)
    out[i] = ew_fn<F>(in[gather_off(g, i)], c) * alpha;
}

// 64 x 64 tiles of (output mode 0 = "a", the output mode that is input mode 0 = "b"), batched
// over the other output modes: reads run along b (input-contiguous), writes along a.
struct transpose_t {
  uint32_t A, B;         // extents of a and b
  uint32_t is_a, os_b;   // input stride of a (b's is 1), output stride of b (a's is 1)
  fastdiv ta;            // tiles along a
  fastdiv tb;            // tiles along b
  fastdiv e0;            // first batch extent (the second is whatever remains)
  uint32_t is0, os0, is1, os1;  // batch strides
};

// V4: A and B are multiples of 4 and both buffers 16-byte aligned (every stride is then a
// multiple of 4 too), so global reads and writes move float4s; the LDS tile stays scalar.
template <int F, bool V4>
__global__ __launch_bounds__(kT) void ew_transpose_kernel(const float *__restrict__ in, float *__restrict__ out,
                                                         transpose_t t, float alpha, float c) {
  __shared__ float tile[64][65];
  uint32_t id = blockIdx.x;
  uint32_t q = fd_div(t.ta, id);
  const uint32_t a0 = (id - q * t.ta.d) * 64;
  id = q;
  q = fd_div(t.tb, id);
  const uint32_t b0 = (id - q * t.tb.d) * 64;
  id = q;
  q = fd_div(t.e0, id);
  const uint32_t c0 = id - q * t.e0.d, c1 = q;
  const float *src = in + c0 * t.is0 + c1 * t.is1;
  float *dst = out + c0 * t.os0 + c1 * t.os1;
  if (V4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int l = threadIdx.x + k * kT, bq = (l & 15) * 4, al = l >> 4;
      const uint32_t a = a0 + al, b = b0 + bq;
      if (a < t.A && b < t.B) {
        const float4 v = *reinterpret_cast<const float4 *>(src + a * t.is_a + b);
        tile[al][bq] = v.x;
        tile[al][bq + 1] = v.y;
        tile[al][bq + 2] = v.z;
        tile[al][bq + 3] = v.w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int l = threadIdx.x + k * kT, aq = (l & 15) * 4, bl = l >> 4;
      const uint32_t a = a0 + aq, b = b0 + bl;
      if (a < t.A && b < t.B) {
        float4 o;
        o.x = ew_fn<F>(tile[aq][bl], c) * alpha;
        o.y = ew_fn<F>(tile[aq + 1][bl], c) * alpha;
        o.z = ew_fn<F>(tile[aq + 2][bl], c) * alpha;
        o.w = ew_fn<F>(tile[aq + 3][bl], c) * alpha;
        store_nt(reinterpret_cast<float4 *>(dst + a + b * t.os_b), o);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int l = threadIdx.x + k * kT, bl = l & 63, al = l >> 6;
    const uint32_t a = a0 + al, b = b0 + bl;
    if (a < t.A && b < t.B) tile[al][bl] = src[a * t.is_a + b];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int l = threadIdx.x + k * kT, al = l & 63, bl = l >> 6;
    const uint32_t a = a0 + al, b = b0 + bl;
    if (a < t.A && b < t.B) dst[a + b * t.os_b] = ew_fn<F>(tile[al][bl], c) * alpha;
  }
}

// ---------------------------------------------------------------- ewb kernels
// SCALE = false is ewb_castable (no alpha / beta factors at all)
template <int OP, bool SCALE>
__device__ __forceinline__ float ewb_val(float l, float r, float alpha, float beta) {
  return SCALE ? bin_op<OP>(alpha * l, beta * r) : bin_op<OP>(l, r);
}

// rank <= 1: each operand is a full vector (stride 1) or one broadcast scalar (stride 0)
template <int OP, bool SCALE>
__global__ __launch_bounds__(kT) void ewb_flat_kernel(const float *__restrict__ l, uint32_t ls, const float *__restrict__ r,
                                                     uint32_t rs, float *out, uint32_t n, float alpha, float beta,
                                                     int vec) {
  if (vec) {  // both operands full, every pointer 16-byte aligned, n % 4 == 0
    const float4 *l4 = reinterpret_cast<const float4 *>(l), *r4 = reinterpret_cast<const float4 *>(r);
    float4 *o4 = reinterpret_cast<float4 *>(out);
    for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n / 4; i += gridDim.x * kT) {
      const float4 a = l4[i], b = r4[i];
      float4 o;
      o.x = ewb_val<OP, SCALE>(a.x, b.x, alpha, beta);
      o.y = ewb_val<OP, SCALE>(a.y, b.y, alpha, beta);
      o.z = ewb_val<OP, SCALE>(a.z, b.z, alpha, beta);
      o.w = ewb_val<OP, SCALE>(a.w, b.w, alpha, beta);
      store_nt(&o4[i], o);
    }
    return;
  }
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT)
    out[i] = ewb_val<OP, SCALE>(l[i * ls], r[i * rs], alpha, beta);
}

struct ewb_t {
  int rank = 0;
  fastdiv od[4];
  uint32_t ls[4] = {}, rs[4] = {};  // operand strides per output mode (0: broadcast)
};

template <int OP>
__global__ __launch_bounds__(kT) void ewb_gather_kernel(const float *__restrict__ l, const float *__restrict__ r,
                                                       float *__restrict__ out, uint32_t n, ewb_t g, float alpha,
                                                       float beta) {
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    uint32_t x = i, lo = 0, ro = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (m < g.rank) {
        const uint32_t q = m + 1 < g.rank ? fd_div(g.od[m], x) : 0u;
        const uint32_t cm = x - q * g.od[m].d;
        lo += cm * g.ls[m];
        ro += cm * g.rs[m];
        x = q;
      }
    }
    out[i] = bin_op<OP>(alpha * l[lo], beta * r[ro]);
  }
}

// Four output elements along output mode 0 per thread: mode 0's extent is a multiple of 4
// and each operand is contiguous (stride 1) or broadcast (stride 0) along it.  g.od[0] holds
// the extent in quads; a contiguous operand's other strides are multiples of 4.
template <int OP>
__global__ __launch_bounds__(kT) void ewb_gather4_kernel(const float *__restrict__ l, const float *__restrict__ r,
                                                        float *__restrict__ out, uint32_t n4, ewb_t g, float alpha,
                                                        float beta) {
  const bool lv = g.ls[0] != 0, rv = g.rs[0] != 0;
  for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n4; i += gridDim.x * kT) {
    uint32_t x = i, lo = 0, ro = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (m < g.rank) {
        const uint32_t q = m + 1 < g.rank ? fd_div(g.od[m], x) : 0u;
        const uint32_t cm = x - q * g.od[m].d;
        lo += m == 0 ? 4 * cm * g.ls[0] : cm * g.ls[m];
        ro += m == 0 ? 4 * cm * g.rs[0] : cm * g.rs[m];
        x = q;
      }
    }
    const float4 a = lv ? *reinterpret_cast<const float4 *>(l + lo) : make_float4(l[lo], l[lo], l[lo], l[lo]);
    const float4 b = rv ? *reinterpret_cast<const float4 *>(r + ro) : make_float4(r[ro], r[ro], r[ro], r[ro]);
    float4 o;
    o.x = bin_op<OP>(alpha * a.x, beta * b.x);
    o.y = bin_op<OP>(alpha * a.y, beta * b.y);
    o.z = bin_op<OP>(alpha * a.z, beta * b.z);
    o.w = bin_op<OP>(alpha * a.w, beta * b.w);
    reinterpret_cast<float4 *>(out)[i] = o;
  }
}

// ---------------------------------------------------------------- reduction kernels
struct red_t {
  int nk = 0, ng = 0;    // kept modes (input order), aggregated modes (input order)
  fastdiv kd[4];         // kept extents, input order
  uint32_t kis[4] = {};  // kept input strides
  uint32_t kos[4] = {};  // kept output strides
  fastdiv od[4];         // output extents, output order (for the wave form)
  uint32_t ois[4] = {};  // input stride of each output mode
  int no = 0;
  uint32_t ge[2] = {1, 1}, gs[2] = {0, 0};  // aggregated extents and strides (<= 2 after merging)
  fastdiv g0;            // ge[0] as a divisor
  uint32_t num_out = 1, A = 1, len = 1, S = 1;  // outputs, fold length, segment length, segments
  int vec = 0;           // wave form: float4 reads of the contiguous run
};

// Input mode 0 kept: one thread per (output, segment), lanes along input mode 0 (coalesced
// reads); the fold runs the aggregated modes in column-major order, exactly the reference's
// left fold when S == 1.  Partials land at [segment][output offset].
template <int OP>
__global__ __launch_bounds__(kT) void red_keep_kernel(const float *__restrict__ in, float *__restrict__ dst, red_t g,
                                                     float alpha) {
  const uint32_t t = blockIdx.x * kT + threadIdx.x;
  if (t >= g.num_out * g.S) return;
  const uint32_t seg = t / g.num_out, o = t - seg * g.num_out;
  uint32_t x = o, base = 0, oo = 0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    if (m < g.nk) {
      const uint32_t q = m + 1 < g.nk ? fd_div(g.kd[m], x) : 0u;
      const uint32_t cm = x - q * g.kd[m].d;
      base += cm * g.kis[m];
      oo += cm * g.kos[m];
      x = q;
    }
  }
  const uint32_t f0 = seg * g.len, f1 = min(g.A, f0 + g.len);
  uint32_t j1 = fd_div(g.g0, f0), j0 = f0 - j1 * g.ge[0];
  const float *p = in + base;
  float acc = p[j0 * g.gs[0] + j1 * g.gs[1]];
  uint32_t left = f1 - f0 - 1;
  if (++j0 == g.ge[0]) {
    j0 = 0;
    ++j1;
  }
  while (left > 0) {
    const uint32_t run = min(g.ge[0] - j0, left);
    const float *q = p + j1 * g.gs[1] + j0 * g.gs[0];
    const uint32_t s0 = g.gs[0];
    uint32_t u = 0;
    for (; u + 4 <= run; u += 4) {
      const float v0 = q[u * s0], v1 = q[(u + 1) * s0], v2 = q[(u + 2) * s0], v3 = q[(u + 3) * s0];
      acc = red_op<OP>(red_op<OP>(red_op<OP>(red_op<OP>(acc, v0), v1), v2), v3);
    }
    for (; u < run; ++u) acc = red_op<OP>(acc, q[u * s0]);
    left -= run;
    j0 = 0;
    ++j1;
  }
  if (g.S == 1) dst[oo] = acc * alpha;
  else dst[seg * g.num_out + oo] = acc;
}

// Input mode 0 aggregated: one wave per (output, segment); lanes stride the segment's
// flattened fold index (contiguous along input mode 0), then a butterfly combines the lanes.
template <int OP>
__global__ __launch_bounds__(kT) void red_wave_kernel(const float *__restrict__ in, float *__restrict__ dst, red_t g,
                                                     float alpha) {
  const uint32_t w = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= g.num_out * g.S) return;  // whole waves leave together
  const uint32_t seg = w / g.num_out, o = w - seg * g.num_out;
  uint32_t x = o, base = 0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    if (m < g.no) {
      const uint32_t q = m + 1 < g.no ? fd_div(g.od[m], x) : 0u;
      base += (x - q * g.od[m].d) * g.ois[m];
      x = q;
    }
  }
  const float *p = in + base;
  const uint32_t f0 = seg * g.len, f1 = min(g.A, f0 + g.len);
  float acc = red_identity<OP>();
  if (g.ng == 1 && g.vec) {  // run length and base offsets multiples of 4, input 16-byte aligned
    for (uint32_t f = f0 + 4 * lane; f < f1; f += 256) {
      const float4 v = *reinterpret_cast<const float4 *>(p + f);
      acc = red_op<OP>(red_op<OP>(red_op<OP>(red_op<OP>(acc, v.x), v.y), v.z), v.w);
    }
  } else if (g.ng == 1) {
    for (uint32_t f = f0 + lane; f < f1; f += 64) acc = red_op<OP>(acc, p[f]);
  } else {
    for (uint32_t f = f0 + lane; f < f1; f += 64) {
      const uint32_t j1 = fd_div(g.g0, f);
      acc = red_op<OP>(acc, p[(f - j1 * g.ge[0]) + j1 * g.gs[1]]);
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) acc = red_op<OP>(acc, __shfl_xor(acc, s));
  if (lane == 0) {
    if (g.S == 1) dst[o] = acc * alpha;
    else dst[seg * g.num_out + o] = acc;
  }
}

// out[o] = alpha * fold over segments, one wave per output: lanes stride the segments, then
// a butterfly combines the lanes
template <int OP>
__global__ __launch_bounds__(kT) void red_combine_wave_kernel(const float *__restrict__ part, float *__restrict__ out,
                                                             uint32_t num_out, uint32_t S, float alpha) {
  const uint32_t o = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= num_out) return;  // whole waves leave together
  float acc = red_identity<OP>();
  for (uint32_t s = lane; s < S; s += 64) acc = red_op<OP>(acc, part[s * num_out + o]);
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) acc = red_op<OP>(acc, __shfl_xor(acc, sh));
  if (lane == 0) out[o] = acc * alpha;
}

// out[o] = alpha * fold over segments, in segment order
template <int OP>
__global__ __launch_bounds__(kT) void red_combine_kernel(const float *__restrict__ part, float *__restrict__ out,
                                                        uint32_t num_out, uint32_t S, float alpha) {
  for (uint32_t o = blockIdx.x * kT + threadIdx.x; o < num_out; o += gridDim.x * kT) {
    float acc = part[o];
    for (uint32_t s = 1; s < S; ++s) acc = red_op<OP>(acc, part[s * num_out + o]);
    out[o] = acc * alpha;
  }
}

// ---------------------------------------------------------------- host side
// grid of a grid-stride kernel
unsigned blocks_for(uint64_t work, uint64_t per_block) {
  uint64_t b = (work + per_block - 1) / per_block;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 16384));
}
// grid of a kernel with one work item per thread (or per wave): every item gets a block
unsigned blocks_exact(uint64_t work, uint64_t per_block) { return (unsigned)std::max<uint64_t>(1, (work + per_block - 1) / per_block); }

float word_f(int32_t w) {
  float f;
  std::memcpy(&f, &w, 4);
  return f;
}

int check_dims(int rank, const int64_t *dims, uint64_t *n) {
  BBTS_REQUIRE(rank >= 0 && rank <= 4, BBTS_E_INVALID, "cutensor rank %d outside 0..4", rank);
  uint64_t p = 1;
  for (int r = 0; r < rank; ++r) {
    BBTS_REQUIRE(dims[r] >= 0, BBTS_E_INVALID, "negative extent");
    p *= (uint64_t)dims[r];
    BBTS_REQUIRE(p < (uint64_t(1) << 31), BBTS_E_UNSUPPORTED, "block of 2^31 or more elements");
  }
  *n = p;
  return BBTS_OK;
}

int check_perm(const std::vector<int> &ord, int rank, const char *what) {
  std::vector<int> seen((size_t)std::max(rank, 1), 0);
  for (int m : ord) {
    BBTS_REQUIRE(m >= 0 && m < rank, BBTS_E_INVALID, "%s: mode %d outside 0..%d", what, m, rank - 1);
    BBTS_REQUIRE(!seen[(size_t)m]++, BBTS_E_INVALID, "%s: mode %d repeated", what, m);
  }
  return BBTS_OK;
}

// ---- ew
struct ew_params {
  int f = 0;
  float c = 0.0f, alpha = 1.0f;
  std::vector<int> ord;
};

// ew.h:21-29 parse + 130-136 checks ("input rank is incorrect")
int parse_ew(const int32_t *w, int nw, int rank, ew_params *p) {
  BBTS_REQUIRE(nw >= 2, BBTS_E_INVALID, "ew: too few parameters");
  int i = 0;
  p->f = w[i++];
  BBTS_REQUIRE(p->f >= 0 && p->f <= 6, BBTS_E_INVALID, "ew: %d is not a valid unary op", p->f);
  if (p->f == 6) {
    BBTS_REQUIRE(nw >= 3, BBTS_E_INVALID, "ew: add-constant needs its constant");
    p->c = word_f(w[i++]);
  }
  p->alpha = word_f(w[i++]);
  p->ord.assign(w + i, w + nw);
  BBTS_REQUIRE((int)p->ord.size() == rank, BBTS_E_INVALID, "ew: input rank is incorrect (ordering of %zu modes, rank %d)",
               p->ord.size(), rank);
  return check_perm(p->ord, rank, "ew ordering");
}

template <int F>
int launch_ew_f(const float *in, float *out, int rank, const int64_t *dims, const std::vector<int> &ord, uint32_t n,
                float alpha, float c, hipStream_t s) {
  // drop extent-1 modes, then merge input neighbours that stay output neighbours
  std::vector<int> pos((size_t)rank);
  for (int r = 0; r < rank; ++r) pos[(size_t)ord[(size_t)r]] = r;
  std::vector<int> keep;
  for (int i = 0; i < rank; ++i)
    if (dims[i] != 1) keep.push_back(i);
  std::vector<int> kout(keep);  // kept modes in output order
  std::sort(kout.begin(), kout.end(), [&](int a, int b) { return pos[(size_t)a] < pos[(size_t)b]; });
  std::vector<int> orank((size_t)rank, -1);
  for (size_t r = 0; r < kout.size(); ++r) orank[(size_t)kout[r]] = (int)r;
  std::vector<uint64_t> gd;  // merged input extents
  std::vector<int> gfirst;   // output rank of each group's first mode
  for (size_t k = 0; k < keep.size(); ++k) {
    const int m = keep[k];
    if (k > 0 && orank[(size_t)m] == orank[(size_t)keep[k - 1]] + 1) {
      gd.back() *= (uint64_t)dims[m];
    } else {
      gd.push_back((uint64_t)dims[m]);
      gfirst.push_back(orank[(size_t)m]);
    }
  }
  const int R = (int)gd.size();
  BBTS_REQUIRE(in != out || R <= 1, BBTS_E_INVALID, "ew: a reordering cannot run in place");
  std::vector<int> perm((size_t)R);  // out mode r = merged input mode perm[r]
  for (int r = 0; r < R; ++r) perm[(size_t)r] = r;
  std::sort(perm.begin(), perm.end(), [&](int a, int b) { return gfirst[(size_t)a] < gfirst[(size_t)b]; });
  std::vector<uint32_t> is((size_t)R);
  uint64_t st = 1;
  for (int m = 0; m < R; ++m) {
    is[(size_t)m] = (uint32_t)st;
    st *= gd[(size_t)m];
  }
  if (R <= 1) {
    const bool vec = (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    if (vec) {
      hipLaunchKernelGGL(ew_flat_kernel<F>, dim3(blocks_for(n / 4 + 1, kT * 4)), dim3(kT), 0, s, in, out, n, alpha, c);
    } else {
      gather_t g;
      g.rank = 1;
      g.od[0] = make_fastdiv(std::max<uint32_t>(n, 1));
      g.is[0] = 1;
      hipLaunchKernelGGL(ew_gather_kernel<F>, dim3(blocks_for(n, kT * 4)), dim3(kT), 0, s, in, out, n, g, alpha, c);
    }
    BBTS_LAUNCH_CHECK();
    return BBTS_OK;
  }
  if (perm[0] == 0) {
    gather_t g;
    g.rank = R;
    for (int r = 0; r < R; ++r) {
      g.od[r] = make_fastdiv((uint32_t)gd[(size_t)perm[(size_t)r]]);
      g.is[r] = is[(size_t)perm[(size_t)r]];
    }
    hipLaunchKernelGGL(ew_gather_kernel<F>, dim3(blocks_for(n, kT * 4)), dim3(kT), 0, s, in, out, n, g, alpha, c);
    BBTS_LAUNCH_CHECK();
    return BBTS_OK;
  }
  // tiled transpose: a = output mode 0, b = the output position of input mode 0
  std::vector<uint32_t> os((size_t)R);
  st = 1;
  for (int r = 0; r < R; ++r) {
    os[(size_t)r] = (uint32_t)st;
    st *= gd[(size_t)perm[(size_t)r]];
  }
  int pb = 0;
  for (int r = 0; r < R; ++r)
    if (perm[(size_t)r] == 0) pb = r;
  transpose_t t{};
  t.A = (uint32_t)gd[(size_t)perm[0]];
  t.B = (uint32_t)gd[0];
  t.is_a = is[(size_t)perm[0]];
  t.os_b = os[(size_t)pb];
  uint32_t ext[2] = {1, 1}, bis[2] = {0, 0}, bos[2] = {0, 0};
  int nb = 0;
  for (int r = 1; r < R; ++r) {
    if (r == pb) continue;
    ext[nb] = (uint32_t)gd[(size_t)perm[(size_t)r]];
    bis[nb] = is[(size_t)perm[(size_t)r]];
    bos[nb] = os[(size_t)r];
    ++nb;
  }
  const uint32_t ta = (t.A + 63) / 64, tb = (t.B + 63) / 64;
  t.ta = make_fastdiv(ta);
  t.tb = make_fastdiv(tb);
  t.e0 = make_fastdiv(ext[0]);
  t.is0 = bis[0];
  t.os0 = bos[0];
  t.is1 = bis[1];
  t.os1 = bos[1];
  const uint64_t tiles = (uint64_t)ta * tb * ext[0] * ext[1];
  BBTS_REQUIRE(tiles < (uint64_t(1) << 31), BBTS_E_UNSUPPORTED, "too many transpose tiles");
  const bool v4 = t.A % 4 == 0 && t.B % 4 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(out) % 16 == 0;
  if (v4) hipLaunchKernelGGL((ew_transpose_kernel<F, true>), dim3((unsigned)tiles), dim3(kT), 0, s, in, out, t, alpha, c);
  else hipLaunchKernelGGL((ew_transpose_kernel<F, false>), dim3((unsigned)tiles), dim3(kT), 0, s, in, out, t, alpha, c);
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}

// parse + the out shape (set_out_meta, ew.h:31-41)
int ew_shape(const int32_t *w, int nw, int rank, const int64_t *dims, ew_params *p, cu_shape_t *so, uint64_t *n) {
  int rc = parse_ew(w, nw, rank, p);
  if (rc) return rc;
  if ((rc = check_dims(rank, dims, n))) return rc;
  *so = cu_shape_t{};
  so->rank = (uint32_t)rank;
  for (int r = 0; r < rank; ++r) so->dims[r] = dims[p->ord[(size_t)r]];
  return BBTS_OK;
}

int run_ew(const int32_t *w, int nw, int rank, const int64_t *dims, const float *in, float *out, hipStream_t s) {
  ew_params p;
  cu_shape_t so;
  uint64_t n;
  int rc = ew_shape(w, nw, rank, dims, &p, &so, &n);
  if (rc) return rc;
  if (n == 0) return BBTS_OK;
  BBTS_REQUIRE(in && out, BBTS_E_INVALID, "null pointer");
  switch (p.f) {
    case 0: return launch_ew_f<0>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    case 1: return launch_ew_f<1>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    case 2: return launch_ew_f<2>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    case 3: return launch_ew_f<3>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    case 4: return launch_ew_f<4>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    case 5: return launch_ew_f<5>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
    default: return launch_ew_f<6>(in, out, rank, dims, p.ord, (uint32_t)n, p.alpha, p.c, s);
  }
}

// ---- ewb
struct ewb_params {
  int op = 0;
  float alpha = 1.0f, beta = 1.0f;
  std::vector<int> ol, orr;
};

// ewb.h:21-36 parse, 38-53 set_out_meta.  The reference takes every word after rank_rhs as
// the rhs ordering and leaves an output mode that neither input names uninitialised; here
// the orderings must match the inputs' ranks and cover every output mode with one extent.
int parse_ewb(const int32_t *w, int nw, int rank_l, const int64_t *dl, int rank_r, const int64_t *dr, ewb_params *p,
              int *rank_o, int64_t *dims_o) {
  BBTS_REQUIRE(nw >= 4, BBTS_E_INVALID, "ewb: too few parameters");
  p->op = w[0];
  BBTS_REQUIRE(p->op >= 0 && p->op <= 5, BBTS_E_INVALID, "ewb: invalid binary op %d", p->op);
  p->alpha = word_f(w[1]);
  p->beta = word_f(w[2]);
  const int rl = w[3];
  BBTS_REQUIRE(rl >= 0 && rl <= 4 && 4 + rl < nw, BBTS_E_INVALID, "ewb: bad lhs rank word %d", rl);
  p->ol.assign(w + 4, w + 4 + rl);
  const int rr = w[4 + rl];
  p->orr.assign(w + 5 + rl, w + nw);
  BBTS_REQUIRE(rl == rank_l && (int)p->orr.size() == rank_r && rr == rank_r, BBTS_E_INVALID,
               "ewb: orderings of %d and %zu modes for inputs of rank %d and %d", rl, p->orr.size(), rank_l, rank_r);
  int ro = 0;
  for (int m : p->ol) ro = std::max(ro, m + 1);
  for (int m : p->orr) ro = std::max(ro, m + 1);
  BBTS_REQUIRE(ro <= 4, BBTS_E_INVALID, "ewb: output rank %d above 4", ro);
  int rc;
  if ((rc = check_perm(p->ol, ro, "ewb lhs ordering")) || (rc = check_perm(p->orr, ro, "ewb rhs ordering"))) return rc;
  for (int r = 0; r < 4; ++r) dims_o[r] = -1;
  for (int x = 0; x < rank_l; ++x) dims_o[p->ol[(size_t)x]] = dl[x];
  for (int x = 0; x < rank_r; ++x) {
    int64_t &d = dims_o[p->orr[(size_t)x]];
    BBTS_REQUIRE(d < 0 || d == dr[x], BBTS_E_SHAPE, "ewb: output mode %d is %lld in lhs and %lld in rhs",
                 p->orr[(size_t)x], (long long)d, (long long)dr[x]);
    d = dr[x];
  }
  for (int r = 0; r < ro; ++r)
    BBTS_REQUIRE(dims_o[r] >= 0, BBTS_E_INVALID, "ewb: output mode %d is in neither input", r);
  *rank_o = ro;
  return BBTS_OK;
}

// One operand ("staged") runs contiguously along an output mode b other than mode 0 and with
// a stride along output mode 0; the other ("direct") is contiguous or broadcast along mode 0.
// 64 x 64 tiles of (a = output mode 0, b), batched over up to two more output modes: the
// staged operand is read along b into LDS, the direct operand and the output run along a.
struct ewbt_t {
  uint32_t A, B;                    // extents of a and b
  uint32_t os_b, ds_a, ds_b, ss_a;  // output b stride; direct a (0 or 1) and b strides; staged a stride
  fastdiv ta, tb, e0;               // tiles along a and b, first batch extent
  uint32_t os0, os1, ds0, ds1, ss0, ss1;  // batch strides: output, direct, staged
};

// V4: A and B are multiples of 4, the buffers 16-byte aligned and every stride a multiple of
// 4 (a broadcast direct operand aside), so global reads and writes move float4s.
template <int OP, bool LSTAGED, bool V4>
__global__ __launch_bounds__(kT) void ewb_transpose_kernel(const float *__restrict__ l, const float *__restrict__ r,
                                                          float *__restrict__ out, ewbt_t t, float alpha, float beta) {
  __shared__ float tile[64][65];
  uint32_t id = blockIdx.x;
  uint32_t q = fd_div(t.ta, id);
  const uint32_t a0 = (id - q * t.ta.d) * 64;
  id = q;
  q = fd_div(t.tb, id);
  const uint32_t b0 = (id - q * t.tb.d) * 64;
  id = q;
  q = fd_div(t.e0, id);
  const uint32_t c0 = id - q * t.e0.d, c1 = q;
  const float *st = (LSTAGED ? l : r) + c0 * t.ss0 + c1 * t.ss1;
  const float *di = (LSTAGED ? r : l) + c0 * t.ds0 + c1 * t.ds1;
  float *dst = out + c0 * t.os0 + c1 * t.os1;
  // op(alpha * lhs, beta * rhs) with the staged value in its operand's place
  auto val = [&](float sv, float dv) {
    return LSTAGED ? bin_op<OP>(alpha * sv, beta * dv) : bin_op<OP>(alpha * dv, beta * sv);
  };
  if (V4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int x = threadIdx.x + k * kT, bq = (x & 15) * 4, al = x >> 4;
      const uint32_t a = a0 + al, b = b0 + bq;
      if (a < t.A && b < t.B) {
        const float4 v = *reinterpret_cast<const float4 *>(st + a * t.ss_a + b);
        tile[al][bq] = v.x;
        tile[al][bq + 1] = v.y;
        tile[al][bq + 2] = v.z;
        tile[al][bq + 3] = v.w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int x = threadIdx.x + k * kT, aq = (x & 15) * 4, bl = x >> 4;
      const uint32_t a = a0 + aq, b = b0 + bl;
      if (a < t.A && b < t.B) {
        float4 d;
        if (t.ds_a) {
          d = *reinterpret_cast<const float4 *>(di + a + b * t.ds_b);
        } else {
          const float v = di[b * t.ds_b];
          d = make_float4(v, v, v, v);
        }
        float4 o;
        o.x = val(tile[aq][bl], d.x);
        o.y = val(tile[aq + 1][bl], d.y);
        o.z = val(tile[aq + 2][bl], d.z);
        o.w = val(tile[aq + 3][bl], d.w);
        store_nt(reinterpret_cast<float4 *>(dst + a + b * t.os_b), o);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int x = threadIdx.x + k * kT, bl = x & 63, al = x >> 6;
    const uint32_t a = a0 + al, b = b0 + bl;
    if (a < t.A && b < t.B) tile[al][bl] = st[a * t.ss_a + b];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int x = threadIdx.x + k * kT, al = x & 63, bl = x >> 6;
    const uint32_t a = a0 + al, b = b0 + bl;
    if (a < t.A && b < t.B) dst[a + b * t.os_b] = val(tile[al][bl], di[a * t.ds_a + b * t.ds_b]);
  }
}

// The tiled form for g when one operand is strided along output mode 0 and contiguous along
// another mode b of at least 32, and the other operand is contiguous or broadcast along mode 0
// (also at least 32); false when g is not of that form.
template <int OP>
bool launch_ewb_transpose(const float *l, const float *r, float *out, const ewb_t &g, float alpha, float beta,
                          hipStream_t s) {
  const uint32_t A = g.od[0].d;
  if (g.rank < 2 || A < 32) return false;
  for (int lst = 0; lst < 2; ++lst) {
    const uint32_t *S = lst ? g.ls : g.rs, *D = lst ? g.rs : g.ls;
    if (D[0] > 1 || S[0] <= 1) continue;
    int b = -1;
    for (int m = 1; m < g.rank; ++m)
      if (S[m] == 1) b = m;
    if (b < 0 || g.od[b].d < 32) continue;
    uint32_t os[4], o = 1;
    for (int m = 0; m < g.rank; ++m) {
      os[m] = o;
      o *= g.od[m].d;
    }
    ewbt_t t{};
    t.A = A;
    t.B = g.od[b].d;
    t.os_b = os[b];
    t.ds_a = D[0];
    t.ds_b = D[b];
    t.ss_a = S[0];
    uint32_t ext[2] = {1, 1};
    int nb = 0;
    for (int m = 1; m < g.rank; ++m) {
      if (m == b) continue;
      ext[nb] = g.od[m].d;
      (nb ? t.os1 : t.os0) = os[m];
      (nb ? t.ds1 : t.ds0) = D[m];
      (nb ? t.ss1 : t.ss0) = S[m];
      ++nb;
    }
    const uint32_t ta = (A + 63) / 64, tb = (t.B + 63) / 64;
    t.ta = make_fastdiv(ta);
    t.tb = make_fastdiv(tb);
    t.e0 = make_fastdiv(ext[0]);
    const uint64_t blocks = (uint64_t)ta * tb * ext[0] * ext[1];
    const auto al16 = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    const float *st = lst ? l : r, *di = lst ? r : l;
    bool v4 = A % 4 == 0 && t.B % 4 == 0 && al16(out) && al16(st) && t.ss_a % 4 == 0 && t.ss0 % 4 == 0 &&
              t.ss1 % 4 == 0 && (t.ds_a == 0 || (al16(di) && t.ds_b % 4 == 0 && t.ds0 % 4 == 0 && t.ds1 % 4 == 0));
    if (lst) {
      if (v4) hipLaunchKernelGGL((ewb_transpose_kernel<OP, true, true>), dim3((uint32_t)blocks), dim3(kT), 0, s, l, r, out, t, alpha, beta);
      else hipLaunchKernelGGL((ewb_transpose_kernel<OP, true, false>), dim3((uint32_t)blocks), dim3(kT), 0, s, l, r, out, t, alpha, beta);
    } else {
      if (v4) hipLaunchKernelGGL((ewb_transpose_kernel<OP, false, true>), dim3((uint32_t)blocks), dim3(kT), 0, s, l, r, out, t, alpha, beta);
      else hipLaunchKernelGGL((ewb_transpose_kernel<OP, false, false>), dim3((uint32_t)blocks), dim3(kT), 0, s, l, r, out, t, alpha, beta);
    }
    return true;
  }
  return false;
}

template <int OP, bool SCALE>
int launch_ewb_op(const float *l, const float *r, float *out, const ewb_t &g, uint32_t n, float alpha, float beta,
                  hipStream_t s) {
  if (g.rank <= 1) {
    const uint32_t ls = g.rank ? g.ls[0] : 0u, rs = g.rank ? g.rs[0] : 0u;
    const int vec = ls == 1 && rs == 1 && n % 4 == 0 && reinterpret_cast<uintptr_t>(l) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(r) % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0;
    hipLaunchKernelGGL((ewb_flat_kernel<OP, SCALE>), dim3(blocks_for(vec ? n / 4 : n, kT * 4)), dim3(kT), 0, s, l, ls,
                       r, rs, out, n, alpha, beta, vec);
  } else if (SCALE && launch_ewb_transpose<OP>(l, r, out, g, alpha, beta, s)) {
    // tiled (one operand transposed against the output)
  } else {
    const auto al16 = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    const uint32_t d0 = g.od[0].d;
    bool quad = d0 % 4 == 0 && g.ls[0] <= 1 && g.rs[0] <= 1 && al16(out) && (g.ls[0] == 0 || al16(l)) &&
                (g.rs[0] == 0 || al16(r));
    for (int m = 1; m < g.rank; ++m) quad = quad && (g.ls[0] == 0 || g.ls[m] % 4 == 0) && (g.rs[0] == 0 || g.rs[m] % 4 == 0);
    if (quad) {
      ewb_t g4 = g;
      g4.od[0] = make_fastdiv(d0 / 4);
      hipLaunchKernelGGL(ewb_gather4_kernel<OP>, dim3(blocks_for(n / 4, kT * 4)), dim3(kT), 0, s, l, r, out, n / 4, g4,
                         alpha, beta);
    } else {
      hipLaunchKernelGGL(ewb_gather_kernel<OP>, dim3(blocks_for(n, kT * 4)), dim3(kT), 0, s, l, r, out, n, g, alpha,
                         beta);
    }
  }
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}

int run_ewb(const int32_t *w, int nw, int rank_l, const int64_t *dl, const float *l, int rank_r, const int64_t *dr,
            const float *r, float *out, cu_shape_t *so, hipStream_t s) {
  ewb_params p;
  int ro;
  int64_t dout[4];
  int rc = parse_ewb(w, nw, rank_l, dl, rank_r, dr, &p, &ro, dout);
  if (rc) return rc;
  uint64_t n, nl, nr;
  if ((rc = check_dims(ro, dout, &n)) || (rc = check_dims(rank_l, dl, &nl)) || (rc = check_dims(rank_r, dr, &nr)))
    return rc;
  if (so) {
    *so = cu_shape_t{};
    so->rank = (uint32_t)ro;
    for (int x = 0; x < ro; ++x) so->dims[x] = dout[x];
  }
  if (n == 0) return BBTS_OK;
  BBTS_REQUIRE(l && r && out, BBTS_E_INVALID, "null pointer");
  // operand strides per output mode (column-major in the operand's own mode order)
  uint64_t ls[4] = {0, 0, 0, 0}, rs[4] = {0, 0, 0, 0};
  uint64_t st = 1;
  for (int x = 0; x < rank_l; ++x) {
    ls[p.ol[(size_t)x]] = st;
    st *= (uint64_t)dl[x];
  }
  st = 1;
  for (int x = 0; x < rank_r; ++x) {
    rs[p.orr[(size_t)x]] = st;
    st *= (uint64_t)dr[x];
  }
  // drop extent-1 output modes; merge output neighbours whose strides continue in both operands
  ewb_t g;
  std::vector<uint64_t> gd, gl, gr;
  for (int m = 0; m < ro; ++m) {
    if (dout[m] == 1) continue;
    if (!gd.empty()) {
      const uint64_t d = gd.back();
      const bool lc = (gl.back() == 0 && ls[m] == 0) || (gl.back() != 0 && ls[m] == gl.back() * d);
      const bool rcn = (gr.back() == 0 && rs[m] == 0) || (gr.back() != 0 && rs[m] == gr.back() * d);
      if (lc && rcn) {
        gd.back() *= (uint64_t)dout[m];
        continue;
      }
    }
    gd.push_back((uint64_t)dout[m]);
    gl.push_back(ls[m]);
    gr.push_back(rs[m]);
  }
  g.rank = (int)gd.size();
  for (int m = 0; m < g.rank; ++m) {
    g.od[m] = make_fastdiv((uint32_t)gd[(size_t)m]);
    g.ls[m] = (uint32_t)gl[(size_t)m];
    g.rs[m] = (uint32_t)gr[(size_t)m];
  }
  const uint32_t nn = (uint32_t)n;
  switch (p.op) {
    case 0: return launch_ewb_op<0, true>(l, r, out, g, nn, p.alpha, p.beta, s);
    case 1: return launch_ewb_op<1, true>(l, r, out, g, nn, p.alpha, p.beta, s);
    case 2: return launch_ewb_op<2, true>(l, r, out, g, nn, p.alpha, p.beta, s);
    case 3: return launch_ewb_op<3, true>(l, r, out, g, nn, p.alpha, p.beta, s);
    case 4: return launch_ewb_op<4, true>(l, r, out, g, nn, p.alpha, p.beta, s);
    default: return launch_ewb_op<5, true>(l, r, out, g, nn, p.alpha, p.beta, s);
  }
}

int run_castable(int op, const float *l, const float *r, float *out, uint64_t n, hipStream_t s) {
  BBTS_REQUIRE(op >= 0 && op <= 3, BBTS_E_INVALID, "ewb_castable: %d is not a castable operator", op);
  BBTS_REQUIRE(n < (uint64_t(1) << 31), BBTS_E_UNSUPPORTED, "block of 2^31 or more elements");
  if (n == 0) return BBTS_OK;
  BBTS_REQUIRE(l && r && out, BBTS_E_INVALID, "null pointer");
  ewb_t g;
  g.rank = 1;
  g.od[0] = make_fastdiv((uint32_t)n);
  g.ls[0] = g.rs[0] = 1;
  switch (op) {
    case 0: return launch_ewb_op<0, false>(l, r, out, g, (uint32_t)n, 1.0f, 1.0f, s);
    case 1: return launch_ewb_op<1, false>(l, r, out, g, (uint32_t)n, 1.0f, 1.0f, s);
    case 2: return launch_ewb_op<2, false>(l, r, out, g, (uint32_t)n, 1.0f, 1.0f, s);
    default: return launch_ewb_op<3, false>(l, r, out, g, (uint32_t)n, 1.0f, 1.0f, s);
  }
}

// ---- reduction
// The partials of a split fold go to a grow-only buffer kept per stream: work on one stream
// runs in order, so the buffer is reused without waiting, and a graph captured with it stays
// valid (an outgrown buffer is kept until shutdown).  The first call on a stream with a new
// largest size allocates, so it belongs outside graph capture.
int stream_scratch(bbts_hip_ctx *ctx, hipStream_t s, size_t bytes, void **p) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto &v = ctx->stream_scratch[reinterpret_cast<uintptr_t>(s)];
  if (v.empty() || v.back().second < bytes) {
    const size_t sz = std::max(bytes, v.empty() ? (size_t(1) << 20) : 2 * v.back().second);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sz);
    if (e != hipSuccess) {
      set_error("reduction scratch of %zu bytes: %s", sz, hipGetErrorString(e));
      return BBTS_E_OOM;
    }
    v.emplace_back(q, sz);
  }
  *p = v.back().first;
  return BBTS_OK;
}

template <int OP>
int launch_red_op(bbts_hip_ctx *ctx, const float *in, float *out, red_t g, bool wave, float alpha, hipStream_t s) {
  float *dst = out;
  if (g.S > 1) {
    void *tmp = nullptr;
    int rc = stream_scratch(ctx, s, (size_t)g.S * g.num_out * sizeof(float), &tmp);
    if (rc) return rc;
    dst = static_cast<float *>(tmp);
  }
  int rc = BBTS_OK;
  if (wave) hipLaunchKernelGGL(red_wave_kernel<OP>, dim3(blocks_exact((uint64_t)g.num_out * g.S, kT / 64)), dim3(kT),
                               0, s, in, dst, g, alpha);
  else hipLaunchKernelGGL(red_keep_kernel<OP>, dim3(blocks_exact((uint64_t)g.num_out * g.S, kT)), dim3(kT), 0, s, in,
                          dst, g, alpha);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && g.S > 1) {
    if (g.S > 16)
      hipLaunchKernelGGL(red_combine_wave_kernel<OP>, dim3(blocks_exact(g.num_out, kT / 64)), dim3(kT), 0, s, dst,
                         out, g.num_out, g.S, alpha);
    else
      hipLaunchKernelGGL(red_combine_kernel<OP>, dim3(blocks_for(g.num_out, kT)), dim3(kT), 0, s, dst, out,
                         g.num_out, g.S, alpha);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    set_error("reduction launch failed: %s", hipGetErrorString(e));
    rc = BBTS_E_HIP;
  }
  return rc;
}

// reduction.h:15-26 parse, 29-37 set_out_meta, 123-125 ("input rank is not great enough")
int run_reduction(bbts_hip_ctx *ctx, const int32_t *w, int nw, int rank, const int64_t *dims, const float *in,
                  float *out, cu_shape_t *so, int segments, hipStream_t s) {
  BBTS_REQUIRE(nw >= 2, BBTS_E_INVALID, "reduction: too few parameters");
  const int op = w[0];
  BBTS_REQUIRE(op >= 0 && op <= 3, BBTS_E_INVALID, "reduction: %d is not a castable operator", op);
  const float alpha = word_f(w[1]);
  std::vector<int> ord(w + 2, w + nw);
  for (int m : ord)
    BBTS_REQUIRE(m < rank, BBTS_E_INVALID, "reduction: input rank is not great enough (mode %d, rank %d)", m, rank);
  int rc = check_perm(ord, rank, "reduction ordering");
  if (rc) return rc;
  uint64_t n;
  if ((rc = check_dims(rank, dims, &n))) return rc;
  const int ro = (int)ord.size();
  if (so) {
    *so = cu_shape_t{};
    so->rank = (uint32_t)ro;
    for (int r = 0; r < ro; ++r) so->dims[r] = dims[ord[(size_t)r]];
  }
  std::vector<int> pos((size_t)rank, -1);
  for (int r = 0; r < ro; ++r) pos[(size_t)ord[(size_t)r]] = r;
  uint64_t num_out = 1, A = 1;
  for (int i = 0; i < rank; ++i) (pos[(size_t)i] >= 0 ? num_out : A) *= (uint64_t)dims[i];
  if (num_out == 0) return BBTS_OK;
  BBTS_REQUIRE(A > 0, BBTS_E_INVALID, "reduction over a mode of extent 0");
  BBTS_REQUIRE(in && out, BBTS_E_INVALID, "null pointer");
  BBTS_REQUIRE(segments >= 0, BBTS_E_INVALID, "negative segment count");
  // input and output strides, then drop extent-1 modes and merge neighbours of one kind
  std::vector<uint64_t> ist((size_t)rank), ost((size_t)std::max(ro, 1));
  uint64_t st = 1;
  for (int i = 0; i < rank; ++i) {
    ist[(size_t)i] = st;
    st *= (uint64_t)dims[i];
  }
  st = 1;
  for (int r = 0; r < ro; ++r) {
    ost[(size_t)r] = st;
    st *= (uint64_t)dims[ord[(size_t)r]];
  }
  struct mode {
    bool kept;
    uint64_t ext, is, os;
    int opos;  // output position of the group's first mode
  };
  std::vector<mode> ms;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] == 1) continue;
    const bool kept = pos[(size_t)i] >= 0;
    if (!ms.empty() && ms.back().kept == kept) {
      mode &b = ms.back();
      const bool contiguous = b.is * b.ext == ist[(size_t)i];
      const bool out_next = !kept || (b.os * b.ext == ost[(size_t)pos[(size_t)i]]);
      if (contiguous && out_next) {
        b.ext *= (uint64_t)dims[i];
        continue;
      }
    }
    ms.push_back({kept, (uint64_t)dims[i], ist[(size_t)i], kept ? ost[(size_t)pos[(size_t)i]] : 0,
                  kept ? pos[(size_t)i] : -1});
  }
  red_t g;
  std::vector<mode> km, gm;
  for (auto &m : ms) (m.kept ? km : gm).push_back(m);
  BBTS_REQUIRE(gm.size() <= 2 && km.size() <= 4, BBTS_E_INVALID, "reduction: mode layout");
  g.nk = (int)km.size();
  for (int k = 0; k < g.nk; ++k) {
    g.kd[k] = make_fastdiv((uint32_t)km[(size_t)k].ext);
    g.kis[k] = (uint32_t)km[(size_t)k].is;
    g.kos[k] = (uint32_t)km[(size_t)k].os;
  }
  std::vector<mode> kout(km);
  std::sort(kout.begin(), kout.end(), [](const mode &a, const mode &b) { return a.opos < b.opos; });
  g.no = (int)kout.size();
  for (int r = 0; r < g.no; ++r) {
    g.od[r] = make_fastdiv((uint32_t)kout[(size_t)r].ext);
    g.ois[r] = (uint32_t)kout[(size_t)r].is;
  }
  g.ng = (int)gm.size();
  for (int k = 0; k < g.ng; ++k) {
    g.ge[k] = (uint32_t)gm[(size_t)k].ext;
    g.gs[k] = (uint32_t)gm[(size_t)k].is;
  }
  g.g0 = make_fastdiv(g.ge[0]);
  g.num_out = (uint32_t)num_out;
  g.A = (uint32_t)A;
  // the wave form when the input's contiguous mode is folded; one thread per output otherwise
  const bool wave = g.ng > 0 && gm[0].is == 1;
  uint64_t S = 1;
  if (segments > 0) {
    S = (uint64_t)segments;
  } else if (wave) {
    // split only below 2048 waves (8 per CU): at 2048 x 2048 -> 2048 one pass ran 5.7 us and
    // the 4-way split with its combine 7.5 us
    if (num_out < 2048 && A > 512) S = std::min<uint64_t>((2048 + num_out - 1) / num_out, (A + 511) / 512);
  } else {
    if (num_out < 262144 && A > 32) S = std::min<uint64_t>((262144 + num_out - 1) / num_out, (A + 15) / 16);
  }
  S = std::max<uint64_t>(1, std::min<uint64_t>(S, A));
  uint64_t len = (A + S - 1) / S;
  if (wave && S > 1) len = (len + 63) / 64 * 64;
  S = (A + len - 1) / len;
  BBTS_REQUIRE(S * num_out < (uint64_t(1) << 31), BBTS_E_UNSUPPORTED, "too many reduction segments");
  g.len = (uint32_t)len;
  g.S = (uint32_t)S;
  g.vec = wave && g.ng == 1 && g.ge[0] % 4 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0;
  switch (op) {
    case 0: return launch_red_op<0>(ctx, in, out, g, wave, alpha, s);
    case 1: return launch_red_op<1>(ctx, in, out, g, wave, alpha, s);
    case 2: return launch_red_op<2>(ctx, in, out, g, wave, alpha, s);
    default: return launch_red_op<3>(ctx, in, out, g, wave, alpha, s);
  }
}

// ---- block forms: look up the blocks, write the out meta, run on the payloads
int cu_block(bbts_hip_ctx *ctx, const bbts_tensor_t *t, block_rec *r, cu_shape_t *sh) {
  int rc = lookup(ctx, t, r);
  if (rc) return rc;
  *sh = cu_meta(r->meta);
  BBTS_REQUIRE(sh->rank <= 4, BBTS_E_FORMAT, "cutensor rank above 4");
  uint64_t n = 1;
  for (uint32_t x = 0; x < sh->rank; ++x) n *= (uint64_t)std::max<int64_t>(sh->dims[x], 0);
  BBTS_REQUIRE(BBTS_META_BYTES + 4 * n <= r->bytes, BBTS_E_SHAPE, "cutensor block smaller than its shape");
  return BBTS_OK;
}

int write_out_meta(bbts_hip_ctx *ctx, bbts_tensor_t *out, const block_rec &ro, const cu_shape_t &so,
                   hipStream_t s) {
  uint64_t n = 1;
  for (uint32_t x = 0; x < so.rank; ++x) n *= (uint64_t)so.dims[x];
  BBTS_REQUIRE(BBTS_META_BYTES + 4 * n <= ro.bytes, BBTS_E_SHAPE,
               "output block of %zu bytes cannot hold %llu elements", ro.bytes, (unsigned long long)n);
  tensor_meta_t mo = ro.meta;
  set_cu_meta(mo, so);
  int rc = update_meta(ctx, out, mo);
  if (rc) return rc;
  BBTS_HIP_CHECK(hipMemcpyAsync(reinterpret_cast<char *>(out) + 4, mo._blob, sizeof(cu_shape_t), hipMemcpyHostToDevice,
                                s));
  BBTS_HIP_CHECK(hipStreamSynchronize(s));  // the host copy of the shape is a stack value
  return BBTS_OK;
}

}  // namespace
}  // namespace bbts_hip

using namespace bbts_hip;

extern "C" {

int bbts_hip_cu_ew_raw(bbts_hip_ctx *ctx, const int32_t *params, int nparams, int rank, const int64_t *dims,
                       const float *in, float *out, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  return run_ew(params, nparams, rank, dims, in, out, (hipStream_t)stream);
}

int bbts_hip_cu_ew(bbts_hip_ctx *ctx, const bbts_tensor_t *in, bbts_tensor_t *out, const int32_t *params, int nparams,
                   bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  block_rec ri, ro;
  cu_shape_t si, so;
  if ((rc = cu_block(ctx, in, &ri, &si)) || (rc = lookup(ctx, out, &ro))) return rc;
  ew_params p;
  uint64_t n;
  if ((rc = ew_shape(params, nparams, (int)si.rank, si.dims, &p, &so, &n))) return rc;
  if ((rc = write_out_meta(ctx, out, ro, so, (hipStream_t)stream))) return rc;
  return run_ew(params, nparams, (int)si.rank, si.dims, payload(in), payload(out), (hipStream_t)stream);
}

int bbts_hip_cu_ewb_raw(bbts_hip_ctx *ctx, const int32_t *params, int nparams, int rank_l, const int64_t *dims_l,
                        const float *lhs, int rank_r, const int64_t *dims_r, const float *rhs, float *out,
                        bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  return run_ewb(params, nparams, rank_l, dims_l, lhs, rank_r, dims_r, rhs, out, nullptr, (hipStream_t)stream);
}

int bbts_hip_cu_ewb(bbts_hip_ctx *ctx, const bbts_tensor_t *lhs, const bbts_tensor_t *rhs, bbts_tensor_t *out,
                    const int32_t *params, int nparams, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  block_rec rl, rr, ro;
  cu_shape_t sl, sr, so;
  if ((rc = cu_block(ctx, lhs, &rl, &sl)) || (rc = cu_block(ctx, rhs, &rr, &sr)) || (rc = lookup(ctx, out, &ro)))
    return rc;
  ewb_params p;
  int ro_rank;
  int64_t dout[4];
  if ((rc = parse_ewb(params, nparams, (int)sl.rank, sl.dims, (int)sr.rank, sr.dims, &p, &ro_rank, dout))) return rc;
  so = cu_shape_t{};
  so.rank = (uint32_t)ro_rank;
  for (int x = 0; x < ro_rank; ++x) so.dims[x] = dout[x];
  if ((rc = write_out_meta(ctx, out, ro, so, (hipStream_t)stream))) return rc;
  return run_ewb(params, nparams, (int)sl.rank, sl.dims, payload(lhs), (int)sr.rank, sr.dims, payload(rhs),
                 payload(out), nullptr, (hipStream_t)stream);
}

int bbts_hip_cu_ewb_castable_raw(bbts_hip_ctx *ctx, int op, const float *lhs, const float *rhs, float *out, size_t n,
                                 bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  return run_castable(op, lhs, rhs, out, n, (hipStream_t)stream);
}

int bbts_hip_cu_ewb_castable(bbts_hip_ctx *ctx, const bbts_tensor_t *lhs, const bbts_tensor_t *rhs,
                             bbts_tensor_t *out, const int32_t *params, int nparams, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  BBTS_REQUIRE(nparams >= 1, BBTS_E_INVALID, "ewb_castable: missing the operator");
  block_rec rl, rr, ro;
  cu_shape_t sl, sr;
  if ((rc = cu_block(ctx, lhs, &rl, &sl)) || (rc = cu_block(ctx, rhs, &rr, &sr)) || (rc = lookup(ctx, out, &ro)))
    return rc;
  uint64_t nl = 1, nr = 1;
  for (uint32_t x = 0; x < sl.rank; ++x) nl *= (uint64_t)sl.dims[x];
  for (uint32_t x = 0; x < sr.rank; ++x) nr *= (uint64_t)sr.dims[x];
  // the reference runs the op over lhs's element count and reads rhs as the same shape
  BBTS_REQUIRE(nl == nr, BBTS_E_SHAPE, "ewb_castable: %llu and %llu elements", (unsigned long long)nl,
               (unsigned long long)nr);
  BBTS_REQUIRE(params[0] >= 0 && params[0] <= 3, BBTS_E_INVALID, "ewb_castable: %d is not a castable operator",
               params[0]);
  if ((rc = write_out_meta(ctx, out, ro, sl, (hipStream_t)stream))) return rc;
  return run_castable(params[0], payload(lhs), payload(rhs), payload(out), nl, (hipStream_t)stream);
}

int bbts_hip_cu_reduction_raw(bbts_hip_ctx *ctx, const int32_t *params, int nparams, int rank, const int64_t *dims,
                              const float *in, float *out, int segments, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  return run_reduction(ctx, params, nparams, rank, dims, in, out, nullptr, segments, (hipStream_t)stream);
}

int bbts_hip_cu_reduction(bbts_hip_ctx *ctx, const bbts_tensor_t *in, bbts_tensor_t *out, const int32_t *params,
                          int nparams, bbts_stream_t stream) {
  int rc = bind(ctx);
  if (rc) return rc;
  block_rec ri, ro;
  cu_shape_t si;
  if ((rc = cu_block(ctx, in, &ri, &si)) || (rc = lookup(ctx, out, &ro))) return rc;
  BBTS_REQUIRE(nparams >= 2, BBTS_E_INVALID, "reduction: too few parameters");
  cu_shape_t so{};
  so.rank = (uint32_t)(nparams - 2);
  BBTS_REQUIRE(so.rank <= si.rank, BBTS_E_INVALID, "reduction: input rank is not great enough");
  for (uint32_t r = 0; r < so.rank; ++r) {
    const int m = params[2 + r];
    BBTS_REQUIRE(m >= 0 && m < (int)si.rank, BBTS_E_INVALID, "reduction: input rank is not great enough");
    so.dims[r] = si.dims[m];
  }
  if ((rc = write_out_meta(ctx, out, ro, so, (hipStream_t)stream))) return rc;
  return run_reduction(ctx, params, nparams, (int)si.rank, si.dims, payload(in), payload(out), nullptr, 0,
                       (hipStream_t)stream);
}

}  // extern "C"
