// Learned cost model and the exhaustive autotuner behind the grouped stream.
//
// Autotuner: for every distinct (M, N, K, dtype, precision) of a batch and
// every tile configuration, it times a grouped persistent launch of R copies
// of that GEMM (R chosen so the launch fills the chip several times over)
// with HIP events, and records cost = time / R -- the op's throughput cost
// inside a packed grouped launch, which is what the stream planner needs.
//
// Cost model: an MLP  nf -> H -> H -> 1  (ReLU) that predicts log(cost in
// microseconds) of (GEMM, tile config) from shape features.  Training runs
// on the GPU through this file's kernels: a batched linear forward, the two
// backward kernels (data and weight gradients), a mean-squared-error loss and
// an Adam update.  Inference (bbts_hip_costmodel_predict) runs the same
// forward kernels on device.  The planner calls it to pick each op's tile
// configuration and the largest-first launch order.
#include "common.h"

#include <cmath>
#include <random>

namespace bbts_hip {
namespace {

constexpr int kNF = 18;
constexpr int kMaxCfgs = 6;
constexpr int kThreads = 256;

// Y[s, o] = act(sum_i X[s, i] * W[o, i] + b[o]);  Z keeps the pre-activation.
__global__ __launch_bounds__(kThreads) void lin_fwd_kernel(const float *__restrict__ X, const float *__restrict__ W,
                                                           const float *__restrict__ b, float *Z, float *Y, int n,
                                                           int in, int out, int relu) {
  int idx = blockIdx.x * kThreads + threadIdx.x;
  if (idx >= n * out) return;
  int s = idx / out, o = idx % out;
  const float *x = X + (size_t)s * in;
  const float *w = W + (size_t)o * in;
  float acc = b[o];
  for (int i = 0; i < in; ++i) acc = fmaf(x[i], w[i], acc);
  if (Z) Z[idx] = acc;
  Y[idx] = relu ? (acc > 0.0f ? acc : 0.0f) : acc;
}

// G[s] = 2 (y[s] - t[s]) / n ;  loss partial sums into lossbuf (block-wise, no atomics)
__global__ __launch_bounds__(kThreads) void mse_kernel(const float *y, const float *t, float *G, float *partial, int n) {
  __shared__ float red[kThreads];
  int s = blockIdx.x * kThreads + threadIdx.x;
  float d = 0.0f;
  if (s < n) {
    d = y[s] - t[s];
    G[s] = 2.0f * d / (float)n;
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void sum_partials_kernel(const float *partial, int nb, int n, float *loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < nb; ++i) s += partial[i];
    *loss = s / (float)n;
  }
}

// Gp[s, i] = (sum_o G[s, o] * W[o, i]) * (Zp[s, i] > 0)
__global__ __launch_bounds__(kThreads) void lin_bwd_data_kernel(const float *__restrict__ G, const float *__restrict__ W,
                                                                const float *__restrict__ Zp, float *Gp, int n, int out,
                                                                int in) {
  int idx = blockIdx.x * kThreads + threadIdx.x;
  if (idx >= n * in) return;
  int s = idx / in, i = idx % in;
  const float *g = G + (size_t)s * out;
  float acc = 0.0f;
  for (int o = 0; o < out; ++o) acc = fmaf(g[o], W[(size_t)o * in + i], acc);
  Gp[idx] = Zp[idx] > 0.0f ? acc : 0.0f;
}

// dW[o, i] = sum_s G[s, o] * X[s, i];  db[o] = sum_s G[s, o]   (one lane per (o, i) / o; deterministic)
__global__ __launch_bounds__(kThreads) void lin_bwd_weight_kernel(const float *__restrict__ G,
                                                                  const float *__restrict__ X, float *dW, float *db,
                                                                  int n, int out, int in) {
  int idx = blockIdx.x * kThreads + threadIdx.x;
  int nw = out * in;
  if (idx < nw) {
    int o = idx / in, i = idx % in;
    float acc = 0.0f;
    for (int s = 0; s < n; ++s) acc = fmaf(G[(size_t)s * out + o], X[(size_t)s * in + i], acc);
    dW[idx] = acc;
  } else if (idx < nw + out) {
    int o = idx - nw;
    float acc = 0.0f;
    for (int s = 0; s < n; ++s) acc += G[(size_t)s * out + o];
    db[o] = acc;
  }
}

__global__ __launch_bounds__(kThreads) void adam_kernel(float *p, const float *g, float *m, float *v, int n, float lr,
                                                        float b1, float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

inline unsigned blocks(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace
}  // namespace bbts_hip

using namespace bbts_hip;

struct bbts_costmodel {
  bbts_hip_ctx *ctx = nullptr;
  int H = 64;
  int nparams = 0;
  // offsets into params
  int oW1, ob1, oW2, ob2, oW3, ob3;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  long long step = 0;
  // activations workspace
  float *ws = nullptr;
  int ws_n = 0;
  float *loss_dev = nullptr, *partial = nullptr;
  int partial_n = 0;
};

namespace {

int ensure_ws(bbts_costmodel *cm, int n) {
  if (n <= cm->ws_n) return BBTS_OK;
  if (cm->ws) BBTS_HIP_CHECK(hipFree(cm->ws));
  // Z1,H1,Z2,H2 (n*H each), y,G3 (n), G2,G1 (n*H)
  size_t bytes = sizeof(float) * ((size_t)n * (6 * (size_t)cm->H + 2));
  BBTS_HIP_CHECK(hipMalloc(&cm->ws, bytes));
  cm->ws_n = n;
  int pn = (n + kThreads - 1) / kThreads;
  if (pn > cm->partial_n) {
    if (cm->partial) BBTS_HIP_CHECK(hipFree(cm->partial));
    BBTS_HIP_CHECK(hipMalloc(&cm->partial, sizeof(float) * (size_t)pn));
    cm->partial_n = pn;
  }
  return BBTS_OK;
}

struct acts {
  float *Z1, *H1, *Z2, *H2, *y, *G3, *G2, *G1;
};
acts view(bbts_costmodel *cm, int n) {
  acts a;
  size_t nh = (size_t)n * cm->H;
  a.Z1 = cm->ws;
  a.H1 = a.Z1 + nh;
  a.Z2 = a.H1 + nh;
  a.H2 = a.Z2 + nh;
  a.G2 = a.H2 + nh;
  a.G1 = a.G2 + nh;
  a.y = a.G1 + nh;
  a.G3 = a.y + n;
  return a;
}

int forward(bbts_costmodel *cm, const float *X, int n, const acts &a, hipStream_t s) {
  const int H = cm->H;
  float *P = cm->params;
  hipLaunchKernelGGL(lin_fwd_kernel, dim3(blocks((long long)n * H)), dim3(kThreads), 0, s, X, P + cm->oW1,
                     P + cm->ob1, a.Z1, a.H1, n, kNF, H, 1);
  hipLaunchKernelGGL(lin_fwd_kernel, dim3(blocks((long long)n * H)), dim3(kThreads), 0, s, a.H1, P + cm->oW2,
                     P + cm->ob2, a.Z2, a.H2, n, H, H, 1);
  hipLaunchKernelGGL(lin_fwd_kernel, dim3(blocks(n)), dim3(kThreads), 0, s, a.H2, P + cm->oW3, P + cm->ob3,
                     (float *)nullptr, a.y, n, H, 1, 0);
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}

int backward(bbts_costmodel *cm, const float *X, const float *t, int n, const acts &a, float *grads, float *loss,
             hipStream_t s) {
  const int H = cm->H;
  float *P = cm->params;
  int nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(kThreads), 0, s, a.y, t, a.G3, cm->partial, n);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s, cm->partial, nb, n, loss);
  // layer 3: out=1, in=H
  hipLaunchKernelGGL(lin_bwd_weight_kernel, dim3(blocks(H + 1)), dim3(kThreads), 0, s, a.G3, a.H2, grads + cm->oW3,
                     grads + cm->ob3, n, 1, H);
  hipLaunchKernelGGL(lin_bwd_data_kernel, dim3(blocks((long long)n * H)), dim3(kThreads), 0, s, a.G3, P + cm->oW3,
                     a.Z2, a.G2, n, 1, H);
  // layer 2: out=H, in=H
  hipLaunchKernelGGL(lin_bwd_weight_kernel, dim3(blocks(H * H + H)), dim3(kThreads), 0, s, a.G2, a.H1,
                     grads + cm->oW2, grads + cm->ob2, n, H, H);
  hipLaunchKernelGGL(lin_bwd_data_kernel, dim3(blocks((long long)n * H)), dim3(kThreads), 0, s, a.G2, P + cm->oW2,
                     a.Z1, a.G1, n, H, H);
  // layer 1: out=H, in=nf
  hipLaunchKernelGGL(lin_bwd_weight_kernel, dim3(blocks(H * kNF + H)), dim3(kThreads), 0, s, a.G1, X, grads + cm->oW1,
                     grads + cm->ob1, n, H, kNF);
  BBTS_LAUNCH_CHECK();
  return BBTS_OK;
}

}  // namespace

namespace bbts_hip {
int costmodel_predict_host(bbts_costmodel *cm, const std::vector<float> &feats, int n, std::vector<float> &out) {
  int rc = bind(cm->ctx);
  if (rc) return rc;
  out.assign((size_t)n, 0.0f);
  if (n == 0) return BBTS_OK;
  float *X = nullptr, *y = nullptr;
  BBTS_HIP_CHECK(hipMalloc(&X, sizeof(float) * feats.size()));
  BBTS_HIP_CHECK(hipMalloc(&y, sizeof(float) * (size_t)n));
  hipError_t e = hipMemcpy(X, feats.data(), sizeof(float) * feats.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) rc = bbts_hip_costmodel_predict(cm, X, n, y, nullptr);
  if (e == hipSuccess && rc == BBTS_OK) e = hipMemcpy(out.data(), y, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(X);
  (void)hipFree(y);
  if (e != hipSuccess) {
    set_error("cost model transfer failed: %s", hipGetErrorString(e));
    return BBTS_E_HIP;
  }
  return rc;
}
}  // namespace bbts_hip

extern "C" {

int bbts_hip_costmodel_num_features(void) { return kNF; }

int bbts_hip_costmodel_features(int M, int N, int K, int dtype, int precision, int cfg, float *f) {
  BBTS_REQUIRE(f, BBTS_E_INVALID, "null feature buffer");
  int bm = 0, bn = 0, bk = 0;
  int rc = bbts_hip_tile_config_shape(cfg, &bm, &bn, &bk);
  if (rc) return rc;
  double tm = std::ceil((double)M / bm), tn = std::ceil((double)N / bn);
  double tiles = tm * tn;
  double eff = (tiles > 0) ? ((double)M * N) / (tiles * bm * bn) : 0.0;
  int slices = dtype == 1 ? 1 : (precision == BBTS_PREC_BF16X3 ? 2 : (precision == BBTS_PREC_BF16X1 ? 1 : 3));
  auto lg = [](double x) { return x > 0 ? std::log2(x) : 0.0; };
  f[0] = (float)(lg(M) / 12.0);
  f[1] = (float)(lg(N) / 12.0);
  f[2] = (float)(lg(K) / 12.0);
  f[3] = (float)(lg(tiles) / 12.0);
  f[4] = (float)eff;
  f[5] = (float)(lg(2.0 * M * N * (double)K) / 40.0);
  f[6] = (float)(slices / 3.0);
  f[7] = (float)dtype;
  for (int c = 0; c < kMaxCfgs; ++c) f[8 + c] = (c == cfg) ? 1.0f : 0.0f;
  f[14] = (float)(lg(bm) / 8.0);
  f[15] = (float)(lg(bn) / 8.0);
  f[16] = (float)(lg(tiles * bm * bn * (double)K * 2.0) / 40.0);  // padded flops
  f[17] = (float)(lg(tiles / 256.0 + 1.0) / 4.0);                  // waves of tiles over 256 CUs
  return BBTS_OK;
}

int bbts_hip_costmodel_create(bbts_hip_ctx *ctx, int hidden, uint64_t seed, bbts_costmodel **out) {
  int rc = bind(ctx);
  if (rc) return rc;
  BBTS_REQUIRE(out && hidden >= 1 && hidden <= 1024, BBTS_E_INVALID, "hidden size must be in [1,1024]");
  auto *cm = new bbts_costmodel();
  cm->ctx = ctx;
  cm->H = hidden;
  const int H = hidden;
  int off = 0;
  cm->oW1 = off; off += H * kNF;
  cm->ob1 = off; off += H;
  cm->oW2 = off; off += H * H;
  cm->ob2 = off; off += H;
  cm->oW3 = off; off += H;
  cm->ob3 = off; off += 1;
  cm->nparams = off;
  std::vector<float> init((size_t)off, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int i = 0; i < H * kNF; ++i) init[(size_t)(cm->oW1 + i)] = nd(rng) * std::sqrt(2.0f / kNF);
  for (int i = 0; i < H * H; ++i) init[(size_t)(cm->oW2 + i)] = nd(rng) * std::sqrt(2.0f / H);
  for (int i = 0; i < H; ++i) init[(size_t)(cm->oW3 + i)] = nd(rng) * std::sqrt(1.0f / H);
  init[(size_t)cm->ob3] = 3.0f;  // log(20 us): a neutral starting level
  size_t bytes = sizeof(float) * (size_t)off;
  hipError_t e = hipMalloc(&cm->params, bytes);
  if (e == hipSuccess) e = hipMalloc(&cm->grads, bytes);
  if (e == hipSuccess) e = hipMalloc(&cm->m, bytes);
  if (e == hipSuccess) e = hipMalloc(&cm->v, bytes);
  if (e == hipSuccess) e = hipMalloc(&cm->loss_dev, sizeof(float) * 4);
  if (e == hipSuccess) e = hipMemcpy(cm->params, init.data(), bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(cm->m, 0, bytes);
  if (e == hipSuccess) e = hipMemset(cm->v, 0, bytes);
  if (e != hipSuccess) {
    set_error("cost model allocation failed: %s", hipGetErrorString(e));
    bbts_hip_costmodel_destroy(cm);
    return BBTS_E_OOM;
  }
  *out = cm;
  return BBTS_OK;
}

int bbts_hip_costmodel_destroy(bbts_costmodel *cm) {
  if (!cm) return BBTS_OK;
  bind(cm->ctx);
  (void)hipDeviceSynchronize();
  for (float *p : {cm->params, cm->grads, cm->m, cm->v, cm->ws, cm->loss_dev, cm->partial})
    if (p) (void)hipFree(p);
  delete cm;
  return BBTS_OK;
}

int bbts_hip_costmodel_num_params(bbts_costmodel *cm) { return cm ? cm->nparams : 0; }

int bbts_hip_costmodel_get_params(bbts_costmodel *cm, float *host, int n) {
  BBTS_REQUIRE(cm && host && n == cm->nparams, BBTS_E_INVALID, "parameter count mismatch");
  int rc = bind(cm->ctx);
  if (rc) return rc;
  BBTS_HIP_CHECK(hipMemcpy(host, cm->params, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
  return BBTS_OK;
}

int bbts_hip_costmodel_set_params(bbts_costmodel *cm, const float *host, int n) {
  BBTS_REQUIRE(cm && host && n == cm->nparams, BBTS_E_INVALID, "parameter count mismatch");
  int rc = bind(cm->ctx);
  if (rc) return rc;
  BBTS_HIP_CHECK(hipMemcpy(cm->params, host, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
  cm->step = 0;
  size_t bytes = sizeof(float) * (size_t)n;
  BBTS_HIP_CHECK(hipMemset(cm->m, 0, bytes));
  BBTS_HIP_CHECK(hipMemset(cm->v, 0, bytes));
  return BBTS_OK;
}

int bbts_hip_costmodel_predict(bbts_costmodel *cm, const float *X, int n, float *y, bbts_stream_t stream) {
  BBTS_REQUIRE(cm && X && y && n >= 0, BBTS_E_INVALID, "bad predict arguments");
  int rc = bind(cm->ctx);
  if (rc) return rc;
  if (n == 0) return BBTS_OK;
  if ((rc = ensure_ws(cm, n))) return rc;
  acts a = view(cm, n);
  hipStream_t s = (hipStream_t)stream;
  if ((rc = forward(cm, X, n, a, s))) return rc;
  BBTS_HIP_CHECK(hipMemcpyAsync(y, a.y, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s));
  BBTS_HIP_CHECK(hipStreamSynchronize(s));
  return BBTS_OK;
}

int bbts_hip_costmodel_predict_host(bbts_costmodel *cm, const float *X_host, int n, float *y_host) {
  BBTS_REQUIRE(cm && X_host && y_host && n >= 0, BBTS_E_INVALID, "bad predict arguments");
  std::vector<float> feats(X_host, X_host + (size_t)n * kNF), out;
  int rc = costmodel_predict_host(cm, feats, n, out);
  if (rc) return rc;
  std::memcpy(y_host, out.data(), sizeof(float) * (size_t)n);
  return BBTS_OK;
}

int bbts_hip_costmodel_grad(bbts_costmodel *cm, const float *X, const float *y, int n, float *grads, float *loss,
                            bbts_stream_t stream) {
  BBTS_REQUIRE(cm && X && y && grads && loss && n >= 1, BBTS_E_INVALID, "bad grad arguments");
  int rc = bind(cm->ctx);
  if (rc) return rc;
  if ((rc = ensure_ws(cm, n))) return rc;
  acts a = view(cm, n);
  hipStream_t s = (hipStream_t)stream;
  if ((rc = forward(cm, X, n, a, s))) return rc;
  if ((rc = backward(cm, X, y, n, a, grads, loss, s))) return rc;
  BBTS_HIP_CHECK(hipStreamSynchronize(s));
  return BBTS_OK;
}

int bbts_hip_costmodel_train(bbts_costmodel *cm, const float *X, const float *y, int n, int batch, float lr,
                             int epochs, float *loss_out, bbts_stream_t stream) {
  BBTS_REQUIRE(cm && X && y && n >= 1 && batch >= 1 && epochs >= 1, BBTS_E_INVALID, "bad train arguments");
  int rc = bind(cm->ctx);
  if (rc) return rc;
  if (batch > n) batch = n;
  if ((rc = ensure_ws(cm, batch))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double loss_sum = 0.0;
  int loss_cnt = 0;
  std::vector<float> lossv;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int start = 0; start < n; start += batch) {
      int nb = std::min(batch, n - start);
      acts a = view(cm, nb);
      const float *Xb = X + (size_t)start * kNF;
      const float *yb = y + start;
      if ((rc = forward(cm, Xb, nb, a, s))) return rc;
      if ((rc = backward(cm, Xb, yb, nb, a, cm->grads, cm->loss_dev, s))) return rc;
      cm->step += 1;
      float c1 = 1.0f - std::pow(b1, (float)cm->step), c2 = 1.0f - std::pow(b2, (float)cm->step);
      hipLaunchKernelGGL(adam_kernel, dim3(blocks(cm->nparams)), dim3(kThreads), 0, s, cm->params, cm->grads, cm->m,
                         cm->v, cm->nparams, lr, b1, b2, eps, c1, c2);
      BBTS_LAUNCH_CHECK();
      if (ep == epochs - 1) {
        float l = 0.0f;
        BBTS_HIP_CHECK(hipMemcpyAsync(&l, cm->loss_dev, sizeof(float), hipMemcpyDeviceToHost, s));
        BBTS_HIP_CHECK(hipStreamSynchronize(s));
        loss_sum += l;
        ++loss_cnt;
      }
    }
  }
  BBTS_HIP_CHECK(hipStreamSynchronize(s));
  if (loss_out) *loss_out = loss_cnt ? (float)(loss_sum / loss_cnt) : 0.0f;
  return BBTS_OK;
}

}  // extern "C"
