// Learned cost model: an fp32 MLP 32 -> 64 -> 64 -> 1 (ReLU) that predicts the
// natural log of a plan's measured milliseconds, with hand-written gfx950
// kernels for inference / forward, backward and Adam.  The reference has an
// analytic cost model only (src/gpudb/CostModel.cu); this one is trained on
// (plan features, measured time) pairs the engine's own kernels produce, and
// is used only where it has proven more accurate than the analytic model on
// fresh measurements (include/mordred_costmodel.h).
//
// Parameter layout (one float array): W1[32][64], b1[64], W2[64][64], b2[64],
// W3[64], b3.  Loss: mean over the minibatch of (y - t)^2 on log-ms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mordred_costmodel.h"

namespace mordred_cm {

constexpr int kIn = MORDRED_COST_FEATURES;  // 32
constexpr int kH = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kParams = kOffB3 + 1;  // 6337
constexpr int kBlock = 256;

// Forward pass, one wavefront per sample: lane o computes hidden unit o of
// both layers (kH = 64 = the wave), the inputs and the first layer's outputs
// reach every lane by shuffles, the output is a wave reduction.  The lane's
// column of W1 sits in registers; W2 is read row by row (each row one
// coalesced 256-byte load, L2-resident).  A few samples -- a launch-size or
// placement decision -- finish in one short dependent chain instead of one
// thread walking all 6,337 parameters.  Writes the prediction and, when
// H1/H2 are given, the post-ReLU activations.
static_assert(kH == 64, "one hidden unit per lane of a wave64");
__global__ __launch_bounds__(kBlock) void mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                             int n, float* __restrict__ Y, float* __restrict__ H1,
                                                             float* __restrict__ H2) {
  const int lane = threadIdx.x & (kH - 1);
  const int nw = gridDim.x * (kBlock / kH);
  int b = (int)(blockIdx.x * kBlock + threadIdx.x) / kH;
  if (b >= n) return;
  float w1[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) w1[i] = P[kOffW1 + i * kH + lane];
  const float b1 = P[kOffB1 + lane], b2 = P[kOffB2 + lane], w3 = P[kOffW3 + lane], b3 = P[kOffB3];
  for (; b < n; b += nw) {
    const float xv = lane < kIn ? X[(size_t)b * kIn + lane] : 0.f;
    float a = b1;
#pragma unroll
    for (int i = 0; i < kIn; i++) a = fmaf(__shfl(xv, i, kH), w1[i], a);
    const float h1 = fmaxf(a, 0.f);
    float c = b2;
#pragma unroll 16
    for (int i = 0; i < kH; i++) c = fmaf(__shfl(h1, i, kH), P[kOffW2 + i * kH + lane], c);
    const float h2 = fmaxf(c, 0.f);
    float y = h2 * w3;
#pragma unroll
    for (int d = kH / 2; d > 0; d >>= 1) y += __shfl_xor(y, d, kH);
    if (H1) H1[(size_t)b * kH + lane] = h1;
    if (H2) H2[(size_t)b * kH + lane] = h2;
    if (lane == 0) Y[b] = y + b3;
  }
}

// Workgroups of the forward pass for n samples (one wave per sample).
static int forward_grid(int n) { return std::max(1, (n + kBlock / kH - 1) / (kBlock / kH)); }

// Backward deltas, one wavefront per sample: lane o holds d2[o] = dL/dz2[o],
// lane i sums d1[i] = sum_o d2[o] W2[i][o] over shuffles of d2, reading W2
// transposed from LDS (consecutive lanes, consecutive words).
__global__ __launch_bounds__(kBlock) void mlp_delta_kernel(const float* __restrict__ P, const float* __restrict__ Y,
                                                           const float* __restrict__ T, const float* __restrict__ H1,
                                                           const float* __restrict__ H2, int n, float* __restrict__ D3,
                                                           float* __restrict__ D2, float* __restrict__ D1) {
  __shared__ float w2t[kH * kH];  // w2t[o * kH + i] = W2[i][o]
  for (int k = threadIdx.x; k < kH * kH; k += kBlock) w2t[(k % kH) * kH + k / kH] = P[kOffW2 + k];
  __syncthreads();
  const int lane = threadIdx.x & (kH - 1);
  const float w3 = P[kOffW3 + lane];
  const int nw = gridDim.x * (kBlock / kH);
  // (b is the same for every lane of a wave: the shuffles see the whole wave)
  for (int b = (int)(blockIdx.x * kBlock + threadIdx.x) / kH; b < n; b += nw) {
    const float d3 = 2.f * (Y[b] - T[b]) / (float)n;
    const float d2 = H2[(size_t)b * kH + lane] > 0.f ? d3 * w3 : 0.f;
    D2[(size_t)b * kH + lane] = d2;
    float a = 0.f;
#pragma unroll 16
    for (int o = 0; o < kH; o++) a = fmaf(__shfl(d2, o, kH), w2t[o * kH + lane], a);
    D1[(size_t)b * kH + lane] = H1[(size_t)b * kH + lane] > 0.f ? a : 0.f;
    if (lane == 0) D3[b] = d3;
  }
}

// Parameter gradients, one workgroup per parameter row: W1 rows (kIn), b1, W2
// rows (kH), b2, W3, b3 -- kGradRows rows of at most kH parameters, lane o
// holding column o.  A row's gradient is sum_b a_b * d_b[o] over the samples;
// the four waves take every fourth sample, then add their partial sums in LDS.
constexpr int kGradRows = kIn + 1 + kH + 1 + 1 + 1;
__global__ __launch_bounds__(kBlock) void mlp_grad_kernel(const float* __restrict__ X, const float* __restrict__ H1,
                                                          const float* __restrict__ H2, const float* __restrict__ D1,
                                                          const float* __restrict__ D2, const float* __restrict__ D3,
                                                          int n, float* __restrict__ G) {
  constexpr int kWaves = kBlock / kH;
  __shared__ float part[kWaves][kH];
  const int row = blockIdx.x, lane = threadIdx.x & (kH - 1), wave = threadIdx.x / kH;
  // a_b: the row's input activation of sample b (1 for a bias row); d: the
  // deltas (or, for W3, the activations H2 against the scalar delta D3)
  const float* A = nullptr;
  int a_stride = 0, a_col = 0;
  const float* Dm = nullptr;
  int out = 0;
  if (row < kIn) {
    A = X, a_stride = kIn, a_col = row, Dm = D1, out = kOffW1 + row * kH;
  } else if (row == kIn) {
    Dm = D1, out = kOffB1;
  } else if (row < kIn + 1 + kH) {
    A = H1, a_stride = kH, a_col = row - kIn - 1, Dm = D2, out = kOffW2 + (row - kIn - 1) * kH;
  } else if (row == kIn + 1 + kH) {
    Dm = D2, out = kOffB2;
  } else if (row == kIn + 2 + kH) {
    A = D3, a_stride = 1, a_col = 0, Dm = H2, out = kOffW3;
  } else {
    out = kOffB3;  // b3: sum_b D3[b], lane 0
  }
  float g0 = 0.f, g1 = 0.f;
  int b = wave;
  for (; b + kWaves < n; b += 2 * kWaves) {
    const float a0 = A ? A[(size_t)b * a_stride + a_col] : 1.f;
    const float a1 = A ? A[(size_t)(b + kWaves) * a_stride + a_col] : 1.f;
    const float d0 = Dm ? Dm[(size_t)b * kH + lane] : (lane == 0 ? D3[b] : 0.f);
    const float d1 = Dm ? Dm[(size_t)(b + kWaves) * kH + lane] : (lane == 0 ? D3[b + kWaves] : 0.f);
    g0 = fmaf(a0, d0, g0);
    g1 = fmaf(a1, d1, g1);
  }
  if (b < n) g0 = fmaf(A ? A[(size_t)b * a_stride + a_col] : 1.f, Dm ? Dm[(size_t)b * kH + lane] : (lane == 0 ? D3[b] : 0.f), g0);
  part[wave][lane] = g0 + g1;
  __syncthreads();
  if (wave == 0) {
    float g = 0.f;
#pragma unroll
    for (int w = 0; w < kWaves; w++) g += part[w][lane];
    if (out == kOffB3) {
      if (lane == 0) G[out] = g;
    } else {
      G[out + lane] = g;
    }
  }
}

// Adam (Kingma & Ba, 2015), bias-corrected, one thread per parameter.
__global__ __launch_bounds__(kBlock) void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                          float* __restrict__ M, float* __restrict__ V, float lr,
                                                          float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= kParams) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace mordred_cm

using namespace mordred_cm;

struct mordred_cost_model {
  int device = 0;
  hipStream_t s = nullptr;
  float* d_p = nullptr;   // parameters
  float* d_g = nullptr;   // gradients
  float* d_m = nullptr;   // Adam first moments
  float* d_v = nullptr;   // Adam second moments
  float* d_work = nullptr;
  size_t work_floats = 0;
  float* h_stage = nullptr;  // pinned staging
  size_t stage_floats = 0;
  long long adam_t = 0;
  std::mt19937_64 rng;
  std::vector<float> xs, ts;  // replay buffer (features, log-ms)
  int capacity = 4096;
  long long seen = 0;
  double ewma_learned = -1, ewma_analytic = -1;
  int scored = 0;
  std::string err;
};

namespace {

int cm_fail(mordred_cost_model* m, hipError_t rc, const char* what) {
  m->err = std::string(what) + ": " + hipGetErrorString(rc);
  return (int)rc;
}

#define CM_OK(call)                                       \
  do {                                                    \
    hipError_t rc_ = (call);                              \
    if (rc_ != hipSuccess) return cm_fail(m, rc_, #call); \
  } while (0)

int reserve_work(mordred_cost_model* m, size_t floats) {
  if (floats <= m->work_floats) return 0;
  if (m->d_work) (void)hipFree(m->d_work);
  m->d_work = nullptr;
  m->work_floats = 0;
  CM_OK(hipMalloc(&m->d_work, floats * sizeof(float)));
  m->work_floats = floats;
  return 0;
}

int reserve_stage(mordred_cost_model* m, size_t floats) {
  if (floats <= m->stage_floats) return 0;
  if (m->h_stage) (void)hipHostFree(m->h_stage);
  m->h_stage = nullptr;
  m->stage_floats = 0;
  CM_OK(hipHostMalloc(reinterpret_cast<void**>(&m->h_stage), floats * sizeof(float), hipHostMallocDefault));
  m->stage_floats = floats;
  return 0;
}

// forward of n samples staged at X (device); y to host
int forward(mordred_cost_model* m, const float* x_host, int n, float* y_host) {
  const size_t need = (size_t)n * kIn + n;
  int rc = reserve_work(m, need);
  if (rc) return rc;
  if ((rc = reserve_stage(m, need))) return rc;
  float* dX = m->d_work;
  float* dY = dX + (size_t)n * kIn;
  std::memcpy(m->h_stage, x_host, (size_t)n * kIn * sizeof(float));
  CM_OK(hipMemcpyAsync(dX, m->h_stage, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, m->s));
  mlp_forward_kernel<<<forward_grid(n), kBlock, 0, m->s>>>(m->d_p, dX, n, dY, nullptr, nullptr);
  CM_OK(hipGetLastError());
  CM_OK(hipMemcpyAsync(m->h_stage, dY, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
  CM_OK(hipStreamSynchronize(m->s));
  std::memcpy(y_host, m->h_stage, n * sizeof(float));
  return 0;
}

// Forward, deltas and parameter gradients of the mean squared loss over n
// samples already on the device (dX: n x kIn, dT: n) into d_g; the
// activations and deltas go to the work area's first n * (5 kH + 2) floats,
// the prediction to *dY_out.
int gradients_dev(mordred_cost_model* m, const float* dX, const float* dT, int n, float** dY_out) {
  float* dY = m->d_work;
  float* H1 = dY + n;
  float* H2 = H1 + (size_t)n * kH;
  float* D1 = H2 + (size_t)n * kH;
  float* D2 = D1 + (size_t)n * kH;
  float* D3 = D2 + (size_t)n * kH;
  mlp_forward_kernel<<<forward_grid(n), kBlock, 0, m->s>>>(m->d_p, dX, n, dY, H1, H2);
  mlp_delta_kernel<<<forward_grid(n), kBlock, 0, m->s>>>(m->d_p, dY, dT, H1, H2, n, D3, D2, D1);
  mlp_grad_kernel<<<kGradRows, kBlock, 0, m->s>>>(dX, H1, H2, D1, D2, D3, n, m->d_g);
  CM_OK(hipGetLastError());
  if (dY_out) *dY_out = dY;
  return 0;
}
size_t activation_floats(int n) { return (size_t)n * (4 * kH + 2); }

// mean squared loss of the n predictions at dY against t_host (synchronizes)
int loss_of(mordred_cost_model* m, const float* dY, const float* t_host, int n, float* loss) {
  CM_OK(hipMemcpyAsync(m->h_stage, dY, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
  CM_OK(hipStreamSynchronize(m->s));
  double l = 0;
  for (int b = 0; b < n; b++) {
    const double d = (double)m->h_stage[b] - (double)t_host[b];
    l += d * d;
  }
  *loss = (float)(l / n);
  return 0;
}

// gradients of the mean squared loss over n samples (X, T on host); leaves
// them in d_g and the loss in *loss
int gradients(mordred_cost_model* m, const float* x_host, const float* t_host, int n, float* loss) {
  const size_t act = activation_floats(n);
  int rc = reserve_work(m, act + (size_t)n * (kIn + 1));
  if (rc) return rc;
  if ((rc = reserve_stage(m, (size_t)n * (kIn + 1)))) return rc;
  float* dX = m->d_work + act;
  float* dT = dX + (size_t)n * kIn;
  std::memcpy(m->h_stage, x_host, (size_t)n * kIn * sizeof(float));
  std::memcpy(m->h_stage + (size_t)n * kIn, t_host, n * sizeof(float));
  CM_OK(hipMemcpyAsync(dX, m->h_stage, (size_t)n * (kIn + 1) * sizeof(float), hipMemcpyHostToDevice, m->s));
  float* dY = nullptr;
  if ((rc = gradients_dev(m, dX, dT, n, &dY))) return rc;
  return loss ? loss_of(m, dY, t_host, n, loss) : 0;
}

int adam_step(mordred_cost_model* m, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m->adam_t++;
  const float c1 = 1.f - (float)std::pow((double)b1, (double)m->adam_t);
  const float c2 = 1.f - (float)std::pow((double)b2, (double)m->adam_t);
  mlp_adam_kernel<<<(kParams + kBlock - 1) / kBlock, kBlock, 0, m->s>>>(m->d_p, m->d_g, m->d_m, m->d_v, lr, b1, b2,
                                                                       eps, c1, c2);
  CM_OK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

mordred_cost_model* mordred_cost_model_create(int device, unsigned long long seed) {
  auto* m = new mordred_cost_model();
  m->device = device;
  m->rng.seed(seed);
  bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreate(&m->s) == hipSuccess;
  for (float** p : {&m->d_p, &m->d_g, &m->d_m, &m->d_v})
    ok = ok && hipMalloc(p, kParams * sizeof(float)) == hipSuccess;
  if (ok) {
    // He initialisation of the weights, zero biases
    std::vector<float> h(kParams, 0.f);
    std::normal_distribution<float> n1(0.f, std::sqrt(2.f / kIn)), n2(0.f, std::sqrt(2.f / kH));
    for (int i = 0; i < kIn * kH; i++) h[kOffW1 + i] = n1(m->rng);
    for (int i = 0; i < kH * kH; i++) h[kOffW2 + i] = n2(m->rng);
    for (int i = 0; i < kH; i++) h[kOffW3 + i] = n2(m->rng) * 0.5f;
    ok = hipMemcpy(m->d_p, h.data(), kParams * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(m->d_m, 0, kParams * sizeof(float)) == hipSuccess &&
         hipMemset(m->d_v, 0, kParams * sizeof(float)) == hipSuccess;
  }
  if (!ok) {
    mordred_cost_model_destroy(m);
    return nullptr;
  }
  return m;
}

void mordred_cost_model_destroy(mordred_cost_model* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  for (float* p : {m->d_p, m->d_g, m->d_m, m->d_v, m->d_work})
    if (p) (void)hipFree(p);
  if (m->h_stage) (void)hipHostFree(m->h_stage);
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
}

const char* mordred_cost_model_last_error(const mordred_cost_model* m) { return m ? m->err.c_str() : "no model"; }

int mordred_cost_model_num_params(void) { return kParams; }

int mordred_cost_model_get_params(mordred_cost_model* m, float* out) {
  if (!m || !out) return -1;
  CM_OK(hipSetDevice(m->device));
  CM_OK(hipMemcpy(out, m->d_p, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mordred_cost_model_set_params(mordred_cost_model* m, const float* in) {
  if (!m || !in) return -1;
  CM_OK(hipSetDevice(m->device));
  CM_OK(hipMemcpy(m->d_p, in, kParams * sizeof(float), hipMemcpyHostToDevice));
  CM_OK(hipMemset(m->d_m, 0, kParams * sizeof(float)));
  CM_OK(hipMemset(m->d_v, 0, kParams * sizeof(float)));
  m->adam_t = 0;
  return 0;
}

int mordred_cost_model_predict(mordred_cost_model* m, const float* x, int n, float* y) {
  if (!m || !x || !y || n < 0) return -1;
  if (n == 0) return 0;
  CM_OK(hipSetDevice(m->device));
  return forward(m, x, n, y);
}

int mordred_cost_model_gradients(mordred_cost_model* m, const float* x, const float* t, int n, float* grads,
                                 float* loss) {
  if (!m || !x || !t || n <= 0) return -1;
  CM_OK(hipSetDevice(m->device));
  int rc = gradients(m, x, t, n, loss);
  if (rc) return rc;
  if (grads) {
    CM_OK(hipMemcpyAsync(grads, m->d_g, kParams * sizeof(float), hipMemcpyDeviceToHost, m->s));
    CM_OK(hipStreamSynchronize(m->s));
  }
  return 0;
}

int mordred_cost_model_step(mordred_cost_model* m, const float* x, const float* t, int n, float lr, float* loss) {
  if (!m || !x || !t || n <= 0 || !(lr > 0)) return -1;
  CM_OK(hipSetDevice(m->device));
  int rc = gradients(m, x, t, n, loss);
  if (rc) return rc;
  if ((rc = adam_step(m, lr))) return rc;
  CM_OK(hipStreamSynchronize(m->s));
  return 0;
}

int mordred_cost_model_observe(mordred_cost_model* m, const float* x, double measured_ms, double analytic_ms) {
  if (!m || !x || !(measured_ms > 0)) return -1;
  const float t = (float)std::log(measured_ms);
  // score the model on the fresh measurement before it is trained on it
  if (m->seen > 0) {
    float y = 0;
    int rc = mordred_cost_model_predict(m, x, 1, &y);
    if (rc) return rc;
    const double el = std::fabs((double)y - t);
    m->ewma_learned = m->ewma_learned < 0 ? el : 0.9 * m->ewma_learned + 0.1 * el;
    m->scored++;
  }
  if (analytic_ms > 0) {
    const double ea = std::fabs(std::log(analytic_ms) - t);
    m->ewma_analytic = m->ewma_analytic < 0 ? ea : 0.9 * m->ewma_analytic + 0.1 * ea;
  }
  const int slot = (int)(m->seen % m->capacity);
  if ((int)m->ts.size() < m->capacity) {
    m->xs.insert(m->xs.end(), x, x + kIn);
    m->ts.push_back(t);
  } else {
    std::memcpy(&m->xs[(size_t)slot * kIn], x, kIn * sizeof(float));
    m->ts[slot] = t;
  }
  m->seen++;
  return 0;
}

int mordred_cost_model_train(mordred_cost_model* m, int steps, int batch, float lr, float* last_loss) {
  if (!m || steps < 0 || batch <= 0 || !(lr > 0)) return -1;
  const int have = (int)m->ts.size();
  if (have == 0 || steps == 0) return 0;
  CM_OK(hipSetDevice(m->device));
  const int n = std::min(batch, have);
  // every step's minibatch staged and copied at once; the steps run back to
  // back on the model's stream with one synchronization at the end (each
  // step's kernels follow the previous step's Adam update on the stream)
  const size_t per = (size_t)n * (kIn + 1), act = activation_floats(n);
  int rc = reserve_work(m, act + per * steps);
  if (rc) return rc;
  if ((rc = reserve_stage(m, std::max(per * steps, (size_t)n)))) return rc;
  std::uniform_int_distribution<int> pick(0, have - 1);
  std::vector<float> t_last(n);
  for (int s = 0; s < steps; s++) {
    float* hx = m->h_stage + per * s;
    float* ht = hx + (size_t)n * kIn;
    for (int b = 0; b < n; b++) {
      const int i = n == have ? b : pick(m->rng);
      std::memcpy(&hx[(size_t)b * kIn], &m->xs[(size_t)i * kIn], kIn * sizeof(float));
      ht[b] = m->ts[i];
    }
    if (s + 1 == steps) std::memcpy(t_last.data(), ht, n * sizeof(float));
  }
  float* dIn = m->d_work + act;
  CM_OK(hipMemcpyAsync(dIn, m->h_stage, per * steps * sizeof(float), hipMemcpyHostToDevice, m->s));
  float* dY = nullptr;
  for (int s = 0; s < steps; s++) {
    const float* dX = dIn + per * s;
    if ((rc = gradients_dev(m, dX, dX + (size_t)n * kIn, n, &dY))) return rc;
    if ((rc = adam_step(m, lr))) return rc;
  }
  // (the staging buffer is free again once the copy above has run: the loss
  // copy below is ordered after it on the stream)
  float loss = 0;
  if ((rc = loss_of(m, dY, t_last.data(), n, &loss))) return rc;
  if (last_loss) *last_loss = loss;
  return 0;
}

int mordred_cost_model_get_stats(const mordred_cost_model* m, mordred_cost_model_stats* st) {
  if (!m || !st) return -1;
  st->samples = m->seen;
  st->scored = m->scored;
  st->ewma_abs_log_error = m->ewma_learned;
  st->ewma_abs_log_error_analytic = m->ewma_analytic;
  st->trusted = m->scored >= 16 && m->ewma_learned >= 0 &&
                (m->ewma_analytic < 0 || m->ewma_learned < m->ewma_analytic) && m->ewma_learned < 0.5;
  return 0;
}

}  // extern "C"
