// Learned push/pull scheduler: a 3-layer FP32 MLP (IN -> 64 -> 64 -> OUT,
// ReLU) whose forward, backward and Adam update are this file's HIP kernels.
// One wave handles one sample with lane k = hidden unit k; gradients are
// reduced per block and then across blocks in a fixed order, so training is
// deterministic for a given input.
//
// Targets are log2 milliseconds of each scheduling option measured by the
// plain method; the loss is the masked mean squared error
//   L = sum_{n,o} M[n,o] (y[n,o] - Y[n,o])^2 / sum M.
#include <cmath>
#include <cstring>

#include "vs_sched.hh"

struct vs_sched_s {
  int refs = 1;  // the caller's handle + one per graph it is attached to
  int device = 0;
  float *d_params = nullptr;
  float *d_m = nullptr;
  float *d_v = nullptr;
  float *d_grad = nullptr;
  float *d_partial = nullptr;
  float *d_loss = nullptr;
  uint32_t step = 0;
  uint32_t grid = 64;
  hipStream_t stream = nullptr;
  // decision domain (host copy of what follows the parameters on the device)
  std::vector<float> domain;
  float radius = 0.f;
};

namespace vs {

constexpr int kMlpBlocks = 64;

__global__ __launch_bounds__(kBlock) void k_mlp_predict(const float *P, const float *X,
                                                       uint32_t n, float *Y) {
  __shared__ float s_h[kWavesPerBlock][2 * kHid];
  __shared__ float s_x[kWavesPerBlock][kIn];
  const int lane = lane_id(), wid = wave_id();
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  for (uint32_t s = blockIdx.x * kWavesPerBlock + wid; s < n; s += nwaves) {
    if (lane < kIn) s_x[wid][lane] = X[static_cast<size_t>(s) * kIn + lane];
    wave_sync();
    float h1, h2, y;
    mlp_forward_wave(P, s_x[wid], s_h[wid], h1, h2, y);
    if (lane < kOut) Y[static_cast<size_t>(s) * kOut + lane] = y;
    wave_sync();
  }
}

// Forward + backward over samples; per-block partial gradients.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float *P, const float *X,
                                                    const float *Y, const float *M,
                                                    uint32_t n, float inv_cnt,
                                                    float *partial, float *loss_partial) {
  __shared__ float s_h[kWavesPerBlock][2 * kHid];
  __shared__ float s_x[kWavesPerBlock][kIn];
  __shared__ float s_d[kWavesPerBlock][kOut];
  __shared__ float s_dh2[kWavesPerBlock][kHid];
  __shared__ float s_red[kNumParams];
  __shared__ float s_loss[kWavesPerBlock];
  const int lane = lane_id(), wid = wave_id();
  const int k = lane;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  float gW1[kIn], gW2[kHid], gW3c[kOut];
#pragma unroll
  for (int i = 0; i < kIn; ++i) gW1[i] = 0.f;
  for (int m = 0; m < kHid; ++m) gW2[m] = 0.f;
#pragma unroll
  for (int o = 0; o < kOut; ++o) gW3c[o] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
  for (uint32_t s = blockIdx.x * kWavesPerBlock + wid; s < n; s += nwaves) {
    if (lane < kIn) s_x[wid][lane] = X[static_cast<size_t>(s) * kIn + lane];
    wave_sync();
    float h1, h2, y;
    mlp_forward_wave(P, s_x[wid], s_h[wid], h1, h2, y);
    float d = 0.f;
    if (lane < kOut) {
      const float mk = M[static_cast<size_t>(s) * kOut + lane];
      const float diff = y - Y[static_cast<size_t>(s) * kOut + lane];
      d = 2.f * mk * diff * inv_cnt;
      loss += mk * diff * diff * inv_cnt;
      s_d[wid][lane] = d;
      gb3 += d;
    }
    wave_sync();
#pragma unroll
    for (int o = 0; o < kOut; ++o) gW3c[o] += s_d[wid][o] * h2;
    float dh2 = 0.f;
    if (h2 > 0.f) {
#pragma unroll
      for (int o = 0; o < kOut; ++o) dh2 += P[kOffW3 + o * kHid + k] * s_d[wid][o];
    }
    s_dh2[wid][k] = dh2;
    gb2 += dh2;
    for (int m = 0; m < kHid; ++m) gW2[m] += dh2 * s_h[wid][m];
    wave_sync();
    float dh1 = 0.f;
    if (h1 > 0.f) {
      for (int m = 0; m < kHid; ++m) dh1 += P[kOffW2 + m * kHid + k] * s_dh2[wid][m];
    }
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < kIn; ++i) gW1[i] += dh1 * s_x[wid][i];
    wave_sync();
  }
  // block reduction in wave order (deterministic)
  for (int w = 0; w < kWavesPerBlock; ++w) {
    if (wid == w) {
      const bool first = w == 0;
#pragma unroll
      for (int i = 0; i < kIn; ++i) {
        float &r = s_red[kOffW1 + k * kIn + i];
        r = first ? gW1[i] : r + gW1[i];
      }
      { float &r = s_red[kOffB1 + k]; r = first ? gb1 : r + gb1; }
      for (int m = 0; m < kHid; ++m) {
        float &r = s_red[kOffW2 + k * kHid + m];
        r = first ? gW2[m] : r + gW2[m];
      }
      { float &r = s_red[kOffB2 + k]; r = first ? gb2 : r + gb2; }
#pragma unroll
      for (int o = 0; o < kOut; ++o) {
        float &r = s_red[kOffW3 + o * kHid + k];
        r = first ? gW3c[o] : r + gW3c[o];
      }
      if (k < kOut) { float &r = s_red[kOffB3 + k]; r = first ? gb3 : r + gb3; }
      // loss: lane sum within the wave in fixed order
      float l = loss;
      for (int m = 1; m < 64; m <<= 1) l += __shfl_xor(l, m);
      if (lane == 0) s_loss[w] = l;
    }
    __syncthreads();
  }
  for (size_t i = threadIdx.x; i < kNumParams; i += blockDim.x)
    partial[blockIdx.x * kNumParams + i] = s_red[i];
  if (threadIdx.x == 0) {
    float l = 0.f;
    for (int w = 0; w < kWavesPerBlock; ++w) l += s_loss[w];
    loss_partial[blockIdx.x] = l;
  }
}

__global__ void k_mlp_reduce(const float *partial, const float *loss_partial,
                             uint32_t nblocks, float *grad, float *loss) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < kNumParams) {
    float acc = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) acc += partial[b * kNumParams + i];
    grad[i] = acc;
  }
  if (i == 0) {
    float l = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) l += loss_partial[b];
    *loss = l;
  }
}

__global__ void k_mlp_adam(float *P, float *Mm, float *Vv, const float *grad, float lr,
                           float beta1, float beta2, float eps, float bc1, float bc2) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= kNumParams) return;
  const float g = grad[i];
  const float m = beta1 * Mm[i] + (1.f - beta1) * g;
  const float v = beta2 * Vv[i] + (1.f - beta2) * g * g;
  Mm[i] = m;
  Vv[i] = v;
  P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

void sched_features(const SchedFeatures &f, float *x) {
  const double rows = f.rows ? static_cast<double>(f.rows) : 1.0;
  x[0] = static_cast<float>(std::log2(1.0 + f.pop) / 32.0);
  x[1] = static_cast<float>(std::log2(1.0 + f.act) / 32.0);
  x[2] = static_cast<float>(std::log2(1.0 + f.actdeg) / 40.0);
  x[3] = static_cast<float>(f.hop / 8.0);
  x[4] = static_cast<float>(static_cast<double>(f.pop) / (rows * (f.V ? f.V : 1)));
  x[5] = static_cast<float>(std::log2(1.0 + f.E) / 40.0);
  x[6] = static_cast<float>(std::log2(1.0 + f.V) / 32.0);
  x[7] = static_cast<float>(f.vis_frac);
}

static vs_status sched_predict_dev(vs_sched_s *s, const float *d_x, uint32_t n,
                                   float *d_y) {
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 3) / 4, 1024));
  hipLaunchKernelGGL(k_mlp_predict, dim3(grid), dim3(kBlock), 0, s->stream,
                     s->d_params, d_x, n, d_y);
  VS_HIP(hipGetLastError());
  return VS_OK;
}

const float *sched_params(vs_sched_s *s) { return s ? s->d_params : nullptr; }

void query_features(const QueryFeatures &f, float *x) {
  x[0] = static_cast<float>(std::log2(1.0 + f.S) / 32.0);
  x[1] = static_cast<float>(std::log2(1.0 + f.vol2) / 48.0);
  x[2] = static_cast<float>(std::log2(1.0 + f.degsum) / 40.0);
  x[3] = 0.f;
  x[4] = 0.f;
  x[5] = static_cast<float>(std::log2(1.0 + f.E) / 40.0);
  x[6] = static_cast<float>(std::log2(1.0 + f.V) / 32.0);
  x[7] = static_cast<float>(f.k_max / 8.0);
}

// Whether x lies within the scheduler's decision domain (no domain: yes).
static bool in_domain(const vs_sched_s *s, const float *x) {
  const size_t n = s->domain.size() / kIn;
  if (n == 0) return true;
  const float r2 = s->radius * s->radius;
  for (size_t i = 0; i < n; ++i) {
    float d = 0.f;
    for (int k = 0; k < kIn; ++k) {
      const float t = s->domain[i * kIn + k] - x[k];
      d += t * t;
    }
    if (d <= r2) return true;
  }
  return false;
}

// One forward pass on the scheduler's device; false if anything fails.
static bool predict_one(vs_sched_s *s, const float *x, float *y) {
  if (hipSetDevice(s->device) != hipSuccess) { (void)hipGetLastError(); return false; }
  // scratch: 2 * 64 floats after the loss slot
  float *d_x = s->d_loss + 1, *d_y = s->d_loss + 1 + kIn;
  if (hipMemcpyAsync(d_x, x, kIn * sizeof(float), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
      sched_predict_dev(s, d_x, 1, d_y) != VS_OK ||
      hipMemcpyAsync(y, d_y, kOut * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

uint32_t sched_tile_words(vs_sched_s *s, const QueryFeatures &f, uint32_t tw_max,
                          uint32_t fallback) {
  if (!s || tw_max < 16) return fallback;
  float x[kIn], y[kOut];
  query_features(f, x);
  if (!in_domain(s, x) || !predict_one(s, x, y)) return fallback;
  auto slot = [](uint32_t tw) { return tw == 16 ? kOptTW16 : tw == 32 ? kOptTW32 : kOptTW64; };
  if (fallback != 16 && fallback != 32 && fallback != 64) return fallback;
  uint32_t best = fallback;
  float bp = y[slot(fallback)];
  if (!std::isfinite(bp)) return fallback;
  for (uint32_t tw = 16; tw <= tw_max && tw <= 64; tw *= 2) {
    const float p = y[slot(tw)];
    if (std::isfinite(p) && p < bp) {
      bp = p;
      best = tw;
    }
  }
  // keep the fallback unless the pick is predicted clearly faster
  if (best != fallback && y[slot(fallback)] - bp < 0.1f) return fallback;
  return best;
}

int sched_decide(vs_sched_s *s, const SchedFeatures &f, const vs_tuning &t) {
  const int fallback = heuristic_decide(f, t);
  if (!s) return fallback;
  float x[kIn], y[kOut];
  sched_features(f, x);
  if (!in_domain(s, x) || !predict_one(s, x, y)) return fallback;
  const float tp = y[kOptPush], tl = y[kOptPull];
  if (!std::isfinite(tp) || !std::isfinite(tl)) return fallback;
  // within 0.1 log2 units (~7%) the prediction cannot separate the options:
  // keep the plain heuristic's choice
  if (std::fabs(tp - tl) < 0.1f) return fallback;
  return tp < tl ? 1 : 2;
}

} // namespace vs

using namespace vs;

extern "C" {

vs_status vs_sched_num_params(size_t *n) try {
  clear_error();
  if (!n) { set_error("vs_sched_num_params: null"); return VS_ERR_INVALID; }
  *n = kNumParams;
  return VS_OK;
} VS_CATCH("vs_sched_num_params")

vs_status vs_sched_create(uint64_t seed, vs_sched_t *out) try {
  clear_error();
  if (!out) { set_error("vs_sched_create: null"); return VS_ERR_INVALID; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    set_error("vs_sched_create: no GPU visible");
    return VS_ERR_NODEVICE;
  }
  std::unique_ptr<vs_sched_s> s(new vs_sched_s);
  VS_HIP(hipGetDevice(&s->device));
  VS_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  VS_HIP(hipMalloc(&s->d_params, kDeviceFloats * 4));
  VS_HIP(hipMemset(s->d_params, 0, kDeviceFloats * 4));  // no domain
  VS_HIP(hipMalloc(&s->d_m, kNumParams * 4));
  VS_HIP(hipMalloc(&s->d_v, kNumParams * 4));
  VS_HIP(hipMalloc(&s->d_grad, kNumParams * 4));
  VS_HIP(hipMalloc(&s->d_partial, static_cast<size_t>(kMlpBlocks) * kNumParams * 4 + kMlpBlocks * 4));
  VS_HIP(hipMalloc(&s->d_loss, (1 + kIn + kOut + 64) * 4));
  VS_HIP(hipMemset(s->d_m, 0, kNumParams * 4));
  VS_HIP(hipMemset(s->d_v, 0, kNumParams * 4));
  // He-uniform weights from the counter-based mixer, zero biases
  std::vector<float> p(kNumParams, 0.f);
  auto fill = [&](size_t off, size_t count, int fan_in) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (size_t i = 0; i < count; ++i) {
      const double u = (mix64(seed, off + i) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = static_cast<float>((2.0 * u - 1.0) * lim);
    }
  };
  fill(kOffW1, static_cast<size_t>(kHid) * kIn, kIn);
  fill(kOffW2, static_cast<size_t>(kHid) * kHid, kHid);
  fill(kOffW3, static_cast<size_t>(kOut) * kHid, kHid);
  VS_HIP(hipMemcpy(s->d_params, p.data(), kNumParams * 4, hipMemcpyHostToDevice));
  s->grid = kMlpBlocks;
  *out = s.release();
  return VS_OK;
} VS_CATCH("vs_sched_create")

} // extern "C"

namespace vs {

void sched_release(vs_sched_s *s) {
  if (!s || --s->refs > 0) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(s->d_params);
  (void)hipFree(s->d_m);
  (void)hipFree(s->d_v);
  (void)hipFree(s->d_grad);
  (void)hipFree(s->d_partial);
  (void)hipFree(s->d_loss);
  (void)hipStreamDestroy(s->stream);
  delete s;
  (void)hipSetDevice(cur);
}

} // namespace vs

extern "C" {

vs_status vs_sched_free(vs_sched_t s) try {
  clear_error();
  sched_release(s);  // attached graphs keep their own reference
  return VS_OK;
} VS_CATCH("vs_sched_free")

vs_status vs_sched_get_params(vs_sched_t s, float *params, size_t n) try {
  clear_error();
  if (!s || !params || n != kNumParams) {
    set_error("vs_sched_get_params: need %zu floats", kNumParams);
    return VS_ERR_INVALID;
  }
  VS_HIP(hipSetDevice(s->device));
  VS_HIP(hipStreamSynchronize(s->stream));
  VS_HIP(hipMemcpy(params, s->d_params, kNumParams * 4, hipMemcpyDeviceToHost));
  return VS_OK;
} VS_CATCH("vs_sched_get_params")

vs_status vs_sched_set_domain(vs_sched_t s, const float *rows, uint32_t n, float radius) try {
  clear_error();
  if (!s || (n && !rows) || n > kMaxDomain || !(radius >= 0.f)) {
    set_error("vs_sched_set_domain: need a scheduler, at most %u rows and a radius >= 0",
              kMaxDomain);
    return VS_ERR_INVALID;
  }
  std::vector<float> dev(2 + static_cast<size_t>(n) * kIn);
  dev[0] = static_cast<float>(n);
  dev[1] = radius;
  if (n) std::memcpy(dev.data() + 2, rows, static_cast<size_t>(n) * kIn * 4);
  VS_HIP(hipSetDevice(s->device));
  VS_HIP(hipStreamSynchronize(s->stream));
  VS_HIP(hipMemcpy(s->d_params + kOffDomain, dev.data(), dev.size() * 4, hipMemcpyHostToDevice));
  s->domain.assign(rows, rows + static_cast<size_t>(n) * kIn);
  s->radius = radius;
  return VS_OK;
} VS_CATCH("vs_sched_set_domain")

vs_status vs_sched_set_params(vs_sched_t s, const float *params, size_t n) try {
  clear_error();
  if (!s || !params || n != kNumParams) {
    set_error("vs_sched_set_params: need %zu floats", kNumParams);
    return VS_ERR_INVALID;
  }
  VS_HIP(hipSetDevice(s->device));
  VS_HIP(hipStreamSynchronize(s->stream));
  VS_HIP(hipMemcpy(s->d_params, params, kNumParams * 4, hipMemcpyHostToDevice));
  VS_HIP(hipMemset(s->d_m, 0, kNumParams * 4));
  VS_HIP(hipMemset(s->d_v, 0, kNumParams * 4));
  s->step = 0;
  return VS_OK;
} VS_CATCH("vs_sched_set_params")

// Uploads a batch; returns device pointers owned by the caller's lease.
struct Batch {
  float *X = nullptr, *Y = nullptr, *M = nullptr;
  float inv_cnt = 0.f;
  ~Batch() {
    (void)hipFree(X);
    (void)hipFree(Y);
    (void)hipFree(M);
  }
};

static vs_status upload_batch(const float *X, const float *Y, const float *M,
                              uint32_t n, Batch &b) {
  if (n == 0 || !X || !Y || !M) { set_error("empty or null training batch"); return VS_ERR_INVALID; }
  double cnt = 0;
  for (size_t i = 0; i < static_cast<size_t>(n) * kOut; ++i) cnt += M[i];
  if (cnt <= 0) { set_error("training mask is all zero"); return VS_ERR_INVALID; }
  b.inv_cnt = static_cast<float>(1.0 / cnt);
  VS_HIP(hipMalloc(&b.X, static_cast<size_t>(n) * kIn * 4));
  VS_HIP(hipMalloc(&b.Y, static_cast<size_t>(n) * kOut * 4));
  VS_HIP(hipMalloc(&b.M, static_cast<size_t>(n) * kOut * 4));
  VS_HIP(hipMemcpy(b.X, X, static_cast<size_t>(n) * kIn * 4, hipMemcpyHostToDevice));
  VS_HIP(hipMemcpy(b.Y, Y, static_cast<size_t>(n) * kOut * 4, hipMemcpyHostToDevice));
  VS_HIP(hipMemcpy(b.M, M, static_cast<size_t>(n) * kOut * 4, hipMemcpyHostToDevice));
  return VS_OK;
}

static vs_status grad_step(vs_sched_s *s, const Batch &b, uint32_t n) {
  float *loss_partial = s->d_partial + static_cast<size_t>(s->grid) * kNumParams;
  hipLaunchKernelGGL(k_mlp_grad, dim3(s->grid), dim3(kBlock), 0, s->stream,
                     s->d_params, b.X, b.Y, b.M, n, b.inv_cnt, s->d_partial,
                     loss_partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNumParams + 255) / 256), dim3(256), 0,
                     s->stream, s->d_partial, loss_partial, s->grid, s->d_grad,
                     s->d_loss);
  VS_HIP(hipGetLastError());
  return VS_OK;
}

vs_status vs_sched_loss_grad(vs_sched_t s, const float *X, const float *Y,
                             const float *M, uint32_t n, float *loss, float *grad) try {
  clear_error();
  if (!s) { set_error("vs_sched_loss_grad: null"); return VS_ERR_INVALID; }
  VS_HIP(hipSetDevice(s->device));
  Batch b;
  VS_TRY(upload_batch(X, Y, M, n, b));
  VS_TRY(grad_step(s, b, n));
  if (grad) VS_HIP(hipMemcpyAsync(grad, s->d_grad, kNumParams * 4, hipMemcpyDeviceToHost, s->stream));
  float l = 0.f;
  VS_HIP(hipMemcpyAsync(&l, s->d_loss, 4, hipMemcpyDeviceToHost, s->stream));
  VS_HIP(hipStreamSynchronize(s->stream));
  if (loss) *loss = l;
  return VS_OK;
} VS_CATCH("vs_sched_loss_grad")

vs_status vs_sched_train(vs_sched_t s, const float *X, const float *Y,
                         const float *M, uint32_t n, uint32_t epochs, float lr,
                         float *loss_curve) try {
  clear_error();
  if (!s) { set_error("vs_sched_train: null"); return VS_ERR_INVALID; }
  VS_HIP(hipSetDevice(s->device));
  Batch b;
  VS_TRY(upload_batch(X, Y, M, n, b));
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  std::vector<float> curve(loss_curve ? epochs : 0);
  for (uint32_t e = 0; e < epochs; ++e) {
    VS_TRY(grad_step(s, b, n));
    s->step++;
    const float bc1 = 1.f - std::pow(beta1, static_cast<float>(s->step));
    const float bc2 = 1.f - std::pow(beta2, static_cast<float>(s->step));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNumParams + 255) / 256), dim3(256), 0,
                       s->stream, s->d_params, s->d_m, s->d_v, s->d_grad, lr, beta1,
                       beta2, eps, bc1, bc2);
    VS_HIP(hipGetLastError());
    if (loss_curve)
      VS_HIP(hipMemcpyAsync(&curve[e], s->d_loss, 4, hipMemcpyDeviceToHost, s->stream));
  }
  VS_HIP(hipStreamSynchronize(s->stream));
  if (loss_curve) std::memcpy(loss_curve, curve.data(), epochs * 4);
  return VS_OK;
} VS_CATCH("vs_sched_train")

vs_status vs_sched_predict(vs_sched_t s, const float *X, uint32_t n, float *Y) try {
  clear_error();
  if (!s || (n > 0 && (!X || !Y))) { set_error("vs_sched_predict: null"); return VS_ERR_INVALID; }
  if (n == 0) return VS_OK;
  VS_HIP(hipSetDevice(s->device));
  float *dX = nullptr, *dY = nullptr;
  VS_HIP(hipMalloc(&dX, static_cast<size_t>(n) * kIn * 4));
  hipError_t e = hipMalloc(&dY, static_cast<size_t>(n) * kOut * 4);
  if (e == hipSuccess) e = hipMemcpy(dX, X, static_cast<size_t>(n) * kIn * 4, hipMemcpyHostToDevice);
  vs_status st = VS_OK;
  if (e == hipSuccess) st = sched_predict_dev(s, dX, n, dY);
  if (e == hipSuccess && st == VS_OK)
    e = hipMemcpyAsync(Y, dY, static_cast<size_t>(n) * kOut * 4, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dX);
  (void)hipFree(dY);
  if (e != hipSuccess) {
    set_error("vs_sched_predict: %s", hipGetErrorString(e));
    return VS_ERR_HIP;
  }
  return st;
} VS_CATCH("vs_sched_predict")

vs_status vs_graph_set_scheduler(vs_graph_t g, vs_sched_t s) try {
  clear_error();
  if (!g) { set_error("vs_graph_set_scheduler: null graph"); return VS_ERR_INVALID; }
  if (s) s->refs++;  // the graph's own reference: the caller may free its handle
  if (g->sched) sched_release(g->sched);
  g->sched = s;
  return VS_OK;
} VS_CATCH("vs_graph_set_scheduler")

} // extern "C"
