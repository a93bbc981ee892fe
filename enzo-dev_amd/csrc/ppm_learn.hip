// ppm_learn.hip -- training and inference kernels of the learned two-shock
// starting pressure (ppm_learn.hpp): features from harvested interface
// records, a forward + backward kernel that sums each workgroup's parameter
// gradients in LDS, Adam, and a forward-only kernel for the loss and the
// predictions.  FP32 throughout; the loss is the mean squared error of
// y = log(p* / p_left).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/enzo_ppm.h"
#include "ppm_sweep.hpp"

using namespace ppm;

namespace {

constexpr int kB = 256;  // samples per workgroup

// records (kHarvestDoubles doubles each) -> features X[n][6], targets Y[n]
// (log(p* / p_ac), clipped to [-kYClip, kYClip]; NaN where the acoustic
// start or p* is not positive -- training leaves those records out)
__global__ void learn_features_kernel(const double *rec, long long n, double gamma, float *X,
                                      float *Y) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double *r = rec + s * kHarvestDoubles;
  float x[kMlpIn];
  mlp_features(r[0], r[1], r[2], r[3], r[4], r[5], gamma, x);
  for (int k = 0; k < kMlpIn; k++) X[s * kMlpIn + k] = x[k];
  const double cl = sqrt(gamma * r[2] * r[0]), cr = sqrt(gamma * r[3] * r[1]);
  const double pac = acoustic_start(r[0], r[1], r[2], r[3], r[4], r[5], cl, cr);
  Y[s] = pac > 0.0 && r[6] > 0.0 ? fminf(fmaxf(mlp_log((float)(r[6] / pac)), -kYClip), kYClip)
                                 : __builtin_nanf("");
}

// forward of one sample: standardised input, hidden layers, output
__device__ inline float forward(const float *w, const float *x, float *xn, float *h1, float *h2) {
  for (int k = 0; k < kMlpIn; k++) xn[k] = (x[k] - w[kOffMean + k]) * w[kOffInvStd + k];
  for (int j = 0; j < kMlpHidden; j++) {
    float a = w[kOffB1 + j];
    for (int k = 0; k < kMlpIn; k++) a += w[kOffW1 + j * kMlpIn + k] * xn[k];
    h1[j] = mlp_tanh(a);
  }
  float y = w[kOffB3];
  for (int j = 0; j < kMlpHidden; j++) {
    float a = w[kOffB2 + j];
    for (int k = 0; k < kMlpHidden; k++) a += w[kOffW2 + j * kMlpHidden + k] * h1[k];
    h2[j] = mlp_tanh(a);
    y += w[kOffW3 + j] * h2[j];
  }
  return y;
}

// gradient of scale * sum (y - target)^2 over samples idx[b0 .. b0+bn) into
// grad (atomically), and the sum of squared errors into *sse
__global__ __launch_bounds__(kB) void learn_grad_kernel(const float *X, const float *Y,
                                                        const int *idx, long long b0, long long bn,
                                                        const float *w, float scale, float *grad,
                                                        float *sse) {
  __shared__ float xs[kB][kMlpIn];
  __shared__ float h1s[kB][kMlpHidden];
  __shared__ float gh2s[kB][kMlpHidden];
  __shared__ float d2s[kB][kMlpHidden];
  __shared__ float d1s[kB][kMlpHidden];
  __shared__ float gs[kB];
  __shared__ float es[kB];
  const int t = threadIdx.x;
  const long long q = (long long)blockIdx.x * kB + t;
  float xn[kMlpIn], h1[kMlpHidden], h2[kMlpHidden];
  if (q < bn) {
    const long long s = idx ? idx[b0 + q] : b0 + q;
    const float y = forward(w, X + s * kMlpIn, xn, h1, h2);
    const float e = y - Y[s];
    const float g = scale * 2.0f * e;
    es[t] = e * e;
    gs[t] = g;
    for (int k = 0; k < kMlpIn; k++) xs[t][k] = xn[k];
    float d2[kMlpHidden];
    for (int j = 0; j < kMlpHidden; j++) {
      h1s[t][j] = h1[j];
      gh2s[t][j] = g * h2[j];
      d2[j] = g * w[kOffW3 + j] * (1.0f - h2[j] * h2[j]);
      d2s[t][j] = d2[j];
    }
    for (int k = 0; k < kMlpHidden; k++) {
      float a = 0.0f;
      for (int j = 0; j < kMlpHidden; j++) a += d2[j] * w[kOffW2 + j * kMlpHidden + k];
      d1s[t][k] = a * (1.0f - h1[k] * h1[k]);
    }
  } else {
    es[t] = 0.0f;
    gs[t] = 0.0f;
    for (int k = 0; k < kMlpIn; k++) xs[t][k] = 0.0f;
    for (int j = 0; j < kMlpHidden; j++) h1s[t][j] = gh2s[t][j] = d2s[t][j] = d1s[t][j] = 0.0f;
  }
  __syncthreads();
  for (int p = t; p < kMlpParams; p += kB) {
    float a = 0.0f;
    if (p < kOffB1) {
      const int j = p / kMlpIn, k = p % kMlpIn;
      for (int s = 0; s < kB; s++) a += d1s[s][j] * xs[s][k];
    } else if (p < kOffW2) {
      const int j = p - kOffB1;
      for (int s = 0; s < kB; s++) a += d1s[s][j];
    } else if (p < kOffB2) {
      const int j = (p - kOffW2) / kMlpHidden, k = (p - kOffW2) % kMlpHidden;
      for (int s = 0; s < kB; s++) a += d2s[s][j] * h1s[s][k];
    } else if (p < kOffW3) {
      const int j = p - kOffB2;
      for (int s = 0; s < kB; s++) a += d2s[s][j];
    } else if (p < kOffB3) {
      const int j = p - kOffW3;
      for (int s = 0; s < kB; s++) a += gh2s[s][j];
    } else {
      for (int s = 0; s < kB; s++) a += gs[s];
    }
    atomicAdd(&grad[p], a);
  }
  if (t == 0) {
    float a = 0.0f;
    for (int s = 0; s < kB; s++) a += es[s];
    atomicAdd(sse, a);
  }
}

__global__ void learn_adam_kernel(float *w, const float *g, float *m, float *v, float lr,
                                  float beta1, float beta2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gi = g[p];
  m[p] = beta1 * m[p] + (1.0f - beta1) * gi;
  v[p] = beta2 * v[p] + (1.0f - beta2) * gi * gi;
  w[p] -= lr * (m[p] / c1) / (sqrtf(v[p] / c2) + eps);
}

__global__ void learn_zero_kernel(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// forward only: predictions into out (if non-null); squared errors over
// the finite targets summed into acc[0], their number into acc[1]
__global__ __launch_bounds__(kB) void learn_eval_kernel(const float *X, const float *Y, long long n,
                                                        const float *w, float *out, float *acc) {
  __shared__ float es[kB], cs[kB];
  const long long s = (long long)blockIdx.x * kB + threadIdx.x;
  float e2 = 0.0f, c = 0.0f;
  if (s < n) {
    float xn[kMlpIn], h1[kMlpHidden], h2[kMlpHidden];
    const float y = forward(w, X + s * kMlpIn, xn, h1, h2);
    if (out) out[s] = y;
    if (Y[s] == Y[s]) {
      e2 = (y - Y[s]) * (y - Y[s]);
      c = 1.0f;
    }
  }
  es[threadIdx.x] = e2;
  cs[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.0f, b = 0.0f;
    for (int k = 0; k < kB; k++) {
      a += es[k];
      b += cs[k];
    }
    atomicAdd(&acc[0], a);
    atomicAdd(&acc[1], b);
  }
}

// per record: the two-shock solve from the reference start and from the
// network's start; out[7] = {updates (reference start), updates (learned
// start, restart included), restarted, p* (reference start), p* (learned
// start), learned start, reference start}
__global__ void learn_newton_kernel(const double *rec, long long n, double gamma, double pmin,
                                    const float *w, double *out) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double *r = rec + s * kHarvestDoubles;
  double pl = r[2], pr = r[3], pbar, ubar, ps0;
  const int it0 = twoshock1(r[0], r[1], pl, pr, r[4], r[5], gamma, pmin, 0, pbar, ubar, ps0);
  double *o = out + s * 7;
  o[0] = it0;
  o[3] = pbar;
  o[6] = ps0;
  pl = r[2];
  pr = r[3];
  int restarted = 0;
  const Acoustic A = acoustic_of(r[0], r[1], pl, pr, r[4], r[5], gamma, pmin);
  const bool gated = learn_gate(w, A, pl, pr, pmin);
  float c = 0.0f;
  if (gated) {
    float x[kMlpIn];
    mlp_features(r[0], r[1], pl, pr, r[4], r[5], gamma, x);
    c = mlp_eval(w, x);
  }
  const int it1 = twoshock_learned(gated, c, A, r[0], r[1], pl, pr, r[4], r[5], gamma, pmin, 0,
                                   pbar, ubar, ps0, restarted);
  o[1] = it1;
  o[2] = restarted;
  o[4] = pbar;
  o[5] = ps0;
}

thread_local char g_learn_error[512] = "";

#define LNCHK(x)                                                                           \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      snprintf(g_learn_error, sizeof(g_learn_error), "%s failed: %s (ppm_learn.hip:%d)", #x, \
               hipGetErrorString(e_), __LINE__);                                           \
      return ENZO_PPM_ERR_HIP;                                                             \
    }                                                                                      \
  } while (0)

// device buffer freed at scope exit
template <class T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)); }
};

// device features and targets of n records
struct Dataset {
  DevBuf<float> X, Y;
  long long n = 0;
};

// a record the network is trained on: positive states, positive p* and a
// positive acoustic start (cavitating interfaces keep the reference start)
bool learnable(const double *r, double gamma) {
  if (!(r[0] > 0.0 && r[1] > 0.0 && r[2] > 0.0 && r[3] > 0.0 && r[6] > 0.0)) return false;
  const double cl = std::sqrt(gamma * r[2] * r[0]), cr = std::sqrt(gamma * r[3] * r[1]);
  return acoustic_start(r[0], r[1], r[2], r[3], r[4], r[5], cl, cr) > 0.0;
}

// features and targets on the device; with `filter` only of the learnable
// records (their order kept), else of all (targets of the others are
// meaningless and left out of the error by evaluate())
int make_dataset(const enzo_float *records, long long n, double gamma, bool filter, Dataset &D) {
  std::vector<double> kept;
  if (filter) {
    long long m = 0;
    for (long long i = 0; i < n; i++) m += learnable(records + i * kHarvestDoubles, gamma) ? 1 : 0;
    if (m == 0) {
      snprintf(g_learn_error, sizeof(g_learn_error), "no learnable records");
      return ENZO_PPM_ERR_ARGUMENT;
    }
    if (m < n) {
      kept.reserve((size_t)m * kHarvestDoubles);
      for (long long i = 0; i < n; i++)
        if (learnable(records + i * kHarvestDoubles, gamma))
          kept.insert(kept.end(), records + i * kHarvestDoubles,
                      records + (i + 1) * kHarvestDoubles);
      records = kept.data();
      n = m;
    }
  }
  DevBuf<double> drec;
  LNCHK(drec.alloc((size_t)kHarvestDoubles * n));
  LNCHK(D.X.alloc((size_t)kMlpIn * n));
  LNCHK(D.Y.alloc((size_t)n));
  LNCHK(hipMemcpy(drec.p, records, sizeof(double) * kHarvestDoubles * n, hipMemcpyHostToDevice));
  D.n = n;
  hipLaunchKernelGGL(learn_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                     drec.p, n, gamma, D.X.p, D.Y.p);
  LNCHK(hipGetLastError());
  LNCHK(hipDeviceSynchronize());
  return ENZO_PPM_SUCCESS;
}

// mean squared error over the records with a finite target (and the
// predictions of all) with device parameters dw
int evaluate(const Dataset &D, const float *dw, float *dout, double *mse) {
  DevBuf<float> acc;  // {sum of squared errors, count}
  LNCHK(acc.alloc(2));
  LNCHK(hipMemset(acc.p, 0, sizeof(float) * 2));
  hipLaunchKernelGGL(learn_eval_kernel, dim3((unsigned)((D.n + kB - 1) / kB)), dim3(kB), 0, 0,
                     D.X.p, D.Y.p, D.n, dw, dout, acc.p);
  LNCHK(hipGetLastError());
  float h[2] = {0.0f, 0.0f};
  LNCHK(hipMemcpy(h, acc.p, sizeof(h), hipMemcpyDeviceToHost));
  *mse = h[1] > 0.0f ? (double)h[0] / (double)h[1] : 0.0;
  return ENZO_PPM_SUCCESS;
}

}  // namespace

extern "C" {

int hipPPMLearnTrain(const enzo_float *records, enzo_long_int n, enzo_float gamma, enzo_int epochs,
                     enzo_int batch, enzo_float lr, enzo_int seed, float *weights,
                     enzo_float *mse_per_epoch) {
  if (!records || n < 1 || epochs < 0 || batch < kB || !weights || !(lr > 0.0))
    return ENZO_PPM_ERR_ARGUMENT;
  Dataset D;
  int st = make_dataset(records, n, gamma, true, D);
  if (st != ENZO_PPM_SUCCESS) return st;
  const long long m = D.n;
  if (m < batch) {
    snprintf(g_learn_error, sizeof(g_learn_error), "%lld learnable records, fewer than a batch", m);
    return ENZO_PPM_ERR_ARGUMENT;
  }
  // standardisation of the inputs: mean and standard deviation of the
  // device-computed features (6 numbers, on the host)
  {
    std::vector<float> hx((size_t)kMlpIn * m);
    LNCHK(hipMemcpy(hx.data(), D.X.p, sizeof(float) * hx.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < kMlpIn; k++) {
      double s1 = 0.0, s2 = 0.0;
      for (long long i = 0; i < m; i++) {
        s1 += hx[i * kMlpIn + k];
        s2 += (double)hx[i * kMlpIn + k] * hx[i * kMlpIn + k];
      }
      const double mu = s1 / m, var = std::max(s2 / m - mu * mu, 1e-12);
      weights[kOffMean + k] = (float)mu;
      weights[kOffInvStd + k] = (float)(1.0 / std::sqrt(var));
    }
  }
  DevBuf<float> dw, dg, dm, dv, dsse;
  DevBuf<int> didx;
  LNCHK(dw.alloc(kMlpFloats));
  LNCHK(dg.alloc(kMlpParams));
  LNCHK(dm.alloc(kMlpParams));
  LNCHK(dv.alloc(kMlpParams));
  LNCHK(dsse.alloc(1));
  LNCHK(didx.alloc((size_t)m));
  LNCHK(hipMemcpy(dw.p, weights, sizeof(float) * kMlpFloats, hipMemcpyHostToDevice));
  LNCHK(hipMemset(dm.p, 0, sizeof(float) * kMlpParams));
  LNCHK(hipMemset(dv.p, 0, sizeof(float) * kMlpParams));
  std::vector<int> perm((size_t)m);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng((unsigned long long)seed);
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const long long steps_total = (long long)epochs * (m / batch);
  long long t = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    LNCHK(hipMemcpy(didx.p, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice));
    for (long long b0 = 0; b0 + batch <= m; b0 += batch) {
      t++;
      hipLaunchKernelGGL(learn_zero_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, 0, dg.p,
                         kMlpParams);
      hipLaunchKernelGGL(learn_zero_kernel, dim3(1), dim3(64), 0, 0, dsse.p, 1);
      hipLaunchKernelGGL(learn_grad_kernel, dim3((unsigned)((batch + kB - 1) / kB)), dim3(kB), 0, 0,
                         D.X.p, D.Y.p, didx.p, b0, (long long)batch, dw.p, 1.0f / (float)batch,
                         dg.p, dsse.p);
      const float c1 = 1.0f - std::pow(beta1, (float)t), c2 = 1.0f - std::pow(beta2, (float)t);
      // learning rate decayed linearly to a tenth over the run
      const float lr_t = (float)(lr * (1.0 - 0.9 * (double)(t - 1) / (double)steps_total));
      hipLaunchKernelGGL(learn_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, 0, dw.p,
                         dg.p, dm.p, dv.p, lr_t, beta1, beta2, eps, c1, c2);
    }
    LNCHK(hipGetLastError());
    if (mse_per_epoch) {
      double mse = 0.0;
      st = evaluate(D, dw.p, nullptr, &mse);
      if (st != ENZO_PPM_SUCCESS) return st;
      mse_per_epoch[ep] = mse;
    }
  }
  LNCHK(hipMemcpy(weights, dw.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return ENZO_PPM_SUCCESS;
}

// full-batch gradient of the mean squared error over the learnable records,
// for checking the backward kernel (grad: kMlpParams doubles, in the
// parameter order of ppm_learn.hpp)
int hipPPMLearnGradient(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                        const float *weights, enzo_float *grad, enzo_float *mse) {
  if (!records || n < 1 || !weights || !grad) return ENZO_PPM_ERR_ARGUMENT;
  Dataset D;
  int st = make_dataset(records, n, gamma, true, D);
  if (st != ENZO_PPM_SUCCESS) return st;
  DevBuf<float> dw, dg, dsse;
  LNCHK(dw.alloc(kMlpFloats));
  LNCHK(dg.alloc(kMlpParams));
  LNCHK(dsse.alloc(1));
  LNCHK(hipMemcpy(dw.p, weights, sizeof(float) * kMlpFloats, hipMemcpyHostToDevice));
  LNCHK(hipMemset(dg.p, 0, sizeof(float) * kMlpParams));
  LNCHK(hipMemset(dsse.p, 0, sizeof(float)));
  hipLaunchKernelGGL(learn_grad_kernel, dim3((unsigned)((D.n + kB - 1) / kB)), dim3(kB), 0, 0,
                     D.X.p, D.Y.p, (const int *)nullptr, 0ll, D.n, dw.p, 1.0f / (float)D.n, dg.p,
                     dsse.p);
  LNCHK(hipGetLastError());
  std::vector<float> g(kMlpParams);
  float sse = 0.0f;
  LNCHK(hipMemcpy(g.data(), dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  LNCHK(hipMemcpy(&sse, dsse.p, sizeof(float), hipMemcpyDeviceToHost));
  for (int p = 0; p < kMlpParams; p++) grad[p] = g[p];
  if (mse) *mse = (double)sse / (double)D.n;
  return ENZO_PPM_SUCCESS;
}

// network outputs (the correction log(p* / p_ac)) for n records, and their
// mean squared error over the learnable ones
int hipPPMLearnPredict(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                       const float *weights, float *y, enzo_float *mse) {
  if (!records || n < 1 || !weights) return ENZO_PPM_ERR_ARGUMENT;
  Dataset D;
  int st = make_dataset(records, n, gamma, false, D);
  if (st != ENZO_PPM_SUCCESS) return st;
  DevBuf<float> dw, dy;
  LNCHK(dw.alloc(kMlpFloats));
  LNCHK(dy.alloc((size_t)n));
  LNCHK(hipMemcpy(dw.p, weights, sizeof(float) * kMlpFloats, hipMemcpyHostToDevice));
  double m = 0.0;
  st = evaluate(D, dw.p, dy.p, &m);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (y) LNCHK(hipMemcpy(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (mse) *mse = m;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLearnNewtonStats(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                           enzo_float pmin, const float *weights, enzo_float *out) {
  if (!records || n < 1 || !weights || !out) return ENZO_PPM_ERR_ARGUMENT;
  DevBuf<double> drec, dout;
  DevBuf<float> dw;
  LNCHK(drec.alloc((size_t)kHarvestDoubles * n));
  LNCHK(dout.alloc((size_t)7 * n));
  LNCHK(dw.alloc(kMlpFloats));
  LNCHK(hipMemcpy(drec.p, records, sizeof(double) * kHarvestDoubles * n, hipMemcpyHostToDevice));
  LNCHK(hipMemcpy(dw.p, weights, sizeof(float) * kMlpFloats, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(learn_newton_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                     drec.p, (long long)n, gamma, pmin, dw.p, dout.p);
  LNCHK(hipGetLastError());
  LNCHK(hipMemcpy(out, dout.p, sizeof(double) * 7 * n, hipMemcpyDeviceToHost));
  return ENZO_PPM_SUCCESS;
}

int hipPPMLearnParameterCount(void) { return kMlpFloats; }

const char *hipPPMLearnLastError(void) { return g_learn_error; }

}  // extern "C"
