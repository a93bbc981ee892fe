// ppm_learn.hpp -- the learned starting pressure of the two-shock Newton
// iteration: a 6 -> 32 -> 32 -> 1 FP32 network with tanh hidden layers that
// predicts the correction c = log(p* / p_ac) to the reference's acoustic
// start p_ac (twoshock.F:156-170), so that log(p_ac / p_left) + c predicts
// log(p* / p_left).  The start is p_ac e^c at interfaces whose acoustic wave
// strength passes the gate (a calibrated threshold stored with the
// weights), p_ac elsewhere: for weak waves the acoustic start's error,
// second order in the strength, is already below the network's own.
// Shared by the sweep kernel (inference per interface), the training
// kernels (ppm_learn.hip) and the CPU builds.
//
// Inputs (dimensionless, so the guess scales with the states):
//   x0 = log(p_r / p_l), x1 = log(rho_r / rho_l),
//   x2 = (u_l - u_r) / c_l, x3 = (u_l - u_r) / c_r,
//   x4 = c_l / (c_l + c_r), x5 = c_r / (c_l + c_r),
// with c = sqrt(gamma p / rho); each standardised with the training set's
// mean and standard deviation, which travel with the weights.
#pragma once

#include <math.h>

// included by ppm_sweep.hpp after PPM_HD is defined

namespace ppm {

constexpr int kMlpIn = 6, kMlpHidden = 32;
// parameter block (floats): W1[H][IN] b1[H] W2[H][H] b2[H] W3[H] b3, mean[IN], inv_std[IN],
// gate
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHidden * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden;
constexpr int kMlpParams = kOffB3 + 1;  // trained parameters: 1313
constexpr int kOffMean = kMlpParams;
constexpr int kOffInvStd = kOffMean + kMlpIn;
constexpr int kOffGate = kOffInvStd + kMlpIn;  // wave-strength threshold for the network
constexpr int kMlpFloats = kOffGate + 1;
// the device copy appends W1 and W2 transposed ([IN][H] and [H][H], input
// index outer), so that the work-items of consecutive hidden units read
// consecutive weights (learn_device_weights)
constexpr int kOffW1T = kMlpFloats;
constexpr int kOffW2T = kOffW1T + kMlpIn * kMlpHidden;
constexpr int kMlpDevFloats = kOffW2T + kMlpHidden * kMlpHidden;

// A parameter block is valid when every value is finite, except the gate,
// which may also be +infinity: the gate closed, no interface evaluates the
// network, every start is the reference one and the library runs its plain
// kernels (bitwise the reference start's results, at no cost).
inline bool learn_gate_closed(const float *w) { return w[kOffGate] == INFINITY; }
inline bool learn_weights_valid(const float *w) {
  for (int k = 0; k < kMlpFloats; k++)
    if (!(std::isfinite(w[k]) || (k == kOffGate && learn_gate_closed(w)))) return false;
  return true;
}

// the device layout of a parameter block w (kMlpFloats) into d (kMlpDevFloats)
inline void learn_device_weights(const float *w, float *d) {
  for (int k = 0; k < kMlpFloats; k++) d[k] = w[k];
  for (int j = 0; j < kMlpHidden; j++)
    for (int k = 0; k < kMlpIn; k++) d[kOffW1T + k * kMlpHidden + j] = w[kOffW1 + j * kMlpIn + k];
  for (int j = 0; j < kMlpHidden; j++)
    for (int k = 0; k < kMlpHidden; k++)
      d[kOffW2T + k * kMlpHidden + j] = w[kOffW2 + j * kMlpHidden + k];
}
// training targets are clipped to [-kYClip, kYClip]: p* at the pressure
// floor gives corrections near -46 and would dominate the squared error
constexpr float kYClip = 8.0f;

// FP32 exponential and logarithm of the network: the hardware's on the
// device (a few ulp), libm's in CPU builds
PPM_HD float mlp_exp(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(a);
#else
  return expf(a);
#endif
}
PPM_HD float mlp_log(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __logf(a);
#else
  return logf(a);
#endif
}

// tanh as 1 - 2 / (1 + e^(2a)): on the device the hardware exponential and
// a fast quotient (about 1e-7 absolute error; the network only supplies a
// starting value that Newton then refines).  Used identically in training
// and inference.
PPM_HD float mlp_tanh(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return 1.0f - __fdividef(2.0f, 1.0f + __expf(2.0f * a));
#else
  return 1.0f - 2.0f / (1.0f + expf(2.0f * a));
#endif
}

// the reference's acoustic start (twoshock.F:156-170) before the floor,
// cl and cr being the Lagrangian wave speeds sqrt(gamma p rho)
PPM_HD double acoustic_start(double dl, double dr, double pl, double pr, double ul, double ur,
                             double cl, double cr) {
  (void)dl;
  (void)dr;
  return (cr * pl + cl * pr + cr * cl * (ul - ur)) / (cr + cl);
}

// features of one interface (FP32 after the FP64 ratios)
PPM_HD void mlp_features(double dl, double dr, double pl, double pr, double ul, double ur,
                         double gamma, float x[kMlpIn]) {
  const double csl = sqrt(gamma * pl / dl), csr = sqrt(gamma * pr / dr);
  const double du = ul - ur;
  x[0] = mlp_log((float)(pr / pl));
  x[1] = mlp_log((float)(dr / dl));
  x[2] = (float)(du / csl);
  x[3] = (float)(du / csr);
  x[4] = (float)(csl / (csl + csr));
  x[5] = (float)(csr / (csl + csr));
}

// network output for standardised-on-the-fly features
PPM_HD float mlp_eval(const float *w, const float x[kMlpIn]) {
  float xn[kMlpIn];
  for (int k = 0; k < kMlpIn; k++) xn[k] = (x[k] - w[kOffMean + k]) * w[kOffInvStd + k];
  float h1[kMlpHidden];
  for (int j = 0; j < kMlpHidden; j++) {
    float a = w[kOffB1 + j];
    for (int k = 0; k < kMlpIn; k++) a += w[kOffW1 + j * kMlpIn + k] * xn[k];
    h1[j] = mlp_tanh(a);
  }
  float y = w[kOffB3];
  for (int j = 0; j < kMlpHidden; j++) {
    float a = w[kOffB2 + j];
    for (int k = 0; k < kMlpHidden; k++) a += w[kOffW2 + j * kMlpHidden + k] * h1[k];
    y += w[kOffW3 + j] * mlp_tanh(a);
  }
  return y;
}

}  // namespace ppm
