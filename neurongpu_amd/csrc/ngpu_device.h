// Device helpers of the per-timestep kernels: FP32 exp/log, the
// Cash-Karp RK5 tableau, Threefry2x32-20 and Philox4x32-10, and the
// synapse update rules.
//
// Every multiply-add that is fused is written as an explicit fmaf; the
// kernels are compiled with -ffp-contract=off, so the FP32 arithmetic is a
// fixed operation sequence (the CPU oracle in oracle/oracle.c follows the
// same sequence and is compared against these kernels bit for bit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ngpu_kernels.h"

namespace ngpu {

// exp(x): Cody-Waite reduction x = k ln2 + r (ln2 split 0x3F317200 +
// 1.42860677e-06), exp(r) by the degree-7 Taylor polynomial on
// |r| <= ln2/2, then 2^k applied by ldexp (one instruction; exact, or
// rounded once where the result is subnormal).
// Written without branches or selects: x is clamped to [-104, 88.8] by
// NaN-propagating minimum/maximum, and the polynomial path already gives
// +inf at 88.8 and +0 at -104 (and NaN for NaN), the values beyond the
// range, so a wave never splits on the argument's range.
__device__ __forceinline__ float dev_expf(float x) {
  const float xc = __builtin_elementwise_minimum(
      __builtin_elementwise_maximum(x, -104.0f), 88.8f);
  float kf = rintf(xc * 1.44269502f);
  float r = __builtin_fmaf(kf, -0.693145751953125f, xc);
  r = __builtin_fmaf(kf, -1.42860677e-06f, r);
  float p = 1.98412701e-04f;
  p = __builtin_fmaf(p, r, 1.38888892e-03f);
  p = __builtin_fmaf(p, r, 8.33333377e-03f);
  p = __builtin_fmaf(p, r, 4.16666679e-02f);
  p = __builtin_fmaf(p, r, 1.66666672e-01f);
  p = __builtin_fmaf(p, r, 0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_amdgcn_ldexpf(p, (int)kf);
}

// log(x) for x > 0: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh((m-1)/(m+1)) by the odd series to s^9.  Branch-free like
// dev_expf: x <= 0, +inf and NaN are selected at the end.
__device__ __forceinline__ float dev_logf(float x) {
  uint32_t b = __float_as_uint(x);
  int e = (int)((b >> 23) & 0xFFu) - 127;
  float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
  const bool hi = m > 1.41421354f;
  m = hi ? m * 0.5f : m;
  e = hi ? e + 1 : e;
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float q = 1.11111112e-01f;
  q = __builtin_fmaf(q, s2, 1.42857149e-01f);
  q = __builtin_fmaf(q, s2, 2.00000003e-01f);
  q = __builtin_fmaf(q, s2, 3.33333343e-01f);
  float two_s = 2.0f * s;
  float lm = __builtin_fmaf(two_s * s2, q, two_s);
  float ef = (float)e;
  float res = __builtin_fmaf(ef, 0.693145751953125f,
                             __builtin_fmaf(ef, 1.42860677e-06f, lm));
  res = x == __uint_as_float(0x7F800000u) ? x : res;
  res = x <= 0.0f ? __uint_as_float(0xFF800000u) : res;
  return x != x ? x : res;
}

// Element i of a device array, addressed by a 32-bit byte offset from the
// array base (wave-uniform, a kernel argument): the global memory
// instructions then take the base in scalar registers and the offset in
// one vector register, instead of a 64-bit vector address per access.
// Valid while i * sizeof(T) < 2^32 (node indices are below 2^26).
template <class T>
__device__ __forceinline__ T &at(T *base, uint32_t i) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                i * (uint32_t)sizeof(T));
}

// err^c for the step-size controller: c = -1/5 after an accepted attempt
// (src/rk5.h:122-146 takes pow(err, -0.20)), c = -1/4 after a rejected one
// (pow(err, -0.25)).  One branch-free sequence for both, so a wave whose
// lanes split between the two cases runs it once, with no division:
// err = 2^k m, m in [sqrt(1/2), sqrt(2)); log2 m = t Q(t), t = m - 1 (exact),
// Q of degree 7 (least-squares fit of log2(1+t)/t, coefficients rounded to
// float); y = c k + c log2 m; 2^y = 2^j R(y - j), j = rint(y), R of degree 5
// (fit of 2^f on [-1/2, 1/2]).  Relative error below 4e-7 against pow over
// err in [1e-4, 1e8]; +inf gives 0 and NaN gives NaN, as pow does.
__device__ __forceinline__ float dev_pow_ctrl(float e, float c) {
  const uint32_t b = __float_as_uint(e);
  int k = (int)((b >> 23) & 0xFFu) - 127;
  float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
  const bool hi = m > 1.41421354f;
  m = hi ? m * 0.5f : m;
  k = hi ? k + 1 : k;
  const float t = m - 1.0f;
  float q = -0.14237648f;
  q = __builtin_fmaf(q, t, 0.23247829f);
  q = __builtin_fmaf(q, t, -0.24932154f);
  q = __builtin_fmaf(q, t, 0.28731275f);
  q = __builtin_fmaf(q, t, -0.36022335f);
  q = __builtin_fmaf(q, t, 0.48091587f);
  q = __builtin_fmaf(q, t, -0.72135293f);
  q = __builtin_fmaf(q, t, 1.442695f);
  const float y = __builtin_fmaf(c, (float)k, c * (t * q));
  const float j = rintf(y);
  const float f = y - j;
  float r = 0.0013266971f;
  r = __builtin_fmaf(r, f, 0.00967546f);
  r = __builtin_fmaf(r, f, 0.055507425f);
  r = __builtin_fmaf(r, f, 0.24022122f);
  r = __builtin_fmaf(r, f, 0.69314694f);
  r = __builtin_fmaf(r, f, 1.0000001f);
  r = __builtin_amdgcn_ldexpf(r, (int)j);
  r = e == __uint_as_float(0x7F800000u) ? 0.0f : r;
  return e != e ? e : r;
}

// Cash-Karp coefficients (src/rk5.cu:23-60), formed in double and rounded
// once to float.
struct CK {
  static constexpr float a21 = (float)0.2;
  static constexpr float a31 = (float)(3.0 / 40.0), a32 = (float)(9.0 / 40.0);
  static constexpr float a41 = (float)0.3, a42 = (float)-0.9, a43 = (float)1.2;
  static constexpr float a51 = (float)(-11.0 / 54.0), a52 = (float)2.5,
                         a53 = (float)(-70.0 / 27.0),
                         a54 = (float)(35.0 / 27.0);
  static constexpr float a61 = (float)(1631.0 / 55296.0),
                         a62 = (float)(175.0 / 512.0),
                         a63 = (float)(575.0 / 13824.0),
                         a64 = (float)(44275.0 / 110592.0),
                         a65 = (float)(253.0 / 4096.0);
  static constexpr float a71 = (float)(37.0 / 378.0),
                         a73 = (float)(250.0 / 621.0),
                         a74 = (float)(125.0 / 594.0),
                         a76 = (float)(512.0 / 1771.0);
  static constexpr float e1 = (float)(37.0 / 378.0 - 2825.0 / 27648.0);
  static constexpr float e3 = (float)(250.0 / 621.0 - 18575.0 / 48384.0);
  static constexpr float e4 = (float)(125.0 / 594.0 - 13525.0 / 55296.0);
  static constexpr float e5 = (float)(-277.00 / 14336.0);
  static constexpr float e6 = (float)(512.0 / 1771.0 - 0.25);
  static constexpr float coeff = (float)0.9;
  static constexpr float err_min = (float)1.889568e-4;
  static constexpr float scal_min = (float)0.1;
};

// Threefry2x32-20 (Salmon, Moraes, Dror, Shaw, SC'11: the Threefish
// rounds of Skein, rotations 13 15 26 6 17 29 16 24, key injection every
// four rounds with the parity word 0x1BD11BDA); returns word 0.  Adds,
// rotations and xors only: the Poisson draw of every neuron-step uses it,
// where Philox's 32-bit multiplies (quarter rate) would cost more.
__device__ __forceinline__ uint32_t dev_threefry2x32_w0(uint32_t c0,
                                                        uint32_t c1,
                                                        uint32_t k0,
                                                        uint32_t k1) {
  const uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
  constexpr int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
  for (int r = 0; r < 20; r++) {
    x0 += x1;
    x1 = (x1 << rot[r & 7]) | (x1 >> (32 - rot[r & 7]));
    x1 ^= x0;
    if ((r & 3) == 3) {
      const uint32_t s = (uint32_t)(r + 1) >> 2;
      x0 += ks[s % 3];
      x1 += ks[(s + 1) % 3] + s;
    }
  }
  return x0;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11); returns word 0.
__device__ __forceinline__ uint32_t dev_philox_w0(uint32_t c0, uint32_t c1,
                                                  uint32_t c2, uint32_t c3,
                                                  uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t lo0 = 0xD2511F53u * c0;
    uint32_t hi0 = __umulhi(0xD2511F53u, c0);
    uint32_t lo1 = 0xCD9E8D57u * c2;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c0;
}

// Synapse rules (src/syn_model.cu:34-46 dispatch):
//   type 1 test_syn_model (src/test_syn_model.cu:24-29): w += offset+fact*Dt
//   type 2 stdp           (src/stdp.cu:24-51), evaluated in double
// param8 = tau_plus, tau_minus, lambda, alpha, mu_plus, mu_minus, Wmax, -
//          (test_syn_model: fact, offset in the first two slots)
__device__ __forceinline__ float dev_syn_update(int type, float wf, float Dt,
                                                const float *param8) {
  if (type == 2) {
    double tau_plus = param8[0], tau_minus = param8[1], lambda = param8[2];
    double alpha = param8[3], mu_plus = param8[4], mu_minus = param8[5];
    double Wmax = param8[6];
    double w = wf, w1;
    if (Dt >= 0) {
      double fact = lambda * exp(-(double)Dt / tau_plus);
      w1 = w + fact * Wmax * pow(1.0 - w / Wmax, mu_plus);
    } else {
      double fact = -alpha * lambda * exp((double)Dt / tau_minus);
      w1 = w + fact * Wmax * pow(w / Wmax, mu_minus);
    }
    w1 = w1 > 0.0 ? w1 : 0.0;
    w1 = w1 < Wmax ? w1 : Wmax;
    return (float)w1;
  }
  if (type == 1) {
    float fact = param8[0], offset = param8[1];
    return wf + (offset + fact * Dt);
  }
  return wf;
}

__device__ __forceinline__ float pf(const PField &f, int i) {
  return f.ptr ? f.ptr[i] : f.val;
}

}  // namespace ngpu
