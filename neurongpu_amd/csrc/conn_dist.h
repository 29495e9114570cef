// Per-connection weights and delays drawn from a distribution, with the
// same value on the host and on the device: one counter-based
// Philox4x32-10 block per (connection, attempt), FP32 arithmetic with no
// contraction (both sides compile with -ffp-contract=off), correctly
// rounded division and square root (IEEE on the host; HIP's default on the
// device), and a logarithm of our own (no libm call whose last bit could
// differ between the two).
//
// Distributions (the reference's Python draws "normal" and
// "normal_clipped" into host arrays, pythonlib/neurongpu.py:1472-1514,
// src/neurongpu.cu:1111-1154; its SynSpec reserves weight_distr_ /
// delay_distr_ for this, src/connect_spec.h:70-73, src/connect_rules.h:55):
//   uniform         low + (high - low) u, u in (0, 1]; with step > 0 the
//                   values low + step j, j uniform in {0 .. n-1},
//                   n = floor((high - low) / step + 0.5) + 1
//   normal          mu + sigma z, z standard normal (Marsaglia's polar
//                   method: pairs (u, v) uniform in (-1, 1] until
//                   0 < s = u^2 + v^2 < 1, z = u sqrt(-2 ln s / s))
//   normal_clipped  normal values outside [low, high] are redrawn; with
//                   step > 0 rounded to low + step round((x - low) / step)
//                   (the reference's vstep rounding)
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define NGPU_HD __host__ __device__
#else
#define NGPU_HD
#endif

namespace ngpu {

enum DistKind {
  kDistNone = 0,
  kDistUniform = 1,
  kDistNormal = 2,
  kDistNormalClipped = 3,
};

// Pairs tried per value before a normal_clipped draw gives up and returns
// the clamped mean (an interval with so little mass is rejected when the
// connection is made).
constexpr uint32_t kDistMaxBlocks = 4096;

struct DistSpec {
  int kind;
  float mu, sigma, low, high, step;
  uint32_t k0, k1, call;   // stream key and call index
};

NGPU_HD inline void dist_philox(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// (0, 1] from the top 24 bits of a word
NGPU_HD inline float dist_u01(uint32_t u) {
  return ((float)(u >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// ln x for 0 < x < 1 (polar-method argument): x = 2^e m with m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m - 1) / (m + 1)) by its odd series
// to the ninth power, e ln 2 with ln 2 split in two words.
NGPU_HD inline float dist_logf(float x) {
  union {
    float f;
    uint32_t u;
  } b;
  b.f = x;
  int e = (int)((b.u >> 23) & 0xFFu) - 127;
  b.u = (b.u & 0x007FFFFFu) | 0x3F800000u;
  float m = b.f;
  if (m > 1.41421354f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float q = 1.11111112e-01f;
  q = fmaf(q, s2, 1.42857149e-01f);
  q = fmaf(q, s2, 2.00000003e-01f);
  q = fmaf(q, s2, 3.33333343e-01f);
  const float two_s = 2.0f * s;
  const float lm = fmaf(two_s * s2, q, two_s);
  const float ef = (float)e;
  return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.42860677e-06f, lm));
}

// Value of connection k (insertion index within its Connect call).
NGPU_HD inline float dist_sample(const DistSpec &d, unsigned long long k) {
  const uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
  if (d.kind == kDistUniform) {
    uint32_t c[4] = {lo, hi, d.call, 0x756E6966u};
    dist_philox(c, d.k0, d.k1);
    if (d.step > 0.0f) {
      const uint32_t n =
          (uint32_t)floorf((d.high - d.low) / d.step + 0.5f) + 1u;
      const uint32_t j = (uint32_t)(((uint64_t)c[0] * n) >> 32);
      return d.low + d.step * (float)j;
    }
    return d.low + (d.high - d.low) * dist_u01(c[0]);
  }
  for (uint32_t a = 0; a < kDistMaxBlocks; a++) {
    uint32_t c[4] = {lo, hi, d.call, 0x4E000000u | a};
    dist_philox(c, d.k0, d.k1);
    for (int p = 0; p < 2; p++) {
      const float u = 2.0f * dist_u01(c[2 * p]) - 1.0f;
      const float v = 2.0f * dist_u01(c[2 * p + 1]) - 1.0f;
      const float s = u * u + v * v;
      if (!(s < 1.0f) || s == 0.0f) continue;
      const float z = u * sqrtf(-2.0f * dist_logf(s) / s);
      float x = d.mu + d.sigma * z;
      if (d.kind == kDistNormalClipped && (x < d.low || x > d.high)) continue;
      if (d.step > 0.0f) x = d.low + d.step * roundf((x - d.low) / d.step);
      return x;
    }
  }
  const float x = d.mu < d.low ? d.low : d.mu;
  return x > d.high ? d.high : x;
}

// A value source of a Connect call: the caller's array, a distribution, or
// one value for every connection.
struct ValSrc {
  const float *arr;
  DistSpec dist;
  float v0;
  NGPU_HD float at(unsigned long long k) const {
    return arr ? arr[k] : (dist.kind != kDistNone ? dist_sample(dist, k) : v0);
  }
};

}  // namespace ngpu
