// Host-side check of the per-connection distributions (conn_dist.h), built
// with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize.py: every kind over many connection indices and two
// call indices, values inside their bounds, finite, and on the step grid
// where a step is given.  Exit status 0 on success.
#include <cmath>
#include <cstdio>

#include "conn_dist.h"

using ngpu::DistSpec;

static int check(const DistSpec &d, const char *name, unsigned long long n) {
  double sum = 0.0;
  for (unsigned long long k = 0; k < n; k++) {
    const float x = ngpu::dist_sample(d, k * 2654435761ull);
    if (!std::isfinite(x)) {
      std::printf("%s: non-finite value at %llu\n", name, k);
      return 1;
    }
    if (d.kind != ngpu::kDistNormal && (x < d.low || x > d.high)) {
      std::printf("%s: %g outside [%g, %g]\n", name, x, d.low, d.high);
      return 1;
    }
    if (d.step > 0.0f) {
      const float j = (x - d.low) / d.step;
      if (std::fabs(j - std::nearbyint(j)) > 1e-3f) {
        std::printf("%s: %g off the step grid\n", name, x);
        return 1;
      }
    }
    sum += x;
  }
  std::printf("%s: mean %.5f over %llu\n", name, sum / (double)n, n);
  return 0;
}

int main() {
  int bad = 0;
  for (unsigned call = 0; call < 2; call++) {
    DistSpec u{ngpu::kDistUniform, 0, 0, 0.1f, 2.0f, 0.0f, 1234u, 0u, call};
    DistSpec us{ngpu::kDistUniform, 0, 0, 0.1f, 2.0f, 0.1f, 1234u, 0u, call};
    DistSpec nn{ngpu::kDistNormal, 1.5f, 0.75f, 0, 0, 0.0f, 99u, 7u, call};
    DistSpec nc{ngpu::kDistNormalClipped, 1.5f, 0.75f, 0.1f, 3.0f, 0.0f,
                5u, 6u, call};
    DistSpec ncs{ngpu::kDistNormalClipped, 0.8f, 0.4f, 0.1f, 2.0f, 0.1f,
                 5u, 6u, call};
    // a narrow interval in the tail: many redrawn pairs per value
    DistSpec tail{ngpu::kDistNormalClipped, 0.0f, 1.0f, 3.0f, 3.5f, 0.0f,
                  11u, 12u, call};
    bad |= check(u, "uniform", 200000);
    bad |= check(us, "uniform step", 200000);
    bad |= check(nn, "normal", 200000);
    bad |= check(nc, "normal_clipped", 200000);
    bad |= check(ncs, "normal_clipped step", 200000);
    bad |= check(tail, "normal_clipped tail", 20000);
  }
  return bad;
}
