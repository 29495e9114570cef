// Device code shared by the per-timestep kernel files: spike append,
// input fold, step values from the step clock, and the adaptive Cash-Karp
// RK5 attempt generic over the neuron model.
#pragma once
#include <hip/hip_runtime.h>

#include "ngpu_device.h"
#include "ngpu_kernels.h"

namespace ngpu {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Step-dependent values of a node-update launch: from the arguments, or
// from the step clock when the launch has one (captured steps).
struct StepNow {
  uint32_t *ring_src;   // ring slot of this step
  float *ring_h;
  uint32_t *ring_cnt;
  uint32_t *ring_n;
  long long T;
  double t;             // neural time at the end of the step
  long long pois_T;     // pending Poisson draw folded at the start
  double pois_t;
};

__device__ __forceinline__ StepNow step_now(const SpikeOut &o, const FoldIn *f,
                                            const StepClock *c) {
  StepNow s;
  if (c) {
    const long long slot = c->gs % o.R;
    s.ring_src = o.ring_src0 + slot * o.ring_cap;
    s.ring_h = o.ring_h0 + slot * o.ring_cap;
    s.ring_cnt = o.ring_cnt0 + slot * o.ring_cap;
    s.ring_n = o.ring_n0 + slot;
    s.T = c->T;
    s.t = c->t;
    s.pois_T = c->pois_T;
    s.pois_t = c->pois_t;
  } else {
    s.ring_src = o.ring_src;
    s.ring_h = o.ring_h;
    s.ring_cnt = o.ring_cnt;
    s.ring_n = o.ring_n;
    s.T = o.T;
    s.t = o.t;
    s.pois_T = f ? f->pois_T : -1;
    s.pois_t = f ? f->pois_t : 0.0;
  }
  return s;
}

// Wavefront-aggregated append of this lane's spikes to the ring slot of the
// step, plus the per-node bookkeeping of PushSpike.
__device__ __forceinline__ void emit_spikes(const SpikeOut &o,
                                            const StepNow &sn, bool active,
                                            int i, int node, int nsp,
                                            float height) {
  bool spk = active && nsp > 0;
  unsigned long long mask = __ballot(spk);
  if (mask == 0ull) return;
  uint32_t leader = (uint32_t)(__ffsll((long long)mask) - 1);
  uint32_t base = 0;
  if (lane_id() == leader)
    base = atomicAdd(sn.ring_n, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  if (!spk) return;
  uint32_t pos = base + lanes_below(mask);
  sn.ring_src[pos] = (uint32_t)node;
  sn.ring_h[pos] = height;
  sn.ring_cnt[pos] = (uint32_t)nsp;
  o.last_spike[node] = sn.T;
  o.last_height[node] = height / (float)nsp;
  if (o.den_idx[node] == 0) o.last_rev[node] = sn.T;
  if (o.spike_count) o.spike_count[i] += nsp;
  if (o.rec_times) {
    int nr = o.n_rec_times[i];
    for (int s = 0; s < nsp; s++) {
      if (nr >= o.max_rec_times - 1) break;
      o.rec_times[(long long)i * o.max_rec_times + nr] = (float)sn.t;
      nr++;
    }
    o.n_rec_times[i] = nr;
  }
}

// Poisson count of a uniform draw u: the number of the table's
// nondecreasing thresholds below u (the table ends with 0xFFFFFFFF and is
// padded with it to kPoissonTableMax entries).  The first 16 thresholds
// are compared at once, with no load depending on a comparison; a longer
// table continues in a loop.
__device__ __forceinline__ int pois_invert(const uint32_t *tab, uint32_t u) {
  int k = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) k += (u > tab[i]) ? 1 : 0;
  if (k == 16)
    while (u > tab[k]) k++;
  return k;
}

// Poisson count of one direct connection (canonical index j, generator g)
// for the pending step: 0 outside the generator's window
// (src/poiss_gen.cu:43-74), else one Threefry draw (counter (j, T mod
// 2^32), key (seed_lo ^ tag, seed_hi + T / 2^32)) inverted over the
// generator's CDF table.
__device__ __forceinline__ int pois_count(const DirectIn &d, uint32_t j,
                                          uint32_t g, float delay,
                                          uint32_t Tlo, uint32_t Thi,
                                          double pois_t) {
  double t_rel = pois_t - (double)d.gen_origin[g] - (double)delay;
  if (!(t_rel >= (double)d.gen_start[g] && t_rel <= (double)d.gen_stop[g]))
    return 0;
  uint32_t ti = d.gen_table[g];
  if (ti == 0xFFFFFFFFu) return 0;
  const uint32_t *tab = d.tables + (size_t)ti * kPoissonTableMax;
  uint32_t u = dev_threefry2x32_w0(j, Tlo, d.key0 ^ kPoissonTag,
                                   d.key1 + Thi);
  return pois_invert(tab, u);
}

// Generator of the compact form's connection j.
__device__ __forceinline__ uint32_t compact_gen(const DirectIn &d,
                                                uint32_t j) {
  return d.gen_mode == 0 ? d.gen0
                         : (d.gen_mode == 1 ? d.gen0 + j : d.gen_of[j]);
}

// Port input of the previous step: delivery accumulator of every port plus
// the Poisson draws of the node's direct connections (canonical order).
// An accumulator entry that is already 0 is not written back.
template <int P>
__device__ __forceinline__ void gather_input(const FoldIn &f,
                                             const StepNow &sn, int node,
                                             float (&a)[P]) {
#pragma unroll
  for (int p = 0; p < P; p++) {
    float &ap = at(f.acc + (long long)p * f.n_total, (uint32_t)node);
    a[p] = ap;
    if (a[p] != 0.0f) ap = 0.0f;
  }
  if (sn.pois_T < 0 || !f.dir.active) return;
  const DirectIn &d = f.dir;
  uint32_t Tlo = (uint32_t)sn.pois_T, Thi = (uint32_t)((uint64_t)sn.pois_T >> 32);
  if (d.compact) {
    const uint32_t j = (uint32_t)node - d.lo;
    if (j >= d.hi - d.lo) return;
    // one generator for all (its parameters and table are then read with
    // scalar loads), or one per node
    const int k = d.gen_mode == 0
                      ? pois_count(d, j, d.gen0, d.c_delay, Tlo, Thi,
                                   sn.pois_t)
                      : pois_count(d, j, compact_gen(d, j), d.c_delay, Tlo,
                                   Thi, sn.pois_t);
    if (k > 0) {
      float add = d.c_weight * (float)k;
#pragma unroll
      for (int p = 0; p < P; p++)
        if (d.c_port == p) a[p] = a[p] + add;
    }
    return;
  }
  uint32_t j0 = d.ptr[node], j1 = d.ptr[node + 1];
  for (uint32_t j = j0; j < j1; j++) {
    const int k = pois_count(d, d.canon[j], d.gen[j], d.delay[j], Tlo, Thi,
                             sn.pois_t);
    if (k > 0) {
      float add = d.weight[j] * (float)k;
      int port = d.port[j];
#pragma unroll
      for (int p = 0; p < P; p++)
        if (port == p) a[p] = a[p] + add;
    }
  }
}

// One port of gather_input, for models whose port count is known only at
// run time: the accumulator of port p, then the Poisson draws of the node's
// direct connections to that port in canonical order.
__device__ __forceinline__ float gather_port(const FoldIn &f, long long pois_T,
                                             double pois_t, int node, int p) {
  float &ap = at(f.acc + (long long)p * f.n_total, (uint32_t)node);
  float a = ap;
  if (a != 0.0f) ap = 0.0f;
  if (pois_T < 0 || !f.dir.active) return a;
  const DirectIn &d = f.dir;
  uint32_t Tlo = (uint32_t)pois_T, Thi = (uint32_t)((uint64_t)pois_T >> 32);
  if (d.compact) {
    const uint32_t j = (uint32_t)node - d.lo;
    if (j >= d.hi - d.lo || d.c_port != p) return a;
    const int k =
        pois_count(d, j, compact_gen(d, j), d.c_delay, Tlo, Thi, pois_t);
    if (k > 0) a = a + d.c_weight * (float)k;
    return a;
  }
  uint32_t j0 = d.ptr[node], j1 = d.ptr[node + 1];
  for (uint32_t j = j0; j < j1; j++) {
    if ((int)d.port[j] != p) continue;
    const int k = pois_count(d, d.canon[j], d.gen[j], d.delay[j], Tlo, Thi,
                             pois_t);
    if (k > 0) a = a + d.weight[j] * (float)k;
  }
  return a;
}

// Start of an adaptive Cash-Karp step (src/rk5.h:57-69) of model M (M::N
// state variables, m.rhs(refr, y, dy) the derivatives): first stage and
// the error scale |y| + |k1 h| + scal_min, which stay fixed over the
// step's retries.
template <class M>
__device__ __forceinline__ void rk5_begin(const M &m, float refr,
                                          const float (&y)[M::N], float h,
                                          float (&k1)[M::N],
                                          float (&scal)[M::N]) {
  constexpr int N = M::N;
  m.rhs(refr, y, k1);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float s = fabsf(y[i]) + fabsf(k1[i] * h);
    scal[i] = s + CK::scal_min;
  }
}

// One attempt of the step with size hh (src/rk5.h:70-135): stages 2-6,
// the fifth-order solution in yn and the error norm relative to the
// tolerance, max_i |hh e_i| / (scal_i eps).  The largest quotient is found
// by cross-multiplication (|hh e_i| scal_m against |hh e_m| scal_i, a NaN
// taking the place, as a NaN quotient does in the reference's running
// maximum), so that only the winner is divided, once.
template <class M>
__device__ __forceinline__ float rk5_attempt(
    const M &m, float refr, const float (&y)[M::N], const float (&k1)[M::N],
    const float (&scal)[M::N], float hh, float eps, float (&yn)[M::N]) {
  constexpr int N = M::N;
  float k2[N], k3[N], k4[N], k5[N], k6[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a21 * k1[i];
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  m.rhs(refr, yn, k2);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a31 * k1[i];
    t = __builtin_fmaf(CK::a32, k2[i], t);
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  m.rhs(refr, yn, k3);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a41 * k1[i];
    t = __builtin_fmaf(CK::a42, k2[i], t);
    t = __builtin_fmaf(CK::a43, k3[i], t);
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  m.rhs(refr, yn, k4);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a51 * k1[i];
    t = __builtin_fmaf(CK::a52, k2[i], t);
    t = __builtin_fmaf(CK::a53, k3[i], t);
    t = __builtin_fmaf(CK::a54, k4[i], t);
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  m.rhs(refr, yn, k5);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a61 * k1[i];
    t = __builtin_fmaf(CK::a62, k2[i], t);
    t = __builtin_fmaf(CK::a63, k3[i], t);
    t = __builtin_fmaf(CK::a64, k4[i], t);
    t = __builtin_fmaf(CK::a65, k5[i], t);
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  m.rhs(refr, yn, k6);
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::a71 * k1[i];
    t = __builtin_fmaf(CK::a73, k3[i], t);
    t = __builtin_fmaf(CK::a74, k4[i], t);
    t = __builtin_fmaf(CK::a76, k6[i], t);
    yn[i] = __builtin_fmaf(hh, t, y[i]);
  }
  float num = 0.0f, den = 1.0f;
#pragma unroll
  for (int i = 0; i < N; i++) {
    float t = CK::e1 * k1[i];
    t = __builtin_fmaf(CK::e3, k3[i], t);
    t = __builtin_fmaf(CK::e4, k4[i], t);
    t = __builtin_fmaf(CK::e5, k5[i], t);
    t = __builtin_fmaf(CK::e6, k6[i], t);
    const float v = fabsf(hh * t);
    const bool take = !(v * den <= num * scal[i]);
    num = take ? v : num;
    den = take ? scal[i] : den;
  }
  return num / (den * eps);
}

// An attempt accepted with a scaled error above 1, which the acceptance
// test allows only at the minimum step size (rare: one atomic each).
__device__ __forceinline__ void count_forced(unsigned long long *forced) {
  if (forced) atomicAdd(forced + (blockIdx.x & (kCounterSlots - 1)), 1ull);
}

// One adaptive Cash-Karp step (src/rk5.h:57-147).  Returns false when the
// attempt cap is reached.
template <class M>
__device__ __forceinline__ bool rk5_step(const M &m, float refr, double &x,
                                         float (&y)[M::N], float &h,
                                         float h_min, double h_min_acc,
                                         float h_max, float eps, int &n_att,
                                         int &n_acc, int &n_rej,
                                         unsigned long long *forced) {
  constexpr int N = M::N;
  float k1[N], yn[N], scal[N];
  float err = 0.0f;
  rk5_begin(m, refr, y, h, k1, scal);
  for (;;) {
    float hh = h;
    if (hh > h_max) hh = h_max;
    if (hh < h_min) hh = h_min;
    h = hh;
    err = rk5_attempt(m, refr, y, k1, scal, hh, eps, yn);
    n_att++;
    if (err <= 1.0f || (double)hh <= h_min_acc) break;
    n_rej++;
    if (n_att > kMaxAttemptsPerStep) return false;
    float h_new = (hh * CK::coeff) * dev_pow_ctrl(err, -0.25f);
    float h_lo = 0.1f * hh;
    h = (h_new > h_lo) ? h_new : h_lo;
  }
  n_acc++;
  if (err > 1.0f) count_forced(forced);
  x += (double)h;
  if (err > CK::err_min) {
    h = (h * CK::coeff) * dev_pow_ctrl(err, -0.2f);
  } else {
    h = 5.0f * h;
  }
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = yn[i];
  return true;
}

// Accepted and rejected RK5 attempts of the wavefront, added to one of
// kCounterSlots counter pairs.
__device__ __forceinline__ void rk5_count(unsigned long long *cnt, int n_acc,
                                          int n_rej) {
  int sa = n_acc, sr = n_rej;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off);
    sr += __shfl_xor(sr, off);
  }
  if (lane_id() == 0 && cnt && (sa | sr)) {
    unsigned long long *c = cnt + 2 * (blockIdx.x & (kCounterSlots - 1));
    atomicAdd(c, (unsigned long long)sa);
    atomicAdd(c + 1, (unsigned long long)sr);
  }
}

}  // namespace
}  // namespace ngpu
