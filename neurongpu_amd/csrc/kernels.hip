// Per-timestep kernels for gfx950 (MI355X).
//
// Step sequence launched by the engine (one HIP stream, no host syncs):
//   node updates   one launch per node group.  Each first folds the port
//                  input of the previous step (delivery accumulators plus
//                  the Poisson draws of that step) into its state, then
//                  advances the state by dt and appends its spikes to the
//                  spike ring slot of this step (wavefront ballot + one
//                  atomic per wavefront).
//   record         gathers recorded variables into a device row buffer
//   delivery       flattened CSR-by-source gather over all (spike, delay)
//                  pairs due this step, float atomics into acc[port][node]
//   post STDP      (plastic networks only) reverse-index sweep
//
// Reference behaviour followed (file:line in the NeuronGPU sources):
//   fold            src/get_spike.cu:116-140, src/poiss_gen.cu:43-74
//   RK5             src/rk5.h:57-195, src/rk5.cu:23-60
//   AdEx RHS/reset  src/aeif_cond_beta_kernel.h:153-214,
//                   src/aeif_cond_alpha_kernel.h:149-211
//   iaf_psc_exp     src/iaf_psc_exp.cu:94-119
//   PushSpike       src/spike_buffer.cu:143-228
//   delivery        src/spike_buffer.cu:233-302, src/nested_loop.cu:96-167,
//                   src/get_spike.cu:48-85
//   post STDP       src/rev_spike.cu:41-75, src/stdp.cu:24-51
#include <hip/hip_runtime.h>

#include "ngpu_device.h"
#include "ngpu_kernels.h"
#include "proposer.h"
#include "step_common.h"

namespace ngpu {

namespace {

// ---------------------------------------------------------------------
// AdEx conductance neurons (aeif_cond_beta / aeif_cond_alpha), RK5.
// ---------------------------------------------------------------------
template <int P>
struct AeifLocal {
  float V_th, Delta_T, inv_Delta_T, g_L, E_L, inv_C_m, a, b, inv_tau_w, I_e,
      V_peak, V_reset, t_ref;
  float E_rev[P], inv_tau_rise[P], inv_tau_decay[P], g0[P];
};

template <int P>
__device__ __forceinline__ void aeif_rhs(const AeifLocal<P> &p, float refr,
                                         const float (&y)[2 + 2 * P],
                                         float (&dy)[2 + 2 * P]) {
  bool rf = refr > 0.0f;
  float V = rf ? p.V_reset : ((y[0] < p.V_peak) ? y[0] : p.V_peak);
  // refractory: dV = 0, by a zero factor rather than a branch (a wave never
  // splits on the refractory state); (t/C_m)*0 and t*0 agree bit for bit
  const float inv_C_m = rf ? 0.0f : p.inv_C_m;
  float I_syn = 0.0f;
#pragma unroll
  for (int i = 0; i < P; i++)
    I_syn = __builtin_fmaf(y[2 + 2 * i], p.E_rev[i] - V, I_syn);
  float V_spike = p.Delta_T * dev_expf((V - p.V_th) * p.inv_Delta_T);
  float t = (V - p.E_L) - V_spike;
  t = (-p.g_L) * t;
  t = t + I_syn;
  t = t - y[1];
  t = t + p.I_e;
  dy[0] = t * inv_C_m;
  float u = p.a * (V - p.E_L);
  u = u - y[1];
  dy[1] = u * p.inv_tau_w;
#pragma unroll
  for (int i = 0; i < P; i++) {
    float g = y[2 + 2 * i], g1 = y[3 + 2 * i];
    dy[3 + 2 * i] = (-g1) * p.inv_tau_rise[i];
    dy[2 + 2 * i] = __builtin_fmaf(-g, p.inv_tau_decay[i], g1);
  }
}

// The AdEx RHS as a model of the generic RK5 attempt.
template <int P>
struct AeifModel {
  static constexpr int N = 2 + 2 * P;
  const AeifLocal<P> &p;
  __device__ __forceinline__ void rhs(float refr, const float (&y)[N],
                                      float (&dy)[N]) const {
    aeif_rhs<P>(p, refr, y, dy);
  }
};

template <int P>
__device__ __forceinline__ void aeif_rk5_begin(const AeifLocal<P> &p,
                                               float refr,
                                               const float (&y)[2 + 2 * P],
                                               float h,
                                               float (&k1)[2 + 2 * P],
                                               float (&scal)[2 + 2 * P]) {
  rk5_begin(AeifModel<P>{p}, refr, y, h, k1, scal);
}

template <int P>
__device__ __forceinline__ float aeif_rk5_attempt(
    const AeifLocal<P> &p, float refr, const float (&y)[2 + 2 * P],
    const float (&k1)[2 + 2 * P], const float (&scal)[2 + 2 * P],
    float hh, float eps, float (&yn)[2 + 2 * P]) {
  return rk5_attempt(AeifModel<P>{p}, refr, y, k1, scal, hh, eps, yn);
}

template <int P>
__device__ __forceinline__ bool aeif_rk5_step(const AeifLocal<P> &p, float refr,
                                              double &x, float (&y)[2 + 2 * P],
                                              float &h, float h_min,
                                              double h_min_acc, float h_max,
                                              float eps, int &n_att,
                                              int &n_acc, int &n_rej,
                                              unsigned long long *forced) {
  return rk5_step(AeifModel<P>{p}, refr, x, y, h, h_min, h_min_acc, h_max,
                  eps, n_att, n_acc, n_rej, forced);
}

template <int P>
__global__ __launch_bounds__(kBlock) void aeif_kernel(AeifArgs A) {
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0;
  int n_acc = 0, n_rej = 0;
  if (active) {
    constexpr int N = 2 + 2 * P;
    AeifLocal<P> p;
    p.V_th = pf(A.p.V_th, i);
    p.Delta_T = pf(A.p.Delta_T, i);
    p.inv_Delta_T = pf(A.p.inv_Delta_T, i);
    p.g_L = pf(A.p.g_L, i);
    p.E_L = pf(A.p.E_L, i);
    p.inv_C_m = pf(A.p.inv_C_m, i);
    p.a = pf(A.p.a, i);
    p.b = pf(A.p.b, i);
    p.inv_tau_w = pf(A.p.inv_tau_w, i);
    p.I_e = pf(A.p.I_e, i);
    p.V_peak = pf(A.p.V_peak, i);
    p.V_reset = pf(A.p.V_reset, i);
    p.t_ref = pf(A.p.t_ref, i);
#pragma unroll
    for (int q = 0; q < P; q++) {
      p.E_rev[q] = pf(A.p.E_rev[q], i);
      p.inv_tau_rise[q] = pf(A.p.inv_tau_rise[q], i);
      p.inv_tau_decay[q] = pf(A.p.inv_tau_decay[q], i);
      p.g0[q] = pf(A.p.g0[q], i);
    }
    float y[N];
    y[0] = A.V[i];
    y[1] = A.w[i];
#pragma unroll
    for (int q = 0; q < P; q++) {
      y[2 + 2 * q] = A.g[q][i];
      y[3 + 2 * q] = A.g1[q][i];
    }
    // fold the previous step's port input into g1
    float in[P];
    gather_input<P>(A.fold, sn, node, in);
#pragma unroll
    for (int q = 0; q < P; q++)
      y[3 + 2 * q] = __builtin_fmaf(in[q], p.g0[q], y[3 + 2 * q]);
    if (A.do_step) {
      double x = A.x[i];
      float h = A.h[i];
      float refr = A.refr[i];
      double x1 = A.clock ? sn.t : A.x1;
      int n_att = 0;
      bool end = false;
      while (!end) {
        float hmax = (float)(x1 - x);
        if (!aeif_rk5_step<P>(p, refr, x, y, h, A.h_min, A.h_min_acc, hmax,
                              A.eps, n_att, n_acc, n_rej, A.forced))
          x = x1;
        end = (x >= x1 - (double)A.h_min);
        // ExternalUpdate
        if (y[0] < -1.0e3f) {
          y[0] = p.V_reset;
          y[1] = 0.0f;
          continue;
        }
        if (y[1] < -1.0e6f || y[1] > 1.0e6f) {
          y[0] = p.V_reset;
          y[1] = 0.0f;
          continue;
        }
        if (refr > 0.0f) {
          y[0] = p.V_reset;
          if (end) refr -= 1.0f;
        } else if (y[0] >= p.V_peak) {
          nsp++;
          y[0] = p.V_reset;
          y[1] = y[1] + p.b;
          float r = roundf(p.t_ref / A.dt);
          int ri = (int)r;
          refr = (float)(ri < 0 ? 0 : ri);
        }
      }
      A.x[i] = x;
      A.h[i] = h;
      A.refr[i] = refr;
    }
    A.V[i] = y[0];
    A.w[i] = y[1];
#pragma unroll
    for (int q = 0; q < P; q++) {
      A.g[q][i] = y[2 + 2 * q];
      A.g1[q][i] = y[3 + 2 * q];
    }
  }
  if (!A.do_step) return;
  emit_spikes(A.out, sn, active, ii, node, nsp, (float)nsp);
  // statistics: accepted / rejected RK5 attempts
  int sa = n_acc, sr = n_rej;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off);
    sr += __shfl_xor(sr, off);
  }
  if (lane_id() == 0 && A.cnt && (sa | sr)) {
    unsigned long long *c = A.cnt + 2 * (blockIdx.x & (kCounterSlots - 1));
    atomicAdd(c, (unsigned long long)sa);
    atomicAdd(c + 1, (unsigned long long)sr);
  }
}

// UNI: every parameter of the group is uniform (PField::ptr == nullptr),
// so the values come straight from the kernel arguments
template <int P, bool UNI>
__device__ __forceinline__ void aeif_load_params(const AeifArgs &A, int i,
                                                 AeifLocal<P> &p) {
  if (UNI) {
    p.V_th = A.p.V_th.val;
    p.Delta_T = A.p.Delta_T.val;
    p.inv_Delta_T = A.p.inv_Delta_T.val;
    p.g_L = A.p.g_L.val;
    p.E_L = A.p.E_L.val;
    p.inv_C_m = A.p.inv_C_m.val;
    p.a = A.p.a.val;
    p.b = A.p.b.val;
    p.inv_tau_w = A.p.inv_tau_w.val;
    p.I_e = A.p.I_e.val;
    p.V_peak = A.p.V_peak.val;
    p.V_reset = A.p.V_reset.val;
    p.t_ref = A.p.t_ref.val;
#pragma unroll
    for (int q = 0; q < P; q++) {
      p.E_rev[q] = A.p.E_rev[q].val;
      p.inv_tau_rise[q] = A.p.inv_tau_rise[q].val;
      p.inv_tau_decay[q] = A.p.inv_tau_decay[q].val;
      p.g0[q] = A.p.g0[q].val;
    }
    return;
  }
  p.V_th = pf(A.p.V_th, i);
  p.Delta_T = pf(A.p.Delta_T, i);
  p.inv_Delta_T = pf(A.p.inv_Delta_T, i);
  p.g_L = pf(A.p.g_L, i);
  p.E_L = pf(A.p.E_L, i);
  p.inv_C_m = pf(A.p.inv_C_m, i);
  p.a = pf(A.p.a, i);
  p.b = pf(A.p.b, i);
  p.inv_tau_w = pf(A.p.inv_tau_w, i);
  p.I_e = pf(A.p.I_e, i);
  p.V_peak = pf(A.p.V_peak, i);
  p.V_reset = pf(A.p.V_reset, i);
  p.t_ref = pf(A.p.t_ref, i);
#pragma unroll
  for (int q = 0; q < P; q++) {
    p.E_rev[q] = pf(A.p.E_rev[q], i);
    p.inv_tau_rise[q] = pf(A.p.inv_tau_rise[q], i);
    p.inv_tau_decay[q] = pf(A.p.inv_tau_decay[q], i);
    p.g0[q] = pf(A.p.g0[q], i);
  }
}

// Resume record of a neuron whose step did not finish within the attempt
// budget of the first pass.  Its state (y, x, current h, refractory count)
// is in the group arrays; h_begin is the step size the current RK5 step
// started with, from which the second pass recomputes the step's first
// stage and error scale exactly.
// Learned step-size proposer (inference): features of a neuron at the
// start of an RK5 step and the MLP with BF16 weights, FP32 accumulation.
template <int P>
__device__ __forceinline__ void prop_features(const AeifLocal<P> &p,
                                              const float (&y)[2 + 2 * P],
                                              float h, float err_last,
                                              float inv_dt,
                                              float (&f)[kPropIn]) {
  f[0] = (y[0] - p.E_L) * 0.1f;
  f[1] = y[1] * 0.01f;
  f[2] = y[2];
  float gi = 0.0f;
#pragma unroll
  for (int q = 1; q < P; q++) gi += y[2 + 2 * q];
  f[3] = gi;
  f[4] = __log2f(h * inv_dt);
  float e = err_last > 1e-8f ? err_last : 1e-8f;
  f[5] = __log2f(e) * (0.30103f * 0.25f);
}

typedef short prop_s4 __attribute__((ext_vector_type(4)));
typedef float prop_f16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short prop_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

// The MLP for the 64 samples of a wavefront (one per lane) on the matrix
// cores: H1^T = W1 X^T and H2^T = W2 H1^T as 32x32x8 BF16 MFMAs with FP32
// accumulation, two 32-sample column blocks each; the accumulator layout of
// a layer is the B-operand layout of the next, so only the input features
// cross lanes.  Every lane of the wave must call it (wave-uniform branch);
// returns the output for the calling lane's own sample.
__device__ __forceinline__ float prop_mlp_mfma(const unsigned char *pk,
                                               const float (&f)[kPropIn]) {
  const uint32_t l = lane_id();
  const bool hi = l >= 32;
  float x[8], xs[8];
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = j < kPropIn ? f[j] : 0.0f;
#pragma unroll
  for (int j = 0; j < 8; j++) xs[j] = __shfl_xor(x[j], 32);
  prop_s4 b0, b1;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    b0[j] = prop_bf16(hi ? xs[4 + j] : x[j]);  // samples 0-31
    b1[j] = prop_bf16(hi ? x[4 + j] : xs[j]);  // samples 32-63
  }
  const prop_s4 a1 = reinterpret_cast<const prop_s4 *>(pk + kPropPkW1)[l];
  const float *bp1 = reinterpret_cast<const float *>(pk + kPropPkB1) + l * 16;
  const float *bp2 = reinterpret_cast<const float *>(pk + kPropPkB2) + l * 16;
  const float *wp3 = reinterpret_cast<const float *>(pk + kPropPkW3) + l * 16;
  prop_f16 c1, c2;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    c1[r] = bp1[r];
    c2[r] = bp2[r];
  }
  prop_f16 d0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, b0, c1, 0, 0, 0);
  prop_f16 d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, b1, c1, 0, 0, 0);
  prop_f16 e0 = c2, e1 = c2;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const prop_s4 a2 =
        reinterpret_cast<const prop_s4 *>(pk + kPropPkW2)[c * 64 + l];
    prop_s4 h0, h1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float u0 = d0[4 * c + j], u1 = d1[4 * c + j];
      h0[j] = prop_bf16(u0 > 0.0f ? u0 : 0.0f);
      h1[j] = prop_bf16(u1 > 0.0f ? u1 : 0.0f);
    }
    e0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, h0, e0, 0, 0, 0);
    e1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, h1, e1, 0, 0, 0);
  }
  float p0 = 0.0f, p1 = 0.0f;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    float w = wp3[r];
    p0 = __builtin_fmaf(w, e0[r] > 0.0f ? e0[r] : 0.0f, p0);
    p1 = __builtin_fmaf(w, e1[r] > 0.0f ? e1[r] : 0.0f, p1);
  }
  p0 += __shfl_xor(p0, 32);
  p1 += __shfl_xor(p1, 32);
  return (hi ? p1 : p0) + *reinterpret_cast<const float *>(pk + kPropPkB3);
}

// Inference on given feature vectors (n padded to whole waves); every
// lane runs the MLP, lanes beyond n use zero features and discard the
// output.
__global__ __launch_bounds__(kBlock) void prop_eval_kernel(
    const unsigned char *pk, const float *feat, int n, float *out) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  float f[kPropIn];
#pragma unroll
  for (int k = 0; k < kPropIn; k++)
    f[k] = i < n ? feat[(size_t)i * kPropIn + k] : 0.0f;
  float o = prop_mlp_mfma(pk, f);
  if (i < n) out[i] = o;
}

constexpr int kStragLists = 64;  // sub-lists of resume records

struct StragRec {
  uint32_t i;
  uint32_t nsp;
  uint32_t n_att;
  uint32_t flags;     // bit 0: saved at an RK5 step boundary
  float h_begin;
  float err_last;     // scaled error of the neuron's last attempt
};
static_assert(sizeof(StragRec) == kStragRecBytes, "resume record size");

// One RK5 attempt of a neuron plus the step-boundary logic of RK5Update
// and ExternalUpdate (src/rk5.h:70-162, src/aeif_cond_beta_kernel.h:
// 178-214).  Same arithmetic, in the same order, as aeif_rk5_step and the
// loop of aeif_kernel.  Returns true when the neuron has reached the end of
// the time step.
template <int P>
__device__ __forceinline__ bool aeif_attempt_and_update(
    const AeifArgs &A, const double x1, const AeifLocal<P> &p,
    float (&y)[2 + 2 * P], const float (&k1)[2 + 2 * P],
    const float (&scal)[2 + 2 * P], double &x, float &h, float &refr,
    int &nsp, int &n_att, bool &need_k1, int &n_acc, int &n_rej,
    float &err_out, float &hh_out) {
  constexpr int N = 2 + 2 * P;
  float yn[N];
  float h_max = (float)(x1 - x);
  float hh = h;
  if (hh > h_max) hh = h_max;
  if (hh < A.h_min) hh = A.h_min;
  h = hh;
  float err = aeif_rk5_attempt<P>(p, refr, y, k1, scal, hh, A.eps, yn);
  err_out = err;
  hh_out = hh;
  n_att++;
  const bool accept = err <= 1.0f || (double)hh <= A.h_min_acc;
  // err^-1/5 (accepted) or err^-1/4 (rejected), one sequence for the wave
  const float pw = dev_pow_ctrl(err, accept ? -0.2f : -0.25f);
  if (accept) {
    n_acc++;
    if (err > 1.0f) count_forced(A.forced);
    x += (double)h;
    if (err > CK::err_min) {
      h = (h * CK::coeff) * pw;
    } else {
      h = 5.0f * h;
    }
#pragma unroll
    for (int k = 0; k < N; k++) y[k] = yn[k];
  } else {
    n_rej++;
    if (n_att > kMaxAttemptsPerStep) {
      x = x1;  // give up on this step
    } else {
      float h_new = (hh * CK::coeff) * pw;
      float h_lo = 0.1f * hh;
      h = (h_new > h_lo) ? h_new : h_lo;
      return false;  // retry with the smaller step
    }
  }
  bool end = (x >= x1 - (double)A.h_min);
  if (y[0] < -1.0e3f) {
    y[0] = p.V_reset;
    y[1] = 0.0f;
  } else if (y[1] < -1.0e6f || y[1] > 1.0e6f) {
    y[0] = p.V_reset;
    y[1] = 0.0f;
  } else if (refr > 0.0f) {
    y[0] = p.V_reset;
    if (end) refr -= 1.0f;
  } else if (y[0] >= p.V_peak) {
    nsp++;
    y[0] = p.V_reset;
    y[1] = y[1] + p.b;
    float r = roundf(p.t_ref / A.dt);
    int ri = (int)r;
    refr = (float)(ri < 0 ? 0 : ri);
  }
  if (!end) need_k1 = true;
  return end;
}

template <int P>
__device__ __forceinline__ void aeif_store_state(const AeifArgs &A, int i,
                                                 const float (&y)[2 + 2 * P],
                                                 double x, float h,
                                                 float refr) {
  const uint32_t u = (uint32_t)i;
  at(A.x, u) = x;
  at(A.h, u) = h;
  at(A.refr, u) = refr;
  at(A.V, u) = y[0];
  at(A.w, u) = y[1];
#pragma unroll
  for (int q = 0; q < P; q++) {
    at(A.g[q], u) = y[2 + 2 * q];
    at(A.g1[q], u) = y[3 + 2 * q];
  }
}

template <int P>
__device__ __forceinline__ void aeif_count(const AeifArgs &A, int n_acc,
                                           int n_rej) {
  rk5_count(A.cnt, n_acc, n_rej);
}

// Sub-list length set of the step (even / odd gs) and the other set.
__device__ __forceinline__ long long aeif_gs(const AeifArgs &A) {
  return A.clock ? A.clock->gs : A.gs;
}

// First pass: one neuron per lane, coalesced, at most A.budget RK5
// attempts.  Neurons that need more (those that spike or sit near
// threshold, which would otherwise hold their 63 wave neighbours idle)
// save their state and are listed per block for the second pass.
// W: the waves per SIMD the register allocation must allow (1: no
// constraint).  With group-uniform parameters and up to two ports the state
// fits in 80 VGPRs, which allows 6 wavefronts per SIMD instead of 5 (7 %
// faster on c1, measured); for more ports or per-neuron parameters the same
// limit would spill to scratch.
// B1: the attempt budget is 1 (the default), known at compile time, so the
// attempt loop is straight-line code.
template <int P, bool UNI, int W, bool OV, bool B1 = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W)))
void aeif_pass1_kernel(AeifArgs A) {
  constexpr int N = 2 + 2 * P;
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  const double x1 = A.clock ? sn.t : A.x1;
  const long long gs = aeif_gs(A);
  const int sel = (int)(gs & 3);
  unsigned int *strag_n = A.strag_n2 + 2 * kStragLists * sel;
  unsigned int *strag_n_next = A.strag_n2 + 2 * kStragLists * ((gs + 1) & 3);
  StragRec *strag = reinterpret_cast<StragRec *>(A.strag) +
                    (size_t)sel * 2 * kStragLists * A.strag_cap;
  __shared__ uint32_t s_n, s_base, s_nc, s_basec;
  if (threadIdx.x == 0) {
    s_n = 0u;
    s_nc = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2 * kStragLists)
    strag_n_next[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t slot = 0xFFFFFFFFu;
  StragRec rec;
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0, n_acc = 0, n_rej = 0;
  bool finished = false, skipped = false;
  if (OV && active && A.skip_prev) {
    const unsigned long long m =
        A.mask[(size_t)((gs - 1) & 1) * A.n_waves + (i >> 6)];
    skipped = (m >> (i & 63)) & 1ull;
  }
  uint32_t cslot = 0xFFFFFFFFu;
  if (skipped) {
    // owned by the previous step's second pass, which may still run:
    // fold only the input, the second pass of this step runs the step
    float in[P];
    gather_input<P>(A.fold, sn, node, in);
    float *carry = A.carry + (size_t)(gs & 1) * P * A.n;
#pragma unroll
    for (int q = 0; q < P; q++) carry[(size_t)q * A.n + i] = in[q];
    cslot = atomicAdd(&s_nc, 1u);
  } else if (active) {
    AeifLocal<P> p;
    aeif_load_params<P, UNI>(A, i, p);
    float y[N], k1[N], scal[N];
    const uint32_t u = (uint32_t)i;
    y[0] = at(A.V, u);
    y[1] = at(A.w, u);
#pragma unroll
    for (int q = 0; q < P; q++) {
      y[2 + 2 * q] = at(A.g[q], u);
      y[3 + 2 * q] = at(A.g1[q], u);
    }
    float in[P];
    gather_input<P>(A.fold, sn, node, in);
#pragma unroll
    for (int q = 0; q < P; q++)
      y[3 + 2 * q] = __builtin_fmaf(in[q], p.g0[q], y[3 + 2 * q]);
    double x = at(A.x, u);
    float h = at(A.h, u);
    float refr = at(A.refr, u);
    int n_att = 0;
    bool need_k1 = true;
    float h_begin = h, err_last = 1.0f, hh_used = h;
    for (;;) {
      if (n_att >= (B1 ? 1 : A.budget)) break;
      if (need_k1) {
        h_begin = h;
        aeif_rk5_begin<P>(p, refr, y, h, k1, scal);
        need_k1 = false;
      }
      if (aeif_attempt_and_update<P>(A, x1, p, y, k1, scal, x, h, refr,
                                     nsp, n_att, need_k1, n_acc, n_rej,
                                     err_last, hh_used)) {
        finished = true;
        break;
      }
    }
    aeif_store_state<P>(A, i, y, x, h, refr);
    if (!finished) {
      slot = atomicAdd(&s_n, 1u);
      rec.i = (uint32_t)i;
      rec.nsp = (uint32_t)nsp;
      rec.n_att = (uint32_t)n_att;
      rec.flags = need_k1 ? 1u : 0u;
      rec.h_begin = need_k1 ? h : h_begin;
      rec.err_last = err_last;
    }
  }
  emit_spikes(A.out, sn, active && finished, ii, node, nsp, (float)nsp);
  rk5_count(A.cnt1 ? A.cnt1 : A.cnt, n_acc, n_rej);
  if (OV && A.mask_write) {
    const unsigned long long m =
        __ballot(slot != 0xFFFFFFFFu || cslot != 0xFFFFFFFFu);
    if (lane_id() == 0 && (unsigned long long)(i >> 6) < A.n_waves)
      A.mask[(size_t)(gs & 1) * A.n_waves + (i >> 6)] = m;
  }
  // this block's resume records go to sub-list blockIdx % kStragLists,
  // at an offset reserved with one atomic per block (and its catch-up
  // entries to the same sub-list of the catch-up lists)
  __syncthreads();
  const uint32_t list = blockIdx.x % kStragLists;
  if (threadIdx.x == 0) {
    s_base = s_n ? atomicAdd(strag_n + list, s_n) : 0u;
    if (OV) s_basec = s_nc ? atomicAdd(strag_n + kStragLists + list, s_nc) : 0u;
  }
  __syncthreads();
  if (slot != 0xFFFFFFFFu)
    strag[(size_t)list * A.strag_cap + s_base + slot] = rec;
  if (OV && cslot != 0xFFFFFFFFu) {
    StragRec c;
    c.i = (uint32_t)i;
    c.nsp = 0u;
    c.n_att = 0u;
    c.flags = 1u;
    c.h_begin = 0.0f;
    c.err_last = 1.0f;
    strag[(size_t)(kStragLists + list) * A.strag_cap + s_basec + cslot] = c;
  }
}

// Second pass: the listed neurons, run to the end of the step.  The
// sub-lists are concatenated (prefix over their counts, in LDS) and lane g
// of the grid takes entries g, g + G, ...; a lane moves on to its next
// entry as soon as the current neuron is done, so lanes stay busy while
// neurons need different numbers of attempts, and waves beyond the list
// end exit at once.
// PROP: the learned proposer's inference is compiled in (its weights in
// LDS, the MFMA registers); without it the pass keeps its occupancy.
template <int P, bool UNI, bool PROP, bool CU>
__device__ __forceinline__ void aeif_pass2_body(const AeifArgs &A,
                                                uint32_t vblock,
                                                uint32_t vgrid) {
  constexpr int N = 2 + 2 * P;
  // the pass is a few long dependent chains: its waves issue first when
  // they share a SIMD with the first pass or the delivery
  if (CU) __builtin_amdgcn_s_setprio(3);
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  const double x1 = A.clock ? sn.t : A.x1;
  const long long gs = aeif_gs(A);
  // this step's set: resume records, then the catch-up entries (neurons
  // the first pass of this step skipped)
  const unsigned int *strag_n = A.strag_n2 + 2 * kStragLists * (gs & 3);
  const unsigned int *strag_n_prev = strag_n + kStragLists;
  __shared__ uint32_t s_pre[kStragLists + 1], s_pre2[kStragLists + 1];
  // proposer weights, staged once per block (read at every step start)
  __shared__ __attribute__((aligned(16)))
  unsigned char s_pk[PROP ? kPropPkBytes : 16];
  if (PROP && A.prop_pk) {
    const uint4 *src = reinterpret_cast<const uint4 *>(A.prop_pk);
    uint4 *dst = reinterpret_cast<uint4 *>(s_pk);
    for (int k = threadIdx.x; k < kPropPkBytes / 16; k += kBlock)
      dst[k] = src[k];
  }
  if (threadIdx.x < (CU ? 128 : 64)) {  // exclusive scans of 64 lengths
    const int l = threadIdx.x & 63;
    const bool second = threadIdx.x >= 64;
    uint32_t c = second ? (A.skip_prev ? strag_n_prev[l] : 0u) : strag_n[l];
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t v = __shfl_up(incl, off);
      if (l >= off) incl += v;
    }
    uint32_t *pre = second ? s_pre2 : s_pre;
    pre[l] = incl - c;
    if (l == 63) pre[kStragLists] = incl;
  }
  __syncthreads();
  const uint32_t n_cont = s_pre[kStragLists];
  const uint32_t total = n_cont + (CU ? s_pre2[kStragLists] : 0u);
  const uint32_t n_lanes = vgrid * kBlock;
  uint32_t j = vblock * kBlock + threadIdx.x;
  if (vblock * kBlock >= total) return;  // whole block beyond the list
  const StragRec *list = reinterpret_cast<const StragRec *>(A.strag) +
                        (size_t)(gs & 3) * 2 * kStragLists * A.strag_cap;
  const StragRec *list_prev = list + (size_t)kStragLists * A.strag_cap;
  const float *carry = A.carry + (size_t)(gs & 1) * P * A.n;
  AeifLocal<P> p;
  float y[N], k1[N], scal[N];
  double x = 0.0;
  float h = 0.0f, refr = 0.0f;
  int nsp = 0, n_att = 0, n_acc = 0, n_rej = 0;
  int i = 0;
  float err_last = 1.0f;
  bool need_k1 = false, have = false;
  for (;;) {
    if (!have && j < total) {
      // entries [0, n_cont) resume this step's deferred neurons, the rest
      // are the catch-up neurons; sub-list of an entry: last l with
      // pre[l] <= its index
      const bool catch_up = CU && j >= n_cont;
      const uint32_t jj = catch_up ? j - n_cont : j;
      const uint32_t *pre = catch_up ? s_pre2 : s_pre;
      int lo = 0, hi = kStragLists;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= jj) lo = mid; else hi = mid;
      }
      StragRec r = (catch_up ? list_prev : list)[(size_t)lo * A.strag_cap +
                                                  (jj - pre[lo])];
      j += n_lanes;
      i = (int)r.i;
      nsp = catch_up ? 0 : (int)r.nsp;
      n_att = catch_up ? 0 : (int)r.n_att;
      err_last = catch_up ? 1.0f : r.err_last;
      aeif_load_params<P, UNI>(A, i, p);
      const uint32_t u = (uint32_t)i;
      y[0] = at(A.V, u);
      y[1] = at(A.w, u);
#pragma unroll
      for (int q = 0; q < P; q++) {
        y[2 + 2 * q] = at(A.g[q], u);
        y[3 + 2 * q] = at(A.g1[q], u);
      }
      x = at(A.x, u);
      h = at(A.h, u);
      refr = at(A.refr, u);
      if (catch_up) {
        // the whole step, as the first pass runs it: input fold, then
        // RK5 steps from the stored step size
#pragma unroll
        for (int q = 0; q < P; q++)
          y[3 + 2 * q] =
              __builtin_fmaf(carry[(size_t)q * A.n + i], p.g0[q], y[3 + 2 * q]);
        need_k1 = true;
      } else if (r.flags & 1u) {
        need_k1 = true;  // saved at a step boundary: a new RK5 step starts
      } else {
        aeif_rk5_begin<P>(p, refr, y, r.h_begin, k1, scal);
        need_k1 = false;
      }
      have = true;
    }
    if (__ballot(have) == 0ull) break;
    bool finished = false, attempted = false;
    float feat[kPropIn];
#pragma unroll
    for (int k = 0; k < kPropIn; k++) feat[k] = 0.0f;
    float err = 1.0f, hh = 0.0f;
    if (PROP && A.prop_pk) {
      // learned step size for the lanes that start an RK5 step; the MLP
      // runs on the whole wave
      bool want = have && need_k1;
      if (want) prop_features<P>(p, y, h, err_last, 1.0f / A.dt, feat);
      if (__ballot(want) != 0ull) {
        float o = prop_mlp_mfma(s_pk, feat);
        if (want) {
          // trust region around the plain controller's step
          float hp = A.dt * exp2f(o) * kPropSafety;
          float lo = h * kPropShrinkMax, hi = h * kPropGrowMax;
          hp = hp > lo ? hp : lo;   // also replaces a NaN output
          hp = hp < hi ? hp : hi;
          h = hp > A.h_min ? hp : A.h_min;
        }
      }
    }
    if (have) {
      if (need_k1) {
        aeif_rk5_begin<P>(p, refr, y, h, k1, scal);
        need_k1 = false;
      }
      if (A.prop_log) prop_features<P>(p, y, h, err_last, 1.0f / A.dt, feat);
      finished = aeif_attempt_and_update<P>(A, x1, p, y, k1, scal, x, h,
                                            refr, nsp, n_att, need_k1, n_acc,
                                            n_rej, err, hh);
      err_last = err;
      attempted = true;
      if (finished) aeif_store_state<P>(A, i, y, x, h, refr);
    }
    if (A.prop_log) {
      // sample: features before the attempt, target log2 of the step that
      // would have given a scaled error of 1
      unsigned long long m = __ballot(attempted);
      if (m != 0ull) {
        uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
        unsigned long long base = 0;
        if (lane_id() == leader)
          base = atomicAdd(A.prop_log_n, (unsigned long long)__popcll(m));
        base = __shfl(base, (int)leader);
        unsigned long long slot2 = base + lanes_below(m);
        if (attempted && slot2 < A.prop_log_max) {
          float e = err > CK::err_min ? err : CK::err_min;
          float t = __log2f(hh / A.dt) - 0.2f * __log2f(e);
          float *o = A.prop_log + slot2 * kPropSample;
#pragma unroll
          for (int k = 0; k < kPropIn; k++) o[k] = feat[k];
          o[kPropIn] = t;
          o[kPropIn + 1] = 0.0f;
        }
      }
    }
    emit_spikes(A.out, sn, finished, finished ? i : 0,
                A.i0 + (finished ? i : 0),
                nsp, (float)nsp);
    if (finished) have = false;
  }
  aeif_count<P>(A, n_acc, n_rej);
}

template <int P, bool UNI, bool PROP, bool CU>
__global__ __launch_bounds__(kBlock) void aeif_pass2_kernel(AeifArgs A) {
  aeif_pass2_body<P, UNI, PROP, CU>(A, blockIdx.x, gridDim.x);
}

template <bool PLASTIC>
__device__ __forceinline__ void delivery_body(const DeliveryArgs &A,
                                              uint32_t vblock,
                                              uint32_t vgrid);

// End of a captured step: block 0 of its last launch writes the next
// step's clock (computed as the host computes it in SimulationStep).
__device__ __forceinline__ void clock_step(const StepClock *c, StepClock *cn) {
  if (!cn || blockIdx.x != 0 || threadIdx.x != 0) return;
  StepClock v = *c;
  v.pois_T = v.T;
  v.pois_t = v.t;
  v.gs += 1;
  v.it += 1;
  v.T = v.T_base + v.it + 1;
  v.t = v.t0 + v.dt * (double)(v.it + 1);
  *cn = v;
}

// Arguments of a delivery or STDP launch with the step values of the
// clock, when the launch has one.
__device__ __forceinline__ DeliveryArgs clocked(const DeliveryArgs &D) {
  DeliveryArgs A = D;
  if (A.clock) {
    A.gs = A.clock->gs;
    A.T = A.clock->T;
  }
  return A;
}
__device__ __forceinline__ RevArgs clocked(const RevArgs &R) {
  RevArgs A = R;
  if (A.clock) {
    A.gs = A.clock->gs;
    A.T = A.clock->T;
  }
  return A;
}

// Second RK5 pass and the delivery of the same step in one launch: blocks
// [0, n2) run the second pass, the others the delivery.  The two are
// independent within a step when no synapse is plastic (the delivery reads
// spikes of earlier steps and writes only the input accumulators), so the
// latency-bound second pass and the atomic-bound delivery share the chip
// without a second stream.
template <int P, bool UNI, bool PROP, bool CU>
__global__ __launch_bounds__(kBlock) void aeif_tail_kernel(AeifArgs A,
                                                           DeliveryArgs D,
                                                           uint32_t n2) {
  clock_step(D.clock, D.clock_next);
  if (blockIdx.x < n2)
    aeif_pass2_body<P, UNI, PROP, CU>(A, blockIdx.x, n2);
  else
    delivery_body<false>(clocked(D), blockIdx.x - n2, gridDim.x - n2);
}

// ---------------------------------------------------------------------
// iaf_psc_exp, exact integration (src/iaf_psc_exp.cu:94-119)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void iaf_kernel(IafArgs A) {
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0;
  if (active) {
    float in[2];
    gather_input<2>(A.fold, sn, node, in);
    float I_ex = A.I_ex[i] + in[0];
    float I_in = A.I_in[i] + in[1];
    if (A.do_step) {
      float V = A.V[i];
      float refr = A.refr[i];
      if (refr > 0.0f) {
        refr -= 1.0f;
      } else {
        float v = V * pf(A.P22, i);
        v = __builtin_fmaf(I_ex, pf(A.P21ex, i), v);
        v = __builtin_fmaf(I_in, pf(A.P21in, i), v);
        v = __builtin_fmaf(pf(A.I_e, i), pf(A.P20, i), v);
        V = v;
      }
      I_ex = I_ex * pf(A.P11ex, i);
      I_in = I_in * pf(A.P11in, i);
      if (V >= pf(A.Theta, i)) {
        nsp = 1;
        V = pf(A.V_reset, i);
        refr = (float)(int)roundf(pf(A.t_ref, i) / A.dt);
      }
      A.V[i] = V;
      A.refr[i] = refr;
    }
    A.I_ex[i] = I_ex;
    A.I_in[i] = I_in;
  }
  if (!A.do_step) return;
  emit_spikes(A.out, sn, active, ii, node, nsp, 1.0f);
}

// ---------------------------------------------------------------------
// spike_detector / parrot_neuron
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void relay_kernel(RelayArgs A) {
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0;
  float height = 0.0f;
  if (active) {
    float in[2];
    if (A.parrot) {
      gather_input<2>(A.fold, sn, node, in);
    } else {
      float in1[1];
      gather_input<1>(A.fold, sn, node, in1);
      in[0] = in1[0];
    }
    float input = A.input[i] + in[0];
    if (A.do_step) {
      float sh = input;
      float hold = pf(A.hold, i);
      if (A.parrot) {
        A.out_var[i] = sh;
        if (sh != 0.0f) {
          if (hold == 0.0f) sh = 1.0f;
          input = 0.0f;
          nsp = 1;
          height = sh;
        }
      } else {
        if (sh != 0.0f) {
          if (hold == 0.0f) sh = 1.0f;
          input = 0.0f;
        }
        A.out_var[i] = sh;
      }
    }
    A.input[i] = input;
  }
  if (!A.do_step || !A.parrot) return;
  emit_spikes(A.out, sn, active, ii, node, nsp, height);
}

// ---------------------------------------------------------------------
// spike_generator (src/spike_generator.cu:37-55)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void spikegen_kernel(SpikeGenArgs A) {
  const StepNow sn = step_now(A.out, nullptr, A.clock);
  const long long it = A.clock ? A.clock->it : A.it;
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int nsp = 0;
  float height = 0.0f;
  if (active) {
    int k = A.next[i];
    if (k < A.n_spikes[i] && (long long)A.time_idx[A.first[i] + k] == it) {
      nsp = 1;
      height = A.height[A.first[i] + k];
      A.next[i] = k + 1;
    }
  }
  emit_spikes(A.out, sn, active, ii, A.i0 + ii, nsp, height);
}

// ---------------------------------------------------------------------
// CreatePoissonGenerator nodes (src/poisson.cu:27-54): Threefry word at
// counter (node, T mod 2^32), key (seed_lo ^ kLegacyPoissonTag, seed_hi +
// T / 2^32), inverted through the CDF table.
// The reference pushes these spikes before it advances the step's time
// (src/neurongpu.cu:461-470), so they carry the previous step's time index
// and neural time; they enter the same ring slot as this step's spikes.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void legacy_pois_kernel(LegacyPoisArgs A) {
  StepNow sn = step_now(A.out, nullptr, A.clock);
  const long long it = A.clock ? A.clock->it : A.it;
  const double t0 = A.clock ? A.clock->t0 : A.t0;
  const double dt = A.clock ? A.clock->dt : A.dt;
  const long long T = sn.T;
  sn.T = T - 1;
  sn.t = t0 + dt * (double)it;
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0;
  float height = 0.0f;
  if (active) {
    uint32_t u = dev_threefry2x32_w0(
        (uint32_t)node, (uint32_t)T, A.key0 ^ kLegacyPoissonTag,
        A.key1 + (uint32_t)((uint64_t)T >> 32));
    int k = 0;
    while (u > A.tab[k]) k++;   // the table ends with 0xFFFFFFFF
    if ((float)k > A.max_count) k = 0;
    if (k > 0) {
      nsp = 1;
      height = (float)k;
    }
  }
  emit_spikes(A.out, sn, active, ii, node, nsp, height);
}

// ---------------------------------------------------------------------
// Delivery.  Pairs (spike, delay slot d) due at global step gs are the
// spikes of ring slot gs-1-d combined with the source's synapse segment of
// delay index d.  Wavefronts take tiles of 64 pairs; a wavefront scans the
// 64 segment sizes, then its lanes stream the flattened synapse range of
// the tile (consecutive lanes read consecutive 8-byte records) and add
// height*weight into the target accumulators.
// ---------------------------------------------------------------------
constexpr int kMaxDelaySlots = 4096;  // checked on the host at calibration

// Block prologue of every delivery kernel: s_pre[d] = number of (spike,
// delay) pairs with delay slot < d due this step (d = 0..D), from the ring
// counts of slots gs-1-d.  Virtual block 0 also clears the ring counter of
// the slot the next step fills (slot gs+1 is never read in this step).
// Returns the number of pairs.  BS: threads per block; s_part holds BS
// entries.
template <int BS = kBlock>
__device__ __forceinline__ uint32_t delivery_prologue(const DeliveryArgs &A,
                                                      uint32_t vblock,
                                                      uint32_t *s_pre,
                                                      uint32_t *s_part) {
  const int D = A.D;
  const int tid = threadIdx.x;
  if (D <= 64) {
    // one wave: scan of the D counts through cross-lane shuffles
    if (tid < 64) {
      uint32_t n = 0;
      if (tid < D) {
        long long s = A.gs - 1 - tid;
        if (s >= 0) n = A.ring_n[s % A.R];
      }
      uint32_t incl = n;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off);
        if (tid >= off) incl += v;
      }
      if (tid < D) s_pre[tid] = incl - n;
      if (tid == 63) s_pre[D] = incl;
    }
    __syncthreads();
    if (vblock == 0 && tid == 0) A.ring_n[(A.gs + 1) % A.R] = 0u;
    return s_pre[D];
  }
  // per-thread partial sums of the due spike counts over d
  const int per = (D + BS - 1) / BS;
  uint32_t run = 0;
  for (int k = 0; k < per; k++) {
    int d = tid * per + k;
    uint32_t n = 0;
    if (d < D) {
      long long s = A.gs - 1 - d;
      if (s >= 0) n = A.ring_n[s % A.R];
    }
    run += n;
  }
  s_part[tid] = run;
  __syncthreads();
  // inclusive scan of the block partials (Hillis-Steele in LDS)
  for (int off = 1; off < BS; off <<= 1) {
    uint32_t v = (tid >= off) ? s_part[tid - off] : 0u;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  {
    uint32_t acc = (tid > 0) ? s_part[tid - 1] : 0u;
    for (int k = 0; k < per; k++) {
      int d = tid * per + k;
      if (d < D) {
        s_pre[d] = acc;
        long long s = A.gs - 1 - d;
        if (s >= 0) acc += A.ring_n[s % A.R];
      }
    }
    if (tid == BS - 1) s_pre[D] = s_part[BS - 1];
  }
  __syncthreads();
  if (vblock == 0 && tid == 0) A.ring_n[(A.gs + 1) % A.R] = 0u;
  return s_pre[D];
}

// Enumerates the synaptic events of this wave's share of the step and calls
// f(ok, syn, rec, height, mult) for U 64-event chunks at a time (all U
// record loads are issued before the calls).  The wave takes tiles of pt
// pairs (the largest power of two in [1, tile_max] that still gives every
// wave of the virtual grid a tile).  When the tile's segments average at
// least kLongSegment synapses they are walked one at a time; otherwise
// the tile's events are spread over the 64 lanes by a lane-level scan of
// the segment lengths and a binary-lift owner search.  The
// enumeration depends only on (vblock, vgrid, tile_max) and the ring, so
// two kernels launched with the same arguments see the same events in the
// same blocks.  Returns the wave's event count.
constexpr uint32_t kLongSegment = 32;

// Synapse segment of (source, delay slot) in the canonical store: its
// records rec[start .. start + cnt).
struct SegLoc {
  __device__ __forceinline__ void operator()(const DeliveryArgs &A,
                                             uint32_t src, int d,
                                             uint32_t &cnt,
                                             unsigned long long &start) const {
    const uint32_t *sg = A.seg + (size_t)src * (A.D + 1) + d;
    const uint32_t o0 = sg[0], o1 = sg[1];
    cnt = o1 - o0;
    start = A.base[src] + o0;
  }
};

// BS: threads per block; Loc(src, d, cnt, start) locates a pair's records
// in `recs`.
template <int U, int BS = kBlock, class F, class Loc = SegLoc>
__device__ __forceinline__ unsigned long long deliver_events(
    const DeliveryArgs &A, uint32_t vblock, uint32_t vgrid,
    const uint32_t *s_pre, uint32_t total, F &&f,
    const unsigned long long *recs = nullptr, Loc loc = Loc{}) {
  const int D = A.D;
  if (!recs) recs = A.rec;
  const uint32_t lane = lane_id();
  const uint32_t wave = (uint32_t)(vblock * (BS / 64) + threadIdx.x / 64);
  const uint32_t n_waves = vgrid * (BS / 64);
  uint32_t pt = (uint32_t)A.tile_max;
  while (pt > 1 && (total + pt - 1) / pt < n_waves) pt >>= 1;
  const uint32_t n_tiles = (total + pt - 1) / pt;
  unsigned long long n_events = 0;
  for (uint32_t wt = wave; wt < n_tiles; wt += n_waves) {
    uint32_t pidx = wt * pt + lane;
    uint32_t cnt = 0, mult = 0;
    unsigned long long start = 0;
    float height = 0.0f;
    if (lane < pt && pidx < total) {
      // delay slot d of this pair: last d with s_pre[d] <= pidx
      int lo = 0, hi = D;  // s_pre[lo] <= pidx < s_pre[hi]
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_pre[mid] <= pidx) lo = mid; else hi = mid;
      }
      int d = lo;
      long long slot = (A.gs - 1 - d) % A.R;
      long long e = slot * A.cap + (pidx - s_pre[d]);
      uint32_t src = A.ring_src[e];
      height = A.ring_h[e];
      mult = A.ring_cnt[e];
      loc(A, src, d, cnt, start);
    }
    // exclusive scan of cnt over the wavefront
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t v = __shfl_up(incl, off);
      if ((int)lane >= off) incl += v;
    }
    uint32_t pre = incl - cnt;
    uint32_t E = __shfl(incl, 63);
    n_events += E;
    const unsigned long long nz = __ballot(cnt > 0u);
    if (E >= (uint32_t)kLongSegment * (uint32_t)__popcll(nz)) {
      // Long segments: one pair at a time, its data broadcast through
      // scalar registers, consecutive lanes on consecutive records.
      unsigned long long m = nz;
      while (m) {
        const int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1ull;
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)cnt, j);
        const uint32_t slo =
            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)start, j);
        const uint32_t shi =
            (uint32_t)__builtin_amdgcn_readlane((int)(start >> 32), j);
        const unsigned long long st = ((unsigned long long)shi << 32) | slo;
        const float h = __uint_as_float((uint32_t)__builtin_amdgcn_readlane(
            (int)__float_as_uint(height), j));
        const uint32_t mu = (uint32_t)__builtin_amdgcn_readlane((int)mult, j);
        for (uint32_t o = 0; o < c; o += 64u * U) {
          unsigned long long syn[U], r[U];
          bool ok[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            uint32_t e = o + 64u * u + lane;
            ok[u] = e < c;
            syn[u] = st + e;
            r[u] = ok[u] ? recs[syn[u]] : 0ull;
          }
#pragma unroll
          for (int u = 0; u < U; u++) f(ok[u], syn[u], r[u], h, mu);
        }
      }
      continue;
    }
    // Short segments: the tile's events are spread over the lanes; each
    // lane finds its pair by a binary lift over the exclusive scan.
    for (uint32_t e0 = 0; e0 < E; e0 += 64u * U) {
      unsigned long long syn[U], r[U];
      float hj[U];
      uint32_t mj[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        uint32_t e = e0 + 64u * u + lane;
        // owner lane j < pt: last lane with pre[j] <= e
        int j = 0;
        for (uint32_t stp = pt >> 1; stp > 0; stp >>= 1) {
          int cand = j + (int)stp;
          uint32_t pv = __shfl(pre, cand);
          if (pv <= e) j = cand;
        }
        uint32_t pj = __shfl(pre, j);
        uint32_t slo = __shfl((uint32_t)start, j);
        uint32_t shi = __shfl((uint32_t)(start >> 32), j);
        hj[u] = __shfl(height, j);
        mj[u] = __shfl(mult, j);
        ok[u] = e < E;
        syn[u] = (((unsigned long long)shi << 32) | slo) + (e - pj);
        r[u] = ok[u] ? recs[syn[u]] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < U; u++) f(ok[u], syn[u], r[u], hj[u], mj[u]);
    }
  }
  return n_events;
}

// Dynamic LDS of the delivery kernels: s_pre[D + 1], then (binned
// phases) one counter per bin.
extern __shared__ __align__(16) uint32_t s_dyn[];

inline size_t delivery_lds_bytes(int D, int nb) {
  return sizeof(uint32_t) * ((size_t)D + 1 + (size_t)nb);
}

template <bool PLASTIC>
__device__ __forceinline__ void delivery_body(const DeliveryArgs &A,
                                              uint32_t vblock,
                                              uint32_t vgrid) {
  uint32_t *s_pre = s_dyn;
  __shared__ uint32_t s_part[kBlock];
  const uint32_t total = delivery_prologue(A, vblock, s_pre, s_part);
  unsigned long long n_events = deliver_events<2>(
      A, vblock, vgrid, s_pre, total,
      [&](bool ok, unsigned long long syn, unsigned long long r, float hj,
          uint32_t mj) {
        if (!ok) return;
        uint32_t tp = (uint32_t)r;
        float w = __uint_as_float((uint32_t)(r >> 32));
        uint32_t tgt = tp & kTargetMask;
        uint32_t port = tp >> kPortShift;
        unsafeAtomicAdd(A.acc + (long long)port * A.n_total + tgt, hj * w);
        if (PLASTIC) {
          int sgp = A.syn_group[syn];
          if (sgp > 0) {
            int type = A.syn_type[sgp - 1];
            const float *prm = A.stdp_param + 8 * (sgp - 1);
            float wn = w;
            for (uint32_t rep = 0; rep < mj; rep++) {
              A.spike_time[syn] = (uint16_t)(A.T & 0xffff);
              long long Dt_int = A.T - A.last_rev[tgt];
              if (Dt_int > 0 && Dt_int < kMaxSynDt)
                wn = dev_syn_update(type, wn, -A.dt * (float)Dt_int, prm);
            }
            if (wn != w)
              reinterpret_cast<uint32_t *>(A.rec + syn)[1] =
                  __float_as_uint(wn);
          }
        }
      });
  if (lane_id() == 0 && n_events) {
    atomicAdd(A.events + (vblock & (kCounterSlots - 1)), n_events);
  }
}

template <bool PLASTIC>
__global__ __launch_bounds__(kBlock) void delivery_kernel(DeliveryArgs A) {
  clock_step(A.clock, A.clock_next);
  delivery_body<PLASTIC>(clocked(A), blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------
// Binned delivery (non-plastic networks).  Float atomics on the
// accumulators run at a fixed rate per device whatever the address
// pattern, so when a step carries many events per accumulator entry the
// events are first grouped by accumulator range (bin of kBinLen entries):
//   1 count    per (sub-bin, block) event counts, LDS histogram
//   2 scan     per sub-bin, exclusive offsets over the blocks, totals
//   3 scatter  the same enumeration again; each event is written as
//              (index within bin, value) at its bin's cursor; block 0
//              also writes the bin and chunk starts.  If the step has more
//              events than the buffer holds, this phase adds them with
//              atomics instead and the later phases do nothing.
//   4 reduce   one block per chunk of events of one bin sums them in LDS
//              (chunk size chosen from the step's event count, see
//              bin_chunk_size, or one chunk per bin when there are at
//              least kSingleChunkBins bins of about equal size; then the
//              sum phase has nothing to do); a bin of one chunk is added to the
//              accumulators directly, otherwise the partial sums are
//              written out
//   5 sum      bins of several chunks: accumulators += sum of partials.
// Count and scatter use the same grid and tile settings, so every block
// enumerates the same events in both.
// ---------------------------------------------------------------------

constexpr int kReduceBlock = 1024;
constexpr int kBinBlock = 512;  // threads per block of the count and scatter

// Float add into LDS by compare-and-swap.  On gfx950 the LDS float atomic
// add runs at about a twelfth of the rate of the integer atomics and of a
// plain read-modify-write; a CAS loop whose first attempt nearly always
// succeeds (random addresses in a table of kBinLen entries) is faster.
__device__ __forceinline__ void lds_add_f32(float *p, float v) {
  uint32_t *u = reinterpret_cast<uint32_t *>(p);
  uint32_t old = *u;
  for (;;) {
    uint32_t nw = __float_as_uint(__uint_as_float(old) + v);
    uint32_t prev = atomicCAS(u, old, nw);
    if (prev == old) break;
    old = prev;
  }
}

__device__ __forceinline__ uint32_t bin_chunk_size(unsigned long long n_ev) {
  unsigned long long ch = (n_ev + kReduceChunks - 1) / kReduceChunks;
  ch = (ch + 1023ull) & ~1023ull;
  if (ch < (unsigned long long)kBinChunkMin) ch = kBinChunkMin;
  if (ch > 0x40000000ull) ch = 0x40000000ull;
  return (uint32_t)ch;
}

// Exclusive scan of val(0..n-1) by one block: out[b] = sum of val(b') for
// b' < b.  Returns the 64-bit total.  s_part holds kBlock entries.
template <int BS = kBlock, class F>
__device__ __forceinline__ unsigned long long block_scan_excl(
    int n, F val, uint32_t *out, unsigned long long *s_part) {
  const int tid = threadIdx.x;
  const int per = (n + BS - 1) / BS;
  unsigned long long run = 0;
  for (int k = 0; k < per; k++) {
    int b = tid * per + k;
    if (b < n) run += val(b);
  }
  s_part[tid] = run;
  __syncthreads();
  for (int off = 1; off < BS; off <<= 1) {
    unsigned long long v = (tid >= off) ? s_part[tid - off] : 0ull;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  unsigned long long acc = (tid > 0) ? s_part[tid - 1] : 0ull;
  for (int k = 0; k < per; k++) {
    int b = tid * per + k;
    if (b < n) {
      unsigned long long v = val(b);  // before out[b]: out may alias the input
      out[b] = (uint32_t)acc;
      acc += v;
    }
  }
  unsigned long long total = s_part[BS - 1];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kBinBlock) void bin_count_kernel(DeliveryArgs A0,
                                                              BinArgs B) {
  const DeliveryArgs A = clocked(A0);
  uint32_t *s_pre = s_dyn;
  uint32_t *s_h = s_dyn + A.D + 1;
  __shared__ uint32_t s_part[kBinBlock];
  for (int b = threadIdx.x; b < B.nsub; b += kBinBlock) s_h[b] = 0u;
  const uint32_t total =
      delivery_prologue<kBinBlock>(A, blockIdx.x, s_pre, s_part);
  deliver_events<4, kBinBlock>(A, blockIdx.x, gridDim.x, s_pre, total,
                 [&](bool ok, unsigned long long, unsigned long long r, float,
                     uint32_t) {
                   if (!ok) return;
                   uint32_t tp = (uint32_t)r;
                   long long idx = (long long)(tp >> kPortShift) * A.n_total +
                                   (tp & kTargetMask);
                   atomicAdd(&s_h[idx >> B.sub_shift], 1u);
                 });
  __syncthreads();
  for (int b = threadIdx.x; b < B.nsub; b += kBinBlock)
    B.hist[(size_t)b * B.grid + blockIdx.x] = s_h[b];
}

// one block per sub-bin
__global__ __launch_bounds__(kBlock) void bin_scan_kernel(BinArgs B) {
  __shared__ unsigned long long s_part[kBlock];
  uint32_t *h = B.hist + (size_t)blockIdx.x * B.grid;
  // in place: each thread reads and rewrites only its own entries
  unsigned long long t = block_scan_excl(
      B.grid, [&](int k) -> unsigned long long { return h[k]; }, h, s_part);
  if (threadIdx.x == 0)
    B.tot[blockIdx.x] = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
}

__global__ __launch_bounds__(kBinBlock) void bin_scatter_kernel(
    DeliveryArgs A0, BinArgs B) {
  const DeliveryArgs A = clocked(A0);
  uint32_t *s_pre = s_dyn;
  uint32_t *s_cur = s_dyn + A.D + 1;
  __shared__ uint32_t s_part[kBinBlock];
  __shared__ unsigned long long s_part64[kBinBlock];
  __shared__ int s_multi;  // some bin has more than one chunk
  if (threadIdx.x == 0) s_multi = 0;
  const unsigned long long n_ev = block_scan_excl<kBinBlock>(
      B.nsub, [&](int b) -> unsigned long long { return B.tot[b]; }, s_cur,
      s_part64);
  const bool binned = n_ev <= (unsigned long long)B.ev_cap;
  if (blockIdx.x == 0) {
    // bin and chunk starts for the reduce and sum phases; the chunk size
    // gives about kReduceChunks chunks in all (at least kBinChunkMin
    // events each), so the reduction fills the device whatever the step's
    // event count
    uint32_t ch = bin_chunk_size(n_ev);
    const int spb = kBinLen >> B.sub_shift;  // sub-bins per bin
    auto start = [&](int b) -> unsigned long long {
      return b * spb < B.nsub ? s_cur[b * spb] : n_ev;
    };
    // With many bins of about equal size, one chunk per bin: the reduction
    // still has a block for most CUs, and no partial sums are written and
    // summed.
    if (B.nb >= kSingleChunkBins) {
      unsigned long long mx = 0;
      for (int b = threadIdx.x; b < B.nb; b += kBinBlock) {
        const unsigned long long nbin = start(b + 1) - start(b);
        mx = nbin > mx ? nbin : mx;
      }
      s_part64[threadIdx.x] = mx;
      __syncthreads();
      for (int off = kBinBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off &&
            s_part64[threadIdx.x + off] > s_part64[threadIdx.x])
          s_part64[threadIdx.x] = s_part64[threadIdx.x + off];
        __syncthreads();
      }
      mx = s_part64[0];
      __syncthreads();
      if (mx * (unsigned long long)B.nb <= 2ull * n_ev) {
        const unsigned long long one = (mx + 1023ull) & ~1023ull;
        if (one > ch && one <= 0x40000000ull) ch = (uint32_t)one;
      }
    }
    for (int b = threadIdx.x; b < B.nb; b += kBinBlock)
      B.bstart[b] = binned ? (uint32_t)start(b) : 0u;
    if (threadIdx.x == 0) {
      B.bstart[B.nb] = binned ? (uint32_t)n_ev : 0u;
      B.meta[0] = ch;
    }
    unsigned long long nc = block_scan_excl<kBinBlock>(
        B.nb,
        [&](int b) -> unsigned long long {
          const unsigned long long c =
              binned ? (start(b + 1) - start(b) + ch - 1u) / ch : 0u;
          if (c > 1u) s_multi = 1;
          return c;
        },
        B.cstart, s_part64);
    if (threadIdx.x == 0) {
      B.cstart[B.nb] = (uint32_t)nc;
      B.meta[1] = (uint32_t)s_multi;
    }
  }
  for (int b = threadIdx.x; b < B.nsub; b += kBinBlock)
    s_cur[b] += B.hist[(size_t)b * B.grid + blockIdx.x];
  // (the prologue's barriers order these LDS writes before their use)
  const uint32_t total =
      delivery_prologue<kBinBlock>(A, blockIdx.x, s_pre, s_part);
  unsigned long long n_events = deliver_events<4, kBinBlock>(
      A, blockIdx.x, gridDim.x, s_pre, total,
      [&](bool ok, unsigned long long, unsigned long long r, float hj,
          uint32_t) {
        if (!ok) return;
        uint32_t tp = (uint32_t)r;
        float w = __uint_as_float((uint32_t)(r >> 32));
        long long idx = (long long)(tp >> kPortShift) * A.n_total +
                        (tp & kTargetMask);
        float v = hj * w;
        if (binned) {
          uint32_t pos = atomicAdd(&s_cur[idx >> B.sub_shift], 1u);
          B.ev[pos] = ((unsigned long long)__float_as_uint(v) << 32) |
                      (uint32_t)(idx & (kBinLen - 1));
        } else {
          unsafeAtomicAdd(A.acc + idx, v);
        }
      });
  if (lane_id() == 0 && n_events) {
    atomicAdd(A.events + (blockIdx.x & (kCounterSlots - 1)), n_events);
  }
}

__global__ __launch_bounds__(kReduceBlock) void bin_reduce_kernel(BinArgs B,
                                                                  float *acc) {
  __shared__ float s_a[kBinLen];
  const uint32_t n_chunk = B.cstart[B.nb];
  const uint32_t ch = B.meta[0];
  for (uint32_t w = blockIdx.x; w < n_chunk; w += gridDim.x) {
    // bin of chunk w: last b with cstart[b] <= w (it has >= 1 chunk)
    int lo = 0, hi = B.nb;  // cstart[lo] <= w < cstart[hi]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (B.cstart[mid] <= w) lo = mid; else hi = mid;
    }
    const int b = lo;
    const uint32_t k = w - B.cstart[b];
    const uint32_t nc = B.cstart[b + 1] - B.cstart[b];
    const uint32_t e_end = B.bstart[b + 1];
    const uint32_t e0 = B.bstart[b] + k * ch;
    const uint32_t e1 = (e_end - e0 < ch) ? e_end : e0 + ch;
    for (int i = threadIdx.x; i < kBinLen; i += kReduceBlock) s_a[i] = 0.0f;
    __syncthreads();
    uint32_t e = e0 + threadIdx.x;
    for (; e + 3u * kReduceBlock < e1; e += 4u * kReduceBlock) {
      unsigned long long v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = B.ev[e + (uint32_t)u * kReduceBlock];
#pragma unroll
      for (int u = 0; u < 4; u++)
        lds_add_f32(&s_a[(uint32_t)v[u]],
                    __uint_as_float((uint32_t)(v[u] >> 32)));
    }
    for (; e < e1; e += kReduceBlock) {
      unsigned long long v = B.ev[e];
      lds_add_f32(&s_a[(uint32_t)v], __uint_as_float((uint32_t)(v >> 32)));
    }
    __syncthreads();
    const long long a0 = (long long)b * kBinLen;
    const long long lim_ll = B.acc_n - a0;
    const int lim = lim_ll < kBinLen ? (int)lim_ll : kBinLen;
    if (nc == 1) {
      for (int i = threadIdx.x; i < lim; i += kReduceBlock) {
        float x = s_a[i];
        if (x != 0.0f) acc[a0 + i] += x;
      }
    } else {
      float *pp = B.part + (size_t)w * kBinLen;
      for (int i = threadIdx.x; i < kBinLen; i += kReduceBlock) pp[i] = s_a[i];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void bin_sum_kernel(
    BinArgs B, float *acc, const StepClock *clock, StepClock *clock_next) {
  clock_step(clock, clock_next);
  if (B.meta[1] == 0u) return;  // every bin was summed by one chunk
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
       i < B.acc_n; i += (long long)gridDim.x * kBlock) {
    const int b = (int)(i >> kBinShift);
    const uint32_t c0 = B.cstart[b], c1 = B.cstart[b + 1];
    if (c1 - c0 <= 1u) continue;
    const int j = (int)(i & (kBinLen - 1));
    float sum = 0.0f;
    for (uint32_t c = c0; c < c1; c++) sum += B.part[(size_t)c * kBinLen + j];
    acc[i] += sum;
  }
}

// ---------------------------------------------------------------------
// Tiled delivery (non-plastic).  Block b owns accumulator tile b / G in
// LDS and takes share b % G of the step's (spike, delay) pairs; for each
// pair it reads only the records of its tile (a contiguous group of the
// tile-major store), so every record is read once per step by one block.
// The same pair enumeration as the other delivery kernels, over the G
// blocks of the tile.
// ---------------------------------------------------------------------
struct TileLoc {
  const uint32_t *off;
  uint32_t T, t;
  __device__ __forceinline__ void operator()(const DeliveryArgs &A,
                                             uint32_t src, int d,
                                             uint32_t &cnt,
                                             unsigned long long &start) const {
    const size_t k = ((size_t)src * (size_t)A.D + (size_t)d) * T + t;
    const uint32_t o0 = off[k], o1 = off[k + T];
    cnt = o1 - o0;
    start = o0;
  }
};

// FIXED: 64-bit fixed-point sums with LDS integer atomics (no return
// value, no retry loop); else float sums by compare-and-swap.
template <bool FIXED>
__global__ __launch_bounds__(kTiledBlock) void tiled_delivery_kernel(
    DeliveryArgs A0, TileArgs B) {
  const DeliveryArgs A = clocked(A0);
  __shared__ uint32_t s_part[kTiledBlock];
  __shared__ int s_big;   // a float atomic went to the accumulator
  // dynamic LDS: the tile (2^tile_shift floats or 64-bit sums), then the
  // pair prefix
  const int TL = 1 << B.tile_shift;
  float *s_acc = reinterpret_cast<float *>(s_dyn);
  unsigned long long *s_q = reinterpret_cast<unsigned long long *>(s_dyn);
  uint32_t *s_pre = s_dyn + (FIXED ? 2 * TL : TL);
  const uint32_t tk = B.blk[blockIdx.x];
  const uint32_t t = tk & 0xFFFFu, k = tk >> 16;
  const uint32_t g0 = B.tg[t], G = B.tg[t + 1] - g0;
  for (int i = threadIdx.x; i < TL; i += kTiledBlock) {
    if (FIXED)
      s_q[i] = 0ull;
    else
      s_acc[i] = 0.0f;
  }
  if (threadIdx.x == 0) s_big = 0;
  // (the prologue's barriers order the clearing before the first add)
  const uint32_t total =
      delivery_prologue<kTiledBlock>(A, blockIdx.x, s_pre, s_part);
  const long long a0 = (long long)t << B.tile_shift;
  auto add = [&](bool ok, unsigned long long, unsigned long long r, float hj,
                 uint32_t) {
    if (!ok) return;
    const uint32_t tp = (uint32_t)r;
    const long long idx = (long long)(tp >> kPortShift) * A.n_total +
                          (tp & kTargetMask);
    const float v = hj * __uint_as_float((uint32_t)(r >> 32));
    if (FIXED) {
      if (fabsf(hj) <= B.h_max) {
        // exact: v * 2^F is v with a larger exponent, below 2^62
        const long long q = __float2ll_rn(v * B.q_scale);
        atomicAdd(&s_q[idx - a0], (unsigned long long)q);
      } else {
        unsafeAtomicAdd(A.acc + idx, v);
        s_big = 1;
      }
    } else {
      lds_add_f32(&s_acc[idx - a0], v);
    }
  };
  const unsigned long long n_events =
      deliver_events<4, kTiledBlock, decltype(add) &, TileLoc>(
          A, k, G, s_pre, total, add, B.trec,
          TileLoc{B.off, (uint32_t)B.T, t});
  __syncthreads();
  const long long lim_ll = B.acc_n - a0;
  const int lim = lim_ll < TL ? (int)lim_ll : TL;
  if (G == 1) {
    // after a float atomic into this tile's entries the flush adds with
    // atomics too (they and plain accesses are not ordered)
    const bool big = FIXED && s_big;
    for (int i = threadIdx.x; i < lim; i += kTiledBlock) {
      const float x =
          FIXED ? (float)(long long)s_q[i] * B.q_inv : s_acc[i];
      if (x == 0.0f) continue;
      if (big)
        unsafeAtomicAdd(A.acc + a0 + i, x);
      else
        A.acc[a0 + i] += x;
    }
  } else if (FIXED) {
    long long *pq = B.part_q + (size_t)(g0 + k) * TL;
    for (int i = threadIdx.x; i < TL; i += kTiledBlock)
      pq[i] = (long long)s_q[i];
  } else {
    float *pp = B.part + (size_t)(g0 + k) * TL;
    for (int i = threadIdx.x; i < TL; i += kTiledBlock) pp[i] = s_acc[i];
  }
  if (lane_id() == 0 && n_events)
    atomicAdd(A.events + (blockIdx.x & (kCounterSlots - 1)), n_events);
}

// Sum of the G partial tiles of every accumulator entry (G > 1); the last
// launch of the step (writes the next step's clock when captured).
__global__ __launch_bounds__(kBlock) void tiled_sum_kernel(
    TileArgs B, float *acc, const StepClock *clock, StepClock *clock_next) {
  clock_step(clock, clock_next);
  if (!B.split) return;
  // four entries a thread (16-byte loads; a tile holds a multiple of 4),
  // the partials of each entry added in block order
  const long long TL = 1ll << B.tile_shift;
  const long long n4 = (B.acc_n + 3) >> 2;
  for (long long i4 = (long long)blockIdx.x * kBlock + threadIdx.x; i4 < n4;
       i4 += (long long)gridDim.x * kBlock) {
    const long long i = i4 << 2;
    const long long t = i >> B.tile_shift;
    const uint32_t g0 = B.tg[t], G = B.tg[t + 1] - g0;
    if (G <= 1) continue;   // that tile's block added its sums
    const size_t o = (size_t)g0 * TL + (size_t)(i & (TL - 1));
    float sum[4];
    if (B.fixed) {
      long long q[4] = {0, 0, 0, 0};
#pragma unroll 4
      for (uint32_t k = 0; k < G; k++) {
        const longlong2 *p =
            reinterpret_cast<const longlong2 *>(B.part_q + o + (size_t)k * TL);
        const longlong2 a = p[0], b = p[1];
        q[0] += a.x;
        q[1] += a.y;
        q[2] += b.x;
        q[3] += b.y;
      }
      for (int j = 0; j < 4; j++) sum[j] = (float)q[j] * B.q_inv;
    } else {
      float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 8
      for (uint32_t k = 0; k < G; k++) {
        const float4 x =
            *reinterpret_cast<const float4 *>(B.part + o + (size_t)k * TL);
        s.x += x.x;
        s.y += x.y;
        s.z += x.z;
        s.w += x.w;
      }
      sum[0] = s.x;
      sum[1] = s.y;
      sum[2] = s.z;
      sum[3] = s.w;
    }
    for (int j = 0; j < 4; j++)
      if (i + j < B.acc_n && sum[j] != 0.0f) acc[i + j] += sum[j];
  }
}

// ---------------------------------------------------------------------
// Post-synaptic STDP (src/rev_spike.cu:41-75): every node whose spike
// reaches its dendrites this step updates its plastic incoming synapses.
// One wavefront per candidate node; lanes stride over its inputs.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void rev_kernel(RevArgs A0) {
  clock_step(A0.clock, A0.clock_next);
  const RevArgs A = clocked(A0);
  // one block per spiking node: its incoming plastic synapses are
  // independent (each has this node as its only target), so all 256 lanes
  // walk the reverse index together
  for (int k = 0; k <= A.max_den; k++) {
    long long s = A.gs - k;
    if (s < 0) break;
    long long slot = s % A.R;
    uint32_t n = A.ring_n[slot];
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
      uint32_t node = A.ring_src[slot * A.cap + e];
      if (A.den_idx[node] != k) continue;
      if (A.last_rev[node] != A.T) continue;
      unsigned long long r0 = A.rev_ptr[node], r1 = A.rev_ptr[node + 1];
      uint16_t ti = (uint16_t)(A.T & 0xffff);
      for (unsigned long long r = r0 + threadIdx.x; r < r1; r += kBlock) {
        uint32_t syn = A.rev_idx[r];
        int sgp = A.syn_group[syn];
        if (sgp <= 0) continue;
        uint16_t st = A.spike_time[syn];
        uint16_t Dt_int = (uint16_t)(ti - st);
        if (Dt_int < kMaxSynDt) {
          uint32_t *wp = reinterpret_cast<uint32_t *>(A.rec + syn) + 1;
          float w = __uint_as_float(*wp);
          float wn = dev_syn_update(A.syn_type[sgp - 1], w,
                                    A.dt * (float)Dt_int,
                                    A.stdp_param + 8 * (sgp - 1));
          *wp = __float_as_uint(wn);
        }
      }
    }
  }
}

// Spikes of nodes with a dendritic delay of k > 0 steps reach the
// dendrites k steps after emission (src/spike_buffer.cu:276-278,299-301):
// mark them before this step's delivery reads last_rev.
__global__ __launch_bounds__(kBlock) void den_rev_kernel(RevArgs A0) {
  const RevArgs A = clocked(A0);
  for (int k = 1; k <= A.max_den; k++) {
    long long s = A.gs - k;
    if (s < 0) break;
    long long slot = s % A.R;
    uint32_t n = A.ring_n[slot];
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < n;
         e += gridDim.x * kBlock) {
      uint32_t node = A.ring_src[slot * A.cap + e];
      if (A.den_idx[node] == k) A.last_rev[node] = A.T;
    }
  }
}

// ---------------------------------------------------------------------
// Step clock of captured steps
// ---------------------------------------------------------------------
__global__ void clock_set_kernel(StepClock *c, StepClock v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *c = v;
}

__global__ void record_kernel(RecordArgs A) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.n_col) return;
  float v;
  float *p = A.ptr[c];
  if (p) {
    v = *p;
  } else {
    int node = A.node[c];
    v = (A.last_spike[node] == A.T) ? A.last_height[node] : 0.0f;
  }
  A.out[c] = v;
}

__global__ void wrap_kernel(uint16_t *st, unsigned long long n, int up) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint16_t v = st[i];
  if (up) {
    if (v >= 0x8000) st[i] = 0;
  } else {
    if (v < 0x8000) st[i] = 0x8000;
  }
}

__global__ void fill_u16_kernel(uint16_t *p, unsigned long long n, uint16_t v) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int P>
void launch_aeif_p(const AeifArgs &a, hipStream_t s) {
  int grid = (a.n + kBlock - 1) / kBlock;
  if (grid <= 0) return;
  if (a.do_step && a.strag) {
    // the overlap bookkeeping is compiled in only where it is used (groups
    // of up to kOverlapMaxPort ports)
    // the default budget of one attempt, compiled in for uniform
    // parameters and up to two ports
    const bool b1 = a.uniform && a.budget == 1 && P <= 2;
    if constexpr (P <= kOverlapMaxPort) {
      if (a.skip_prev || a.mask_write) {
        if (b1)
          hipLaunchKernelGGL(
              (aeif_pass1_kernel<P, true, (P <= 2 ? 6 : 1), true, true>),
              dim3(grid), dim3(kBlock), 0, s, a);
        else if (a.uniform)
          hipLaunchKernelGGL((aeif_pass1_kernel<P, true, (P <= 2 ? 6 : 1), true>),
                             dim3(grid), dim3(kBlock), 0, s, a);
        else
          hipLaunchKernelGGL((aeif_pass1_kernel<P, false, 1, true>), dim3(grid),
                             dim3(kBlock), 0, s, a);
        return;
      }
    }
    if (b1)
      hipLaunchKernelGGL(
          (aeif_pass1_kernel<P, true, (P <= 2 ? 6 : 1), false, true>),
          dim3(grid), dim3(kBlock), 0, s, a);
    else if (a.uniform)
      hipLaunchKernelGGL((aeif_pass1_kernel<P, true, (P <= 2 ? 6 : 1), false>),
                         dim3(grid), dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL((aeif_pass1_kernel<P, false, 1, false>), dim3(grid),
                         dim3(kBlock), 0, s, a);
  } else {
    hipLaunchKernelGGL(aeif_kernel<P>, dim3(grid), dim3(kBlock), 0, s, a);
  }
}

template <int P, bool UNI, bool PROP, bool CU>
void launch_pass2_v(const AeifArgs &a, const DeliveryArgs *d, int grid2,
                    int d_grid, hipStream_t s) {
  if (d)
    hipLaunchKernelGGL((aeif_tail_kernel<P, UNI, PROP, CU>), dim3(grid2 + d_grid),
                       dim3(kBlock), delivery_lds_bytes(d->D, 0), s, a, *d,
                       (uint32_t)grid2);
  else
    hipLaunchKernelGGL((aeif_pass2_kernel<P, UNI, PROP, CU>), dim3(grid2),
                       dim3(kBlock), 0, s, a);
}

template <int P>
void launch_aeif_pass2_p(const AeifArgs &a, const DeliveryArgs *d,
                         int d_grid, hipStream_t s) {
  int grid2 = a.pass2_grid > 0 ? a.pass2_grid : 256;
  // the proposer's inference exists for groups of up to kPropMaxPort ports
  constexpr bool kProp = P <= kPropMaxPort;
  // the catch-up of neurons an overlapped step left
  if constexpr (P <= kOverlapMaxPort) {
    if (a.skip_prev || a.mask_write) {
      if (a.prop_pk && kProp) {
        if (a.uniform)
          launch_pass2_v<P, true, kProp, true>(a, d, grid2, d_grid, s);
        else
          launch_pass2_v<P, false, kProp, true>(a, d, grid2, d_grid, s);
      } else if (a.uniform) {
        launch_pass2_v<P, true, false, true>(a, d, grid2, d_grid, s);
      } else {
        launch_pass2_v<P, false, false, true>(a, d, grid2, d_grid, s);
      }
      return;
    }
  }
  if (a.prop_pk && kProp) {
    if (a.uniform)
      launch_pass2_v<P, true, kProp, false>(a, d, grid2, d_grid, s);
    else
      launch_pass2_v<P, false, kProp, false>(a, d, grid2, d_grid, s);
  } else {
    if (a.uniform)
      launch_pass2_v<P, true, false, false>(a, d, grid2, d_grid, s);
    else
      launch_pass2_v<P, false, false, false>(a, d, grid2, d_grid, s);
  }
}

}  // namespace

void LaunchPropEval(const unsigned char *pk, const float *feat, int n,
                    float *out, void *stream) {
  int grid = (n + kBlock - 1) / kBlock;
  if (grid > 0)
    hipLaunchKernelGGL(prop_eval_kernel, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, pk, feat, n, out);
}

// Staging of the allocation-free inference: the feature vectors and the
// outputs of one call, in a file-scope device array (one copy per GPU, held
// until the process ends).  Host code never uses the array's name as an
// address: each device's copy is found once with hipGetSymbolAddress.
namespace {
constexpr size_t kPropScratchFloats = (size_t)kPropEvalNoAllocMax * (kPropIn + 1);
__device__ __align__(256) float prop_scratch[kPropScratchFloats];
constexpr int kMaxDevices = 64;
float *prop_scratch_dev[kMaxDevices];   // per device, null until looked up
}  // namespace

const char *PropEvalNoAlloc(const unsigned char *pk, const float *feat,
                            int n, float *out, void *stream) {
  if (n < 0) return "negative number of feature vectors";
  if (n > kPropEvalNoAllocMax)
    return "more feature vectors than the allocation-free inference holds "
           "(kPropEvalNoAllocMax)";
  if (n == 0) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return "no usable HIP device";
  if (!prop_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(prop_scratch)) != hipSuccess ||
        !p)
      return "hipGetSymbolAddress of the inference staging failed";
    prop_scratch_dev[dev] = static_cast<float *>(p);
  }
  float *f_dev = prop_scratch_dev[dev];
  float *o_dev = f_dev + (size_t)kPropEvalNoAllocMax * kPropIn;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(f_dev, feat, (size_t)n * kPropIn * sizeof(float),
                     hipMemcpyHostToDevice, s) != hipSuccess)
    return "copy of the feature vectors failed";
  LaunchPropEval(pk, f_dev, n, o_dev, stream);
  if (hipGetLastError() != hipSuccess) return "inference launch failed";
  if (hipMemcpyAsync(out, o_dev, (size_t)n * sizeof(float),
                     hipMemcpyDeviceToHost, s) != hipSuccess)
    return "copy of the outputs failed";
  if (hipStreamSynchronize(s) != hipSuccess) return "inference failed";
  return nullptr;
}

void LaunchAeifPass2(const AeifArgs &a, int n_port, const DeliveryArgs *d,
                     int d_grid, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  switch (n_port) {
#define NGPU_CASE(P) \
  case P: launch_aeif_pass2_p<P>(a, d, d_grid, s); break;
    NGPU_CASE(1) NGPU_CASE(2) NGPU_CASE(3) NGPU_CASE(4) NGPU_CASE(5)
    NGPU_CASE(6) NGPU_CASE(7) NGPU_CASE(8) NGPU_CASE(9) NGPU_CASE(10)
    NGPU_CASE(11) NGPU_CASE(12) NGPU_CASE(13) NGPU_CASE(14) NGPU_CASE(15)
    NGPU_CASE(16) NGPU_CASE(17) NGPU_CASE(18) NGPU_CASE(19) NGPU_CASE(20)
#undef NGPU_CASE
    default: break;
  }
}

void LaunchAeif(const AeifArgs &a, int n_port, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  switch (n_port) {
#define NGPU_CASE(P) \
  case P: launch_aeif_p<P>(a, s); break;
    NGPU_CASE(1) NGPU_CASE(2) NGPU_CASE(3) NGPU_CASE(4) NGPU_CASE(5)
    NGPU_CASE(6) NGPU_CASE(7) NGPU_CASE(8) NGPU_CASE(9) NGPU_CASE(10)
    NGPU_CASE(11) NGPU_CASE(12) NGPU_CASE(13) NGPU_CASE(14) NGPU_CASE(15)
    NGPU_CASE(16) NGPU_CASE(17) NGPU_CASE(18) NGPU_CASE(19) NGPU_CASE(20)
#undef NGPU_CASE
    default: break;
  }
}

void LaunchIaf(const IafArgs &a, void *stream) {
  int grid = (a.n + kBlock - 1) / kBlock;
  if (grid > 0)
    hipLaunchKernelGGL(iaf_kernel, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
}

void LaunchRelay(const RelayArgs &a, void *stream) {
  int grid = (a.n + kBlock - 1) / kBlock;
  if (grid > 0)
    hipLaunchKernelGGL(relay_kernel, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
}

void LaunchLegacyPoisson(const LegacyPoisArgs &a, void *stream) {
  int grid = (a.n + kBlock - 1) / kBlock;
  if (grid > 0)
    hipLaunchKernelGGL(legacy_pois_kernel, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
}

void LaunchSpikeGen(const SpikeGenArgs &a, void *stream) {
  int grid = (a.n + kBlock - 1) / kBlock;
  if (grid > 0)
    hipLaunchKernelGGL(spikegen_kernel, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
}

void LaunchDelivery(const DeliveryArgs &a, int grid, void *stream) {
  const size_t lds = delivery_lds_bytes(a.D, 0);
  if (a.syn_group)
    hipLaunchKernelGGL(delivery_kernel<true>, dim3(grid), dim3(kBlock), lds,
                       (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(delivery_kernel<false>, dim3(grid), dim3(kBlock), lds,
                       (hipStream_t)stream, a);
}

void LaunchDeliveryTiled(const DeliveryArgs &a, const TileArgs &b,
                         void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t lds =
      sizeof(uint32_t) * (((size_t)(b.fixed ? 2 : 1) << b.tile_shift) +
                          (size_t)a.D + 1);
  static bool lds_attr = false;   // the tile may exceed the default limit
  if (!lds_attr) {
    // the largest tiles: 2^15 floats, or 2^14 64-bit sums (the host uses
    // fixed-point sums only where the tile fits in LDS)
    const int most = (int)(sizeof(uint32_t) *
                           (((size_t)1 << kTileShiftMax) + kMaxDelaySlots + 1));
    (void)hipFuncSetAttribute(
        reinterpret_cast<const void *>(tiled_delivery_kernel<false>),
        hipFuncAttributeMaxDynamicSharedMemorySize, most);
    (void)hipFuncSetAttribute(
        reinterpret_cast<const void *>(tiled_delivery_kernel<true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, most);
    lds_attr = true;
  }
  if (b.fixed)
    hipLaunchKernelGGL(tiled_delivery_kernel<true>, dim3((unsigned)b.n_blocks),
                       dim3(kTiledBlock), lds, s, a, b);
  else
    hipLaunchKernelGGL(tiled_delivery_kernel<false>,
                       dim3((unsigned)b.n_blocks), dim3(kTiledBlock), lds, s,
                       a, b);
  long long g = ((b.acc_n + 3) / 4 + kBlock - 1) / kBlock;
  if (g > 2048) g = 2048;
  if (!b.split) g = 1;
  hipLaunchKernelGGL(tiled_sum_kernel, dim3((unsigned)g), dim3(kBlock), 0, s,
                     b, a.acc, a.clock, a.clock_next);
}

void LaunchDeliveryBinned(const DeliveryArgs &a, const BinArgs &b,
                          int reduce_grid, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t lds = delivery_lds_bytes(a.D, b.nsub);
  hipLaunchKernelGGL(bin_count_kernel, dim3(b.grid), dim3(kBinBlock), lds, s,
                     a, b);
  hipLaunchKernelGGL(bin_scan_kernel, dim3(b.nsub), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(b.grid), dim3(kBinBlock), lds, s,
                     a, b);
  hipLaunchKernelGGL(bin_reduce_kernel, dim3(reduce_grid), dim3(kReduceBlock),
                     0, s, b, a.acc);
  long long g = (b.acc_n + kBlock - 1) / kBlock;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(bin_sum_kernel, dim3((unsigned)g), dim3(kBlock), 0, s, b,
                     a.acc, a.clock, a.clock_next);
}

void LaunchClockSet(StepClock *c, const StepClock &v, void *stream) {
  hipLaunchKernelGGL(clock_set_kernel, dim3(1), dim3(64), 0,
                     (hipStream_t)stream, c, v);
}

void LaunchRevStdp(const RevArgs &a, int grid, void *stream) {
  hipLaunchKernelGGL(rev_kernel, dim3(grid), dim3(kBlock), 0,
                     (hipStream_t)stream, a);
}

void LaunchDenRev(const RevArgs &a, int grid, void *stream) {
  hipLaunchKernelGGL(den_rev_kernel, dim3(grid), dim3(kBlock), 0,
                     (hipStream_t)stream, a);
}

void LaunchRecord(const RecordArgs &a, void *stream) {
  int grid = (a.n_col + 255) / 256;
  if (grid > 0)
    hipLaunchKernelGGL(record_kernel, dim3(grid), dim3(256), 0,
                       (hipStream_t)stream, a);
}

void LaunchSpikeTimeWrap(uint16_t *spike_time, unsigned long long n, int up,
                         void *stream) {
  unsigned long long grid = (n + 255) / 256;
  if (grid > 0)
    hipLaunchKernelGGL(wrap_kernel, dim3((unsigned)grid), dim3(256), 0,
                       (hipStream_t)stream, spike_time, n, up);
}

void LaunchFillU16(uint16_t *p, unsigned long long n, uint16_t v, void *stream) {
  unsigned long long grid = (n + 255) / 256;
  if (grid > 0)
    hipLaunchKernelGGL(fill_u16_kernel, dim3((unsigned)grid), dim3(256), 0,
                       (hipStream_t)stream, p, n, v);
}

}  // namespace ngpu
