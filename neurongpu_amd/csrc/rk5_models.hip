// Node update of the other RK5 neuron models: one neuron per lane, the
// whole time step in one launch (adaptive Cash-Karp substeps with the
// model's ExternalUpdate after each, as RK5Update does).
//
//   aeif_psc_alpha  src/aeif_psc_alpha_kernel.h:132-201, aeif_psc_alpha.cu:24-68
//   aeif_psc_exp    src/aeif_psc_exp_kernel.h:126-193, aeif_psc_exp.cu:24-62
//   aeif_psc_delta  src/aeif_psc_delta_kernel.h:118-180, aeif_psc_delta.cu:24-64
//   izhikevich_cond_beta  src/izhikevich_cond_beta_kernel.h:145-206,
//                   izhikevich_cond_beta.cu:24-86
//   RK5Update       src/rk5.h:149-162
//
// State and parameters are the group's SoA arrays [index][n] in the
// reference's enum order; every parameter is read per neuron.  The
// derivatives multiply by precomputed reciprocals where the reference
// divides (as the AdEx conductance kernels do); izhikevich's membrane
// equation is evaluated in double, as the reference's double literals
// make it.
#include <hip/hip_runtime.h>

#include "ngpu_device.h"
#include "ngpu_kernels.h"
#include "step_common.h"

namespace ngpu {

namespace {

// AdEx with current-based synapses: alpha (I_syn, I1_syn per port),
// exponential (I_syn per port) or delta (input added to V_m).
template <int K, int P>
struct PscModel {
  static constexpr int VP =
      K == kRk5PscAlpha ? 2 : (K == kRk5PscExp ? 1 : 0);
  static constexpr int N = 2 + VP * P;
  float V_th, Delta_T, inv_Delta_T, g_L, E_L, inv_C_m, a, b, inv_tau_w, I_e,
      V_peak, V_reset, t_ref;
  float inv_tau_syn[P], w_in[P];

  __device__ __forceinline__ void load(const float *prm, int n, int i) {
    V_th = prm[0 * (size_t)n + i];
    Delta_T = prm[1 * (size_t)n + i];
    inv_Delta_T = 1.0f / Delta_T;
    g_L = prm[2 * (size_t)n + i];
    E_L = prm[3 * (size_t)n + i];
    inv_C_m = 1.0f / prm[4 * (size_t)n + i];
    a = prm[5 * (size_t)n + i];
    b = prm[6 * (size_t)n + i];
    inv_tau_w = 1.0f / prm[7 * (size_t)n + i];
    I_e = prm[8 * (size_t)n + i];
    V_peak = prm[9 * (size_t)n + i];
    V_reset = prm[10 * (size_t)n + i];
    t_ref = prm[11 * (size_t)n + i];
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (K == kRk5PscAlpha) {  // port params tau_syn, I0
        inv_tau_syn[q] = 1.0f / prm[(size_t)(14 + 2 * q) * n + i];
        w_in[q] = prm[(size_t)(15 + 2 * q) * n + i];
      } else if (K == kRk5PscExp) {  // port param tau_syn, input weight 1
        inv_tau_syn[q] = 1.0f / prm[(size_t)(14 + q) * n + i];
        w_in[q] = 1.0f;
      } else {
        inv_tau_syn[q] = 0.0f;
        w_in[q] = 1.0f;
      }
    }
  }

  __device__ __forceinline__ void rhs(float refr, const float (&y)[N],
                                      float (&dy)[N]) const {
    const bool rf = refr > 0.0f;
    const float V = rf ? V_reset : ((y[0] < V_peak) ? y[0] : V_peak);
    float I_syn = 0.0f;
    if (VP > 0) {
#pragma unroll
      for (int q = 0; q < P; q++) I_syn = I_syn + y[2 + VP * q];
    }
    const float V_spike =
        Delta_T == 0.0f ? 0.0f
                        : Delta_T * dev_expf((V - V_th) * inv_Delta_T);
    float t = (V - E_L) - V_spike;
    t = (-g_L) * t;
    t = t + I_syn;
    t = t - y[1];
    t = t + I_e;
    dy[0] = rf ? 0.0f : t * inv_C_m;
    float u = a * (V - E_L);
    u = u - y[1];
    dy[1] = u * inv_tau_w;
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (K == kRk5PscAlpha) {
        const float I = y[2 + 2 * q], I1 = y[3 + 2 * q];
        dy[3 + 2 * q] = (-I1) * inv_tau_syn[q];
        dy[2 + 2 * q] = __builtin_fmaf(-I, inv_tau_syn[q], I1);
      } else if (K == kRk5PscExp) {
        dy[2 + q] = (-y[2 + q]) * inv_tau_syn[q];
      }
    }
  }

  // port input of the previous step (src/get_spike.cu:116-140): alpha into
  // I1_syn times I0, exponential into I_syn, delta into V_m (weight 1)
  __device__ __forceinline__ void fold(float (&y)[N], const float (&in)[P]) {
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (K == kRk5PscAlpha)
        y[3 + 2 * q] = __builtin_fmaf(in[q], w_in[q], y[3 + 2 * q]);
      else if (K == kRk5PscExp)
        y[2 + q] = y[2 + q] + in[q];
      else
        y[0] = y[0] + in[q];
    }
  }

  static constexpr int kRefrParam = 12;
  __device__ __forceinline__ float reset_V() const { return V_reset; }
  __device__ __forceinline__ float thresh() const { return V_peak; }
  __device__ __forceinline__ float jump() const { return b; }
  __device__ __forceinline__ float tref() const { return t_ref; }
};

// Izhikevich neuron with beta-shaped conductances.
template <int P>
struct IzhModel {
  static constexpr int N = 2 + 2 * P;
  float V_th, a, b, c, d, I_e, t_ref;
  float E_rev[P], inv_tau_rise[P], inv_tau_decay[P], g0[P];

  __device__ __forceinline__ void load(const float *prm, int n, int i) {
    V_th = prm[0 * (size_t)n + i];
    a = prm[1 * (size_t)n + i];
    b = prm[2 * (size_t)n + i];
    c = prm[3 * (size_t)n + i];
    d = prm[4 * (size_t)n + i];
    I_e = prm[5 * (size_t)n + i];
    t_ref = prm[6 * (size_t)n + i];
#pragma unroll
    for (int q = 0; q < P; q++) {  // port params E_rev, tau_rise, tau_decay, g0
      E_rev[q] = prm[(size_t)(9 + 4 * q) * n + i];
      inv_tau_rise[q] = 1.0f / prm[(size_t)(10 + 4 * q) * n + i];
      inv_tau_decay[q] = 1.0f / prm[(size_t)(11 + 4 * q) * n + i];
      g0[q] = prm[(size_t)(12 + 4 * q) * n + i];
    }
  }

  __device__ __forceinline__ void rhs(float refr, const float (&y)[N],
                                      float (&dy)[N]) const {
    const bool rf = refr > 0.0f;
    const float V = rf ? c : y[0];
    float I_syn = 0.0f;
#pragma unroll
    for (int q = 0; q < P; q++)
      I_syn = __builtin_fmaf(y[2 + 2 * q], E_rev[q] - V, I_syn);
    const double Vd = (double)V;
    double t = 0.04 * Vd;
    t = t * Vd;
    t = __builtin_fma(5.0, Vd, t);
    t = t + 140.0;
    t = t - (double)y[1];
    t = t + (double)I_syn;
    t = t + (double)I_e;
    dy[0] = rf ? 0.0f : (float)t;
    dy[1] = a * __builtin_fmaf(b, V, -y[1]);
#pragma unroll
    for (int q = 0; q < P; q++) {
      const float g = y[2 + 2 * q], g1 = y[3 + 2 * q];
      dy[3 + 2 * q] = (-g1) * inv_tau_rise[q];
      dy[2 + 2 * q] = __builtin_fmaf(-g, inv_tau_decay[q], g1);
    }
  }

  __device__ __forceinline__ void fold(float (&y)[N], const float (&in)[P]) {
#pragma unroll
    for (int q = 0; q < P; q++)
      y[3 + 2 * q] = __builtin_fmaf(in[q], g0[q], y[3 + 2 * q]);
  }

  static constexpr int kRefrParam = 7;
  __device__ __forceinline__ float reset_V() const { return c; }
  __device__ __forceinline__ float thresh() const { return V_th; }
  __device__ __forceinline__ float jump() const { return d; }
  __device__ __forceinline__ float tref() const { return t_ref; }
};

template <class M, int P>
__global__ __launch_bounds__(kBlock) void rk5_model_kernel(Rk5ModelArgs A) {
  constexpr int N = M::N;
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  const double x1 = A.clock ? sn.t : A.x1;
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0, n_acc = 0, n_rej = 0;
  if (active) {
    M m;
    m.load(A.param, A.n, i);
    float y[N];
#pragma unroll
    for (int k = 0; k < N; k++) y[k] = A.var[(size_t)k * A.n + i];
    float in[P];
    gather_input<P>(A.fold, sn, node, in);
    m.fold(y, in);
    if (A.do_step) {
      float *refr_p = A.param + (size_t)M::kRefrParam * A.n + i;
      double x = A.x[i];
      float h = A.h[i];
      float refr = *refr_p;
      int n_att = 0;
      bool end = false;
      while (!end) {
        float hmax = (float)(x1 - x);
        if (!rk5_step(m, refr, x, y, h, A.h_min, A.h_min_acc, hmax, A.eps,
                      n_att, n_acc, n_rej, A.forced))
          x = x1;
        end = (x >= x1 - (double)A.h_min);
        // ExternalUpdate
        if (y[0] < -1.0e3f) {
          y[0] = m.reset_V();
          y[1] = 0.0f;
          continue;
        }
        if (y[1] < -1.0e6f || y[1] > 1.0e6f) {
          y[0] = m.reset_V();
          y[1] = 0.0f;
          continue;
        }
        if (refr > 0.0f) {
          y[0] = m.reset_V();
          if (end) refr -= 1.0f;
        } else if (y[0] >= m.thresh()) {
          nsp++;
          y[0] = m.reset_V();
          y[1] = y[1] + m.jump();
          float r = roundf(m.tref() / A.dt);
          int ri = (int)r;
          refr = (float)(ri < 0 ? 0 : ri);
        }
      }
      A.x[i] = x;
      A.h[i] = h;
      *refr_p = refr;
    }
#pragma unroll
    for (int k = 0; k < N; k++) A.var[(size_t)k * A.n + i] = y[k];
  }
  if (!A.do_step) return;
  emit_spikes(A.out, sn, active, ii, node, nsp, (float)nsp);
  rk5_count(A.cnt, n_acc, n_rej);
}

template <int P>
void launch_p(const Rk5ModelArgs &a, dim3 grid, hipStream_t s) {
  switch (a.kind) {
    case kRk5PscAlpha:
      hipLaunchKernelGGL((rk5_model_kernel<PscModel<kRk5PscAlpha, P>, P>),
                         grid, dim3(kBlock), 0, s, a);
      break;
    case kRk5PscExp:
      hipLaunchKernelGGL((rk5_model_kernel<PscModel<kRk5PscExp, P>, P>),
                         grid, dim3(kBlock), 0, s, a);
      break;
    case kRk5PscDelta:
      hipLaunchKernelGGL((rk5_model_kernel<PscModel<kRk5PscDelta, P>, P>),
                         grid, dim3(kBlock), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((rk5_model_kernel<IzhModel<P>, P>), grid,
                         dim3(kBlock), 0, s, a);
      break;
  }
}

}  // namespace

const char *LaunchRk5Model(const Rk5ModelArgs &a, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (a.n <= 0) return nullptr;
  dim3 grid((a.n + kBlock - 1) / kBlock);
  switch (a.n_port) {
    case 0:
    case 1: launch_p<1>(a, grid, s); break;
    case 2: launch_p<2>(a, grid, s); break;
    case 3: launch_p<3>(a, grid, s); break;
    case 4: launch_p<4>(a, grid, s); break;
    case 5: launch_p<5>(a, grid, s); break;
    case 6: launch_p<6>(a, grid, s); break;
    case 7: launch_p<7>(a, grid, s); break;
    case 8: launch_p<8>(a, grid, s); break;
    default: return "more receptor ports than this model supports";
  }
  return nullptr;
}

}  // namespace ngpu
