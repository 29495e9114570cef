// Node update of the fixed-step point models, one neuron per lane:
//
//   iaf_psc_exp_g   src/iaf_psc_exp_g.cu:34-61, 136-150 (exact integration,
//                   propagators from the group parameters)
//   iaf_psc_exp_hc  src/iaf_psc_exp_hc.cu:29-53 with the propagators of
//                   src/iaf_psc_exp_hc_params.h (double constants, so the
//                   update is evaluated in double)
//   izhikevich      src/izhikevich.cu:35-72 (forward Euler; the input is
//                   added to V_m and cleared)
//   izhikevich_psc_exp, _2s, _5s
//                   src/izhikevich_psc_exp.cu:40-76,
//                   src/izhikevich_psc_exp_2s.cu:41-83 (forward Euler on 1,
//                   2 or 5 sub-steps of dt, exponentially decaying current)
//   ext_neuron      src/ext_neuron.cu:23-39 (per-port input moved to
//                   port_value for the host; no spikes)
//
// The iaf and izhikevich models have one input port whose accumulator is
// added to I_syn with weight 1 before the update.  The izhikevich membrane equations carry the
// reference's double literals, so they are evaluated in double with the
// reference's operation order; everything else is FP32.
#include <hip/hip_runtime.h>

#include "ngpu_device.h"
#include "ngpu_kernels.h"
#include "step_common.h"

namespace ngpu {

namespace {

// src/iaf_psc_exp_hc_params.h
constexpr double kHcP11 = 8.1873075E-01;
constexpr double kHcP22 = 9.9004983E-01;
constexpr double kHcP21 = 3.6067175E-04;
constexpr double kHcP20 = 3.9800665E-04;
constexpr double kHcTheta = 1.5000000E+01;
constexpr double kHcVReset = 0.0;
constexpr int kHcRefrSteps = 20;

// dV/dt of the Izhikevich model without the input term, in double and in
// the reference's order: 0.04 v v + 5 v + 140 - u
__device__ __forceinline__ double izh_drive(float v, float u) {
  double vd = (double)v;
  double s = 0.04 * vd;
  s = s * vd;
  s = s + 5.0 * vd;
  s = s + 140.0;
  s = s - (double)u;
  return s;
}

template <int KIND, int STEPS>
__global__ __launch_bounds__(kBlock) void point_kernel(PointArgs A) {
  const StepNow sn = step_now(A.out, &A.fold, A.clock);
  int i = blockIdx.x * kBlock + threadIdx.x;
  bool active = i < A.n;
  int ii = active ? i : 0;
  int node = A.i0 + ii;
  int nsp = 0;
  if (active) {
    float in[1];
    gather_input<1>(A.fold, sn, node, in);
    float I = A.I_syn[i] + in[0];
    if (A.do_step) {
      float V = A.V[i];
      float refr = A.refr[i];
      const float I_e = pf(A.I_e, i);
      if (KIND == kPointIafG || KIND == kPointIafHc) {
        if (refr > 0.0f) {
          refr -= 1.0f;
        } else if (KIND == kPointIafHc) {
          double v = (double)V * kHcP22;
          v = v + (double)I * kHcP21;
          v = v + (double)I_e * kHcP20;
          V = (float)v;
        } else {
          float v = V * A.P22;
          v = __builtin_fmaf(I, A.P21, v);
          v = __builtin_fmaf(I_e, A.P20, v);
          V = v;
        }
        if (KIND == kPointIafHc)
          I = (float)((double)I * kHcP11);
        else
          I = I * A.P11;
        const bool fire = KIND == kPointIafHc ? (double)V >= kHcTheta
                                              : V >= A.Theta;
        if (fire) {
          nsp = 1;
          V = KIND == kPointIafHc ? (float)kHcVReset : A.V_reset;
          refr = (float)(KIND == kPointIafHc ? kHcRefrSteps : A.n_refr);
        }
      } else {
        float u = A.u[i];
        const float ha = A.h * A.a;
        if (KIND == kPointIzh) {
          if (refr > 0.0f) {
            refr -= 1.0f;
          } else {
            const float v0 = V, u0 = u;
            double t = (double)A.h * (izh_drive(v0, u0) + (double)I_e);
            t = t + (double)I;
            V = (float)((double)V + t);
            u = u + ha * (A.b * v0 - u0);
          }
          I = 0.0f;
        } else {
          if (refr > 0.0f) {
            refr -= 1.0f;
#pragma unroll
            for (int s = 0; s < STEPS; s++) I = I * A.C_syn;
          } else {
#pragma unroll
            for (int s = 0; s < STEPS; s++) {
              const float v0 = V, u0 = u;
              double r = izh_drive(v0, u0) + (double)I;
              r = r + (double)I_e;
              V = (float)((double)V + (double)A.h * r);
              u = u + ha * (A.b * v0 - u0);
              I = I * A.C_syn;
            }
          }
        }
        if (V >= A.V_th) {
          nsp = 1;
          V = A.c;
          u = u + A.d;
          refr = (float)(A.n_refr < 0 ? 0 : A.n_refr);
        }
        A.u[i] = u;
      }
      A.V[i] = V;
      A.refr[i] = refr;
    }
    A.I_syn[i] = I;
  }
  if (!A.do_step) return;
  emit_spikes(A.out, sn, active, ii, node, nsp, 1.0f);
}

// ext_neuron, one thread per (node, port), nodes fastest
__global__ __launch_bounds__(kBlock) void ext_kernel(ExtArgs A) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)A.n * A.n_port) return;
  const int p = (int)(t / A.n), i = (int)(t % A.n);
  const long long pois_T = A.clock ? A.clock->pois_T : A.fold.pois_T;
  const double pois_t = A.clock ? A.clock->pois_t : A.fold.pois_t;
  const float in = gather_port(A.fold, pois_T, pois_t, A.i0 + i, p);
  float *pin = A.var + (size_t)(2 * p) * A.n + i;
  float x = __builtin_fmaf(in, A.port_weight[(size_t)p * A.n + i], *pin);
  if (A.do_step) {
    A.var[(size_t)(2 * p + 1) * A.n + i] = x;
    x = 0.0f;
  }
  *pin = x;
}

template <int KIND, int STEPS>
void launch(const PointArgs &a, hipStream_t s) {
  hipLaunchKernelGGL((point_kernel<KIND, STEPS>),
                     dim3((a.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
}

}  // namespace

const char *LaunchPoint(const PointArgs &a, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (a.n <= 0) return nullptr;
  switch (a.kind) {
    case kPointIafG: launch<kPointIafG, 1>(a, s); return nullptr;
    case kPointIafHc: launch<kPointIafHc, 1>(a, s); return nullptr;
    case kPointIzh: launch<kPointIzh, 1>(a, s); return nullptr;
    case kPointIzhPsc:
      if (a.substeps == 1) { launch<kPointIzhPsc, 1>(a, s); return nullptr; }
      if (a.substeps == 2) { launch<kPointIzhPsc, 2>(a, s); return nullptr; }
      if (a.substeps == 5) { launch<kPointIzhPsc, 5>(a, s); return nullptr; }
      return "izhikevich_psc_exp: sub-steps must be 1, 2 or 5";
    default:
      return "unknown point model";
  }
}

void LaunchExt(const ExtArgs &a, void *stream) {
  long long total = (long long)a.n * a.n_port;
  if (total <= 0) return;
  hipLaunchKernelGGL(ext_kernel, dim3((unsigned)((total + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
}

}  // namespace ngpu
