// Learned RK5 step-size proposer: a 6 -> 32 -> 32 -> 1 ReLU MLP that
// predicts, from a neuron's state at the start of an RK5 step, the largest
// step the embedded error estimate would accept.  The RK5 error test still
// accepts or rejects every attempt, so accuracy is that of the plain
// controller; only the number of attempts changes.
//
// It has no counterpart in the reference.  Used only in the second pass of
// the two-pass AdEx update (the neurons that need more than one attempt in
// a time step).
#pragma once
#include <stdint.h>

namespace ngpu {

constexpr int kPropIn = 6;
constexpr int kPropHid = 32;
// parameter layout (FP32 master copy and BF16 inference copy)
constexpr int kPropW1 = 0;                          // [32][6]
constexpr int kPropB1 = kPropW1 + kPropHid * kPropIn;   // [32]
constexpr int kPropW2 = kPropB1 + kPropHid;             // [32][32]
constexpr int kPropB2 = kPropW2 + kPropHid * kPropHid;  // [32]
constexpr int kPropW3 = kPropB2 + kPropHid;             // [32]
constexpr int kPropB3 = kPropW3 + kPropHid;             // [1]
constexpr int kPropNParam = kPropB3 + 1;                // 1313
constexpr int kPropSample = 8;   // floats per logged sample: 6 features,
                                 // target, spare
constexpr float kPropSafety = 0.9f;
// the second RK5 pass carries the proposer for groups of up to this many
// receptor ports
constexpr int kPropMaxPort = 4;
// trust region of a proposal relative to the plain controller's step
constexpr float kPropShrinkMax = 0.5f;
constexpr float kPropGrowMax = 32.0f;

// Inference copy in the lane layout of the 32x32x8 BF16 MFMA
// (v_mfma_f32_32x32x8_bf16): the hidden layers are computed as
// H^T = W X^T with the 64 samples of a wavefront as two 32-column blocks.
//   w1p[64][4]      bf16  A operand of layer 1: lane l holds
//                         W1[l % 32][4 (l / 32) + j], inputs 6, 7 zero
//   w2p[4][64][4]   bf16  A operands of layer 2, K chunk c: lane l holds
//                         W2[l % 32][8 c + 4 (l / 32) + j]
//   b1p, b2p, w3p [64][16] f32  in the accumulator layout: lane l, value r
//                         belongs to hidden unit 8 (r / 4) + 4 (l / 32) + r % 4
//   b3               f32
constexpr int kPropPkW1 = 0;                 // bytes
constexpr int kPropPkW2 = kPropPkW1 + 64 * 4 * 2;
constexpr int kPropPkB1 = kPropPkW2 + 4 * 64 * 4 * 2;
constexpr int kPropPkB2 = kPropPkB1 + 64 * 16 * 4;
constexpr int kPropPkW3 = kPropPkB2 + 64 * 16 * 4;
constexpr int kPropPkB3 = kPropPkW3 + 64 * 16 * 4;
constexpr int kPropPkBytes = kPropPkB3 + 16;

// Features of a neuron at the start of an RK5 step, from quantities the
// plain controller has at hand:
//   (V - E_L) / 10 mV, w / 100 pA, g_exc (nS), g_inh (nS, sum of ports >= 1),
//   log2(h / dt) of the step the controller would take, and
//   log10(err) / 4 of the last attempt's scaled error.
// Target of a sample (one RK5 attempt of size hh with scaled error err):
//   log2(hh * err^(-1/5) / dt), the step that would give err = 1.

// Training on logged samples (proposer.hip): minibatches of `batch` samples
// drawn with Philox from [0, n), mean squared error, SGD or Adam.
struct PropTrainArgs {
  const float *samples;        // [n][kPropSample]
  unsigned long long n;
  float *w;                    // FP32 master weights
  float *grad;                 // [kPropNParam] accumulated gradient
  float *loss;                 // [1] accumulated loss
  float *m, *v;                // Adam moments
  uint16_t *w_bf16;            // inference copy
  int batch;
  uint32_t iter;               // Philox counter of the minibatch
  uint32_t seed;
  float lr;
  int adam;
  int full_batch;              // 1: samples 0..batch-1 in order
};

const char *LaunchPropTrainStep(const PropTrainArgs &a, void *stream);
void LaunchPropToBf16(const float *w, uint16_t *w_bf16, void *stream);
// FP32 master weights -> MFMA-layout inference copy (kPropPkBytes)
void LaunchPropPack(const float *w, unsigned char *packed, void *stream);

}  // namespace ngpu
