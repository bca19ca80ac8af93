// Launch interface of the narrow-band network kernels (host side).
//
// A Fourier-feature MLP (99 features padded to 112, three hidden layers of
// 64 ReLU units, one output) predicts the truncated SDF at 4^3-brick
// centres.  Matrix products run on MFMA (v_mfma_f32_16x16x16_bf16): bf16
// weights and activations, f32 accumulation; master weights and the Adam
// moments are f32.  The parameter vector layout:
//   W1[64][112] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iso {
namespace mlp {

constexpr int kHidden = 64;
constexpr int kFreq = 16;
constexpr int kFeat = 3 + 6 * kFreq;  // 99
constexpr int kFeatPad = 112;         // 7 K-blocks of 16
constexpr int kKB1 = kFeatPad / 16;

constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kFeatPad;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden * kHidden;
constexpr int kOffW4 = kOffB3 + kHidden;
constexpr int kOffB4 = kOffW4 + kHidden;
constexpr int kParams = kOffB4 + 1;

// bf16 MFMA operand fragments: F1 (W1), F2 (W2), F3 (W3), T2 (W2^T), T3 (W3^T),
// each [m-block][k-block][lane][4] with element
// W[16 mb + lane % 16][16 kb + 4 (lane / 16) + j].
constexpr int kFragF1 = 0;
constexpr int kFragF2 = kFragF1 + kHidden * kFeatPad;
constexpr int kFragF3 = kFragF2 + kHidden * kHidden;
constexpr int kFragT2 = kFragF3 + kHidden * kHidden;
constexpr int kFragT3 = kFragT2 + kHidden * kHidden;
constexpr int kTrainFrags = kFragT3 + kHidden * kHidden;  // F1 F2 F3 T2 T3
// F1I: W1 with its columns in the inference forward's feature order (row
// 16 kb + 4 g + j of the features holds, for kb < 6, sin (kb < 3) or cos of
// coordinate kb % 3 at frequency 4 j + g; rows 96-98 the coordinates; the
// rest padding), so each K block's features share one function per
// wavefront.
constexpr int kFragF1I = kTrainFrags;
constexpr int kFrags = kFragF1I + kHidden * kFeatPad;

// Device state of one model.
struct Model {
  float* theta;     // kParams master weights
  float* m;         // Adam first moments
  float* v;         // Adam second moments
  float* grad;      // kParams gradient accumulator (cleared by the Adam step)
  uint16_t* frags;  // kFrags bf16 fragments of theta
  float* loss;      // [0]: sum of weighted squared errors of the last batch
};

// Activations of one training batch, hidden-major bf16 ([unit][sample]).
struct Batch {
  uint16_t* x;    // kFeatPad x n
  uint16_t* h1;   // 64 x n
  uint16_t* h2;
  uint16_t* h3;
  uint16_t* dz1;  // 64 x n
  uint16_t* dz2;
  uint16_t* dz3;
  float* g;       // n: dL/df per sample
  int64_t n;      // batch size, a multiple of 16
};

// Training samples (device): points (3 per sample), SDF targets, weights
// (null: all 1).  Batch sample i of step `step` is sample hash(seed, step, i) % count.
struct Samples {
  const float* points;
  const float* sdf;
  const float* weight;
  uint64_t count;
  float clamp;  // targets are clamp(sdf, -clamp, clamp)
  float overshoot;  // weight of the error beyond the interval between 0 and the target
  uint32_t seed;
  int32_t gather;  // 0: batch sample i is sample i (exact gradients for tests)
};

// Writes the bf16 fragments of theta.
hipError_t launch_pack(const Model& M, hipStream_t s);
// f for n points (3 floats each), written to out.
hipError_t launch_forward(const Model& M, const float* points, int64_t n, float* out,
                          hipStream_t s);
// Forward + backward-data of one batch (step `step`) and the weight
// gradients accumulated into M.grad; M.loss[0] gets the batch loss sum.
hipError_t launch_train_batch(const Model& M, const Batch& B, const Samples& S, uint32_t step,
                              hipStream_t s);
// Adam step (bias-corrected, step counted from 1) with learning rate lr;
// clears the gradient and refreshes the fragments.
hipError_t launch_adam(const Model& M, uint32_t step, float lr, float beta1, float beta2,
                       float eps, hipStream_t s);
// Clears M.grad and M.loss.
hipError_t launch_zero_grad(const Model& M, hipStream_t s);

}  // namespace mlp
}  // namespace iso
