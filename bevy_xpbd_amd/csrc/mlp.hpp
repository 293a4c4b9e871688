// Learned warm start: a per-vertex MLP (30 -> 64 -> 64 -> 3, GELU) that
// predicts the converged constraint correction of a substep from the
// predicted positions, so fewer Jacobi iterations reach a given residual.
// BF16 weights and activations on MFMA (v_mfma_f32_16x16x16_bf16), FP32
// accumulate.  No reference counterpart.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace xpbd {

constexpr int MLP_FEAT = 30;   // features used
constexpr int MLP_IN = 32;     // padded input width
constexpr int MLP_HID = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_OUTP = 16;   // padded output width

// f32 master parameters, flattened (the gradient and Adam state share it)
constexpr int OFF_W1 = 0;                         // [32][64]  (input, hidden1)
constexpr int OFF_B1 = OFF_W1 + MLP_IN * MLP_HID;  // [64]
constexpr int OFF_W2 = OFF_B1 + MLP_HID;           // [64][64]  (hidden1, hidden2)
constexpr int OFF_B2 = OFF_W2 + MLP_HID * MLP_HID; // [64]
constexpr int OFF_W3 = OFF_B2 + MLP_HID;           // [64][16]  (hidden2, output)
constexpr int OFF_B3 = OFF_W3 + MLP_HID * MLP_OUTP;// [16]
constexpr int MLP_PARAMS = OFF_B3 + MLP_OUTP;      // 7312
constexpr int MLP_PSTRIDE = 7424;                  // per-wave partial stride (params + loss, padded)

// bf16 copies of the weights in the layouts the MFMA operands read
constexpr int BF_W1T = 0;                          // [64][32]  W1 transposed (forward A)
constexpr int BF_W2T = BF_W1T + MLP_HID * MLP_IN;  // [64][64]  W2 transposed (forward A)
constexpr int BF_W3T = BF_W2T + MLP_HID * MLP_HID; // [16][64]  W3 transposed (forward A)
constexpr int BF_W2N = BF_W3T + MLP_OUTP * MLP_HID;// [64][64]  W2 (backward A)
constexpr int BF_W3N = BF_W2N + MLP_HID * MLP_HID; // [64][16]  W3 (backward A)
constexpr int MLP_BF = BF_W3N + MLP_HID * MLP_OUTP;// 12288 bf16

constexpr uint32_t MLP_TRAIN_BLOCKS = 128;         // 4 waves each
constexpr uint32_t MLP_TRAIN_WAVES = MLP_TRAIN_BLOCKS * 4;

struct WarmFeatArgs {
    const float4* xp;      // predicted positions (w = inverse mass)
    const float4* vel;     // state velocities
    const float4* rest4;
    const int32_t* nbr8;   // 8 grid neighbours per vertex, -1 where absent
    const float4* vaux;    // scaled log10 compliances (stretch, shear, bend), w = 1 / spacing
    float dt;
    uint32_t n;
    uint16_t* feat;        // n x 32 bf16
    uint8_t* tame_out;     // k_warm_infer: per-slice tame bytes of x0 (nullable)
};

// features of the predicted state
void launch_warm_features(const WarmFeatArgs& a, hipStream_t s);
// dataset targets: (converged - predicted) / spacing, f32 x 3 per vertex
void launch_warm_targets(const float4* xp, const float4* xconv, const float4* vaux, uint32_t n, float* t,
                         hipStream_t s);
// forward of `rows` feature rows -> out f32 x 3 per row (model units)
void launch_mlp_forward(const uint16_t* feat, uint32_t rows, const float* P, const uint16_t* bf, float* out,
                        hipStream_t s);
// forward + backward over `rows` rows; per-wave partial gradients (and loss)
void launch_mlp_train(const uint16_t* feat, const float* targets, uint32_t rows, float target_scale,
                      const float* P, const uint16_t* bf, float* partial, hipStream_t s);
// sum the partials; optionally write the gradient and loss; optionally an
// Adam step (lr > 0) that also refreshes the bf16 copies
void launch_mlp_reduce_adam(const float* partial, float* P, float* m, float* v, uint16_t* bf, float* grad_out,
                            float* loss_out, float lr, float beta1_t, float beta2_t, hipStream_t s);
// bf16 operand copies from the f32 master parameters
void launch_mlp_pack(const float* P, uint16_t* bf, hipStream_t s);
// warm start of one substep: features of a.xp (a.feat unused), the forward,
// x0 = xp + spacing * MLP / out_scale (pinned vertices keep xp); count
// (nullable): count[0] = 0, count[1] += 1
void launch_warm_infer(const WarmFeatArgs& a, const float* P, const uint16_t* bf, float inv_out_scale, float4* x0,
                       uint32_t* count, hipStream_t s);

}  // namespace xpbd
