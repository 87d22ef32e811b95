// Learned densification scorer kernels (mlp.hip) and their launchers.
#pragma once

#include "internal.hpp"

namespace splat {

constexpr int kMlpParams = 9347;  // W1[64][12] b1 W2[64][64] b2 W3[64][64] b3 W4[3][64] b4
constexpr int kMlpFeatures = 12;

size_t mlp_packed_bytes();
// Binary32 master weights -> bfloat16 matrix-core operand blocks.
void mlp_pack(splat_ctx* ctx, const float* params, void* packed);
// logits[n][3] of feats[n][12].
void mlp_forward(splat_ctx* ctx, const void* packed, const float* params, const float* feats,
                 size_t n, float* logits);
// Rows of partial gradients (and partial losses) mlp_grad writes for `count` samples.
unsigned mlp_grad_rows(size_t count);
// Weighted BCE gradient over samples order[0..count) (or 0..count when order is null),
// scaled by inv_batch, as mlp_grad_rows(count) partial rows of kMlpParams floats.
void mlp_grad(splat_ctx* ctx, const void* packed, const float* params, const float* feats,
              const uint8_t* labels, const float* weights, const uint32_t* order, size_t count,
              float inv_batch, float* partial, float* partial_loss);
// Sum of the partial rows (in row order) -> grad_out (optional); Adam on the master weights
// when `update`.
void mlp_reduce_update(splat_ctx* ctx, float* params, float* m, float* v, const float* partial,
                       unsigned rows, float* grad_out, bool update, float lr, float c1, float c2);
void mlp_loss_sum(splat_ctx* ctx, const float* partial_loss, unsigned rows, float* out);

}  // namespace splat
