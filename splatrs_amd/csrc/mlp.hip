// Learned densification scorer: 12 features -> 64 -> 64 -> 64 (tanh GELU) -> 3 logits
// (clone, split, prune) on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: bfloat16
// operands, binary32 accumulation).  No counterpart in the reference, whose densification
// is the threshold heuristic of src/densify.rs:125-221.
//
// Orientation: every layer computes D[unit][sample] = W[unit][in] * X[in][sample] + b, so a
// wave holds 16 samples (the 16 columns of the 16x16 tile) and the output units are the
// rows.  MFMA lane layouts (16x16x32): A: lane l holds row l%16, k = 8(l/16)..+7;
// B: column l%16, k = 8(l/16)..+7; D: column l%16, rows 4(l/16)..+3.  A layer's D values
// become the next layer's B operand without leaving the lane: slot (kb, g = l/16, q) of
// the contraction is taken to be unit h(kb, g, q) = 16(2kb + q/4) + 4g + q%4, exactly the
// units the lane holds, and the weight operands are packed in the same permuted order.
//
// Weight gradients contract over samples, so the error signals and activations of a
// 64-sample chunk are staged in LDS transposed ([unit][sample]) and multiplied there; a
// constant-one input row turns the bias gradients into one more output column.
//
// Parameter layout (binary32 master copy): W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[3][64] b4[3] (9347 values).
#include "internal.hpp"
#include "mlp.hpp"

namespace splat {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 12, kHid = 64, kOut = 3;
constexpr int kW1 = 0, kB1 = 768, kW2 = 832, kB2 = 4928, kW3 = 4992, kB3 = 9088, kW4 = 9152,
              kB4 = 9344;
// Packed operand blocks (64 lanes x 8 bfloat16 each): forward A operands of layers 1-4
// (4x1, 4x2, 4x2, 1x2 blocks of M x K), then backward A operands (transposed weights) of
// layers 4, 3, 2 (4x1, 4x2, 4x2).
constexpr int kF1 = 0, kF2 = 4, kF3 = 12, kF4 = 20, kT4 = 22, kT3 = 26, kT2 = 34,
              kBlocks = 42;
constexpr int kChunk = 64;  // samples per workgroup pass of the training kernel

__device__ __forceinline__ __bf16 to_bf16(float x) {  // round to nearest, ties to even
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// gelu(z) = 0.5 z (1 + tanh(u)), u = c (z + 0.044715 z^3); also its derivative.
__device__ __forceinline__ float gelu(float z, float* dg) {
  const float c = 0.7978845608028654f, k = 0.044715f;
  const float u = c * (z + k * z * z * z);
  const float t = 1.0f - 2.0f / (1.0f + splat_expf(2.0f * u));
  const float du = c * (1.0f + 3.0f * k * z * z);
  *dg = 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
  return 0.5f * z * (1.0f + t);
}

__device__ __forceinline__ bf16x8 blockop(const bf16x8* __restrict__ packed, int block, int lane) {
  return packed[block * 64 + lane];
}

// B operand (slot kb) from the D values of four M-blocks.
__device__ __forceinline__ bf16x8 b_from_d(const f32x4* v, int kb) {
  bf16x8 r;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    r[q] = to_bf16(v[2 * kb][q]);
    r[q + 4] = to_bf16(v[2 * kb + 1][q]);
  }
  return r;
}

// Bias in D layout for M-block mb: rows 16 mb + 4g + r.
__device__ __forceinline__ f32x4 bias_d(const float* __restrict__ b, int nout, int mb, int g) {
  f32x4 c;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int o = 16 * mb + 4 * g + r;
    c[r] = o < nout ? b[o] : 0.0f;
  }
  return c;
}

// Feature operand of one lane: features 8g..8g+7 of its sample (zero beyond 12).
__device__ __forceinline__ bf16x8 feature_operand(const float* f, int g) {
  bf16x8 x;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = 8 * g + q;
    x[q] = to_bf16(k < kIn ? f[k] : 0.0f);
  }
  return x;
}

struct Fwd {
  f32x4 z4;                   // logits (D layout, rows 0..2 in lanes 0..15)
  f32x4 d1[4], d2[4], d3[4];  // GELU derivatives
  f32x4 a1[4], a2[4], a3[4];  // activations (binary32, rounded when used)
};

__device__ __forceinline__ void forward16(const bf16x8* __restrict__ packed,
                                          const float* __restrict__ p, bf16x8 x, int lane,
                                          Fwd& F) {
  const int g = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    const f32x4 z = mfma(blockop(packed, kF1 + mb, lane), x, bias_d(p + kB1, kHid, mb, g));
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float d;
      F.a1[mb][r] = gelu(z[r], &d);
      F.d1[mb][r] = d;
    }
  }
  bf16x8 b0 = b_from_d(F.a1, 0), b1 = b_from_d(F.a1, 1);
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    f32x4 z = bias_d(p + kB2, kHid, mb, g);
    z = mfma(blockop(packed, kF2 + 2 * mb, lane), b0, z);
    z = mfma(blockop(packed, kF2 + 2 * mb + 1, lane), b1, z);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float d;
      F.a2[mb][r] = gelu(z[r], &d);
      F.d2[mb][r] = d;
    }
  }
  b0 = b_from_d(F.a2, 0);
  b1 = b_from_d(F.a2, 1);
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    f32x4 z = bias_d(p + kB3, kHid, mb, g);
    z = mfma(blockop(packed, kF3 + 2 * mb, lane), b0, z);
    z = mfma(blockop(packed, kF3 + 2 * mb + 1, lane), b1, z);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float d;
      F.a3[mb][r] = gelu(z[r], &d);
      F.d3[mb][r] = d;
    }
  }
  b0 = b_from_d(F.a3, 0);
  b1 = b_from_d(F.a3, 1);
  f32x4 acc = bias_d(p + kB4, kOut, 0, g);
  acc = mfma(blockop(packed, kF4, lane), b0, acc);
  F.z4 = mfma(blockop(packed, kF4 + 1, lane), b1, acc);
}

// Packs the binary32 master weights into the bfloat16 operand blocks.
__global__ void __launch_bounds__(256) mlp_pack_kernel(const float* __restrict__ p,
                                                       bf16x8* __restrict__ packed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one lane of one block
  if (t >= kBlocks * 64) return;
  const int block = t / 64, lane = t % 64, g = lane >> 4, row16 = lane & 15;
  bf16x8 out;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    float v = 0.0f;
    const int hperm = 4 * g + (q & 3);  // + 16 (2 kb + q / 4)
    if (block < kF2) {  // layer 1 forward: K = features in natural order
      const int o = 16 * block + row16, k = 8 * g + q;
      v = k < kIn ? p[kW1 + o * kIn + k] : 0.0f;
    } else if (block < kT4) {  // layers 2-4 forward
      int L, mb, kb;
      if (block < kF3) {
        L = 2, mb = (block - kF2) / 2, kb = (block - kF2) % 2;
      } else if (block < kF4) {
        L = 3, mb = (block - kF3) / 2, kb = (block - kF3) % 2;
      } else {
        L = 4, mb = 0, kb = block - kF4;
      }
      const int o = 16 * mb + row16, h = 16 * (2 * kb + q / 4) + hperm;
      const int nout = L == 4 ? kOut : kHid;
      const int base = L == 2 ? kW2 : (L == 3 ? kW3 : kW4);
      v = o < nout ? p[base + o * kHid + h] : 0.0f;
    } else {  // backward: transposed weights, rows = input units of layer L
      int L, mb, kb;
      if (block < kT3) {
        L = 4, mb = block - kT4, kb = 0;
      } else if (block < kT2) {
        L = 3, mb = (block - kT3) / 2, kb = (block - kT3) % 2;
      } else {
        L = 2, mb = (block - kT2) / 2, kb = (block - kT2) % 2;
      }
      const int i = 16 * mb + row16, h = 16 * (2 * kb + q / 4) + hperm;
      const int nout = L == 4 ? kOut : kHid;
      const int base = L == 2 ? kW2 : (L == 3 ? kW3 : kW4);
      v = h < nout ? p[base + h * kHid + i] : 0.0f;
    }
    out[q] = to_bf16(v);
  }
  packed[t] = out;
}

// Logits of n feature rows ([n][12] -> [n][3]); one wave per 16 samples, grid-stride.
__global__ void __launch_bounds__(256) mlp_forward_kernel(const bf16x8* __restrict__ packed,
                                                          const float* __restrict__ p,
                                                          const float* __restrict__ feats,
                                                          uint32_t n, float* __restrict__ logits) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  for (uint32_t base = wave * 16; base < n; base += waves * 16) {
    const uint32_t s = base + (lane & 15);
    float f[kIn];
#pragma unroll
    for (int k = 0; k < kIn; k++) f[k] = s < n ? feats[(size_t)s * kIn + k] : 0.0f;
    Fwd F;
    forward16(packed, p, feature_operand(f, g), lane, F);
    if (g == 0 && s < n) {
#pragma unroll
      for (int o = 0; o < kOut; o++) logits[(size_t)s * kOut + o] = F.z4[o];
    }
  }
}

__device__ __forceinline__ float softplus(float z) {
  return z > 0.0f ? z + log1pf(splat_expf(-z)) : log1pf(splat_expf(z));
}

// One 16x16 weight-gradient tile over the 64 staged samples: rows = output units
// 16 mb.., columns = input units 16 nb.. (LDS rows of 64 bfloat16 samples).
__device__ __forceinline__ f32x4 wgrad_tile(const uint16_t* __restrict__ sE,
                                            const uint16_t* __restrict__ sA, int mb, int nb,
                                            int lane, f32x4 acc) {
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int kb = 0; kb < 2; kb++) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(sE + (16 * mb + c) * kChunk + 32 * kb + 8 * g);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(sA + (16 * nb + c) * kChunk + 32 * kb + 8 * g);
    acc = mfma(a, b, acc);
  }
  return acc;
}

// Weight-gradient tiles: layer 4 (1 x 5), layer 3 (4 x 5), layer 2 (4 x 5), layer 1 (4 x 1);
// tile t belongs to wave t % 4.  Column block 4 of layers 2-4 (and column 12 of layer 1)
// is the constant-one input, i.e. the bias gradient.
constexpr int kTiles = 49;
__device__ __forceinline__ void tile_info(int t, int* L, int* mb, int* nb) {
  if (t < 5) {
    *L = 4, *mb = 0, *nb = t;
  } else if (t < 25) {
    *L = 3, *mb = (t - 5) / 5, *nb = (t - 5) % 5;
  } else if (t < 45) {
    *L = 2, *mb = (t - 25) / 5, *nb = (t - 25) % 5;
  } else {
    *L = 1, *mb = t - 45, *nb = 0;
  }
}

// Stage D-layout values (4 M-blocks, or 1 for layer 4) of one wave's 16 samples into an
// LDS [unit][sample] array.
__device__ __forceinline__ void stage(uint16_t* s, const f32x4* v, int mblocks, int lane,
                                      int wave) {
  const int g = lane >> 4, col = 16 * wave + (lane & 15);
  for (int mb = 0; mb < mblocks; mb++)
#pragma unroll
    for (int r = 0; r < 4; r++) s[(16 * mb + 4 * g + r) * kChunk + col] = bf16_bits(v[mb][r]);
}

// Gradient of the weighted binary cross-entropy over `count` samples (ids order[0..count)
// or 0..count), scaled by inv_batch.  Each workgroup accumulates its chunks' weight
// gradients in registers and writes one partial gradient row (kMlpParams floats) plus its
// partial loss; mlp_adam_kernel sums the rows in workgroup order.
__global__ void __launch_bounds__(256) mlp_grad_kernel(
    const bf16x8* __restrict__ packed, const float* __restrict__ p,
    const float* __restrict__ feats, const uint8_t* __restrict__ labels,
    const float* __restrict__ weights, const uint32_t* __restrict__ order, uint32_t count,
    float inv_batch, float* __restrict__ partial, float* __restrict__ partial_loss) {
  __shared__ __attribute__((aligned(16))) uint16_t sE[64 * kChunk];
  __shared__ __attribute__((aligned(16))) uint16_t sA[80 * kChunk];
  __shared__ float sloss[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
  // constant rows: sA[64] = 1 (bias column of layers 2-4), sA[65..79] = 0
  for (int k = threadIdx.x; k < 16 * kChunk; k += blockDim.x)
    sA[64 * kChunk + k] = k < kChunk ? (uint16_t)0x3F80 : (uint16_t)0;
  f32x4 acc[13];
#pragma unroll
  for (int k = 0; k < 13; k++) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float lsum = 0.0f;
  const uint32_t chunks = (count + kChunk - 1) / kChunk;
  for (uint32_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint32_t idx = ch * kChunk + 16 * wave + (lane & 15);
    const bool valid = idx < count;
    const uint32_t s = valid ? (order ? order[idx] : idx) : 0;
    float f[kIn];
#pragma unroll
    for (int k = 0; k < kIn; k++) f[k] = valid ? feats[(size_t)s * kIn + k] : 0.0f;
    const bf16x8 x = feature_operand(f, g);
    Fwd F;
    forward16(packed, p, x, lane, F);
    // output error (rows 0..2 live in lanes g == 0)
    f32x4 e4 = {0.0f, 0.0f, 0.0f, 0.0f};
    const float w = valid ? weights[s] : 0.0f;
    const uint32_t y = valid ? labels[s] : 0u;
    if (g == 0 && valid) {
#pragma unroll
      for (int o = 0; o < kOut; o++) {
        const float z = F.z4[o];
        const float yo = (y >> o) & 1u ? 1.0f : 0.0f;
        lsum += w * (softplus(z) - yo * z) * inv_batch;
        const float sg = 1.0f / (1.0f + splat_expf(-z));
        e4[o] = w * (sg - yo) * inv_batch;
      }
    }
    // round the error signal as the matrix-core operand will see it
#pragma unroll
    for (int r = 0; r < 4; r++) e4[r] = __builtin_bit_cast(float, (uint32_t)bf16_bits(e4[r]) << 16);
    bf16x8 E;
#pragma unroll
    for (int q = 0; q < 8; q++) E[q] = to_bf16(q < 4 ? e4[q] : 0.0f);
    f32x4 e3[4], e2[4], e1[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      const f32x4 t = mfma(blockop(packed, kT4 + mb, lane), E, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int r = 0; r < 4; r++) e3[mb][r] = t[r] * F.d3[mb][r];
    }
    bf16x8 E0 = b_from_d(e3, 0), E1 = b_from_d(e3, 1);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 t = mfma(blockop(packed, kT3 + 2 * mb, lane), E0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      t = mfma(blockop(packed, kT3 + 2 * mb + 1, lane), E1, t);
#pragma unroll
      for (int r = 0; r < 4; r++) e2[mb][r] = t[r] * F.d2[mb][r];
    }
    E0 = b_from_d(e2, 0);
    E1 = b_from_d(e2, 1);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 t = mfma(blockop(packed, kT2 + 2 * mb, lane), E0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      t = mfma(blockop(packed, kT2 + 2 * mb + 1, lane), E1, t);
#pragma unroll
      for (int r = 0; r < 4; r++) e1[mb][r] = t[r] * F.d1[mb][r];
    }
    // weight gradients, layer by layer through LDS
#pragma unroll
    for (int L = 4; L >= 1; L--) {
      __syncthreads();
      if (L == 4) {
        stage(sE, &e4, 1, lane, wave);
        stage(sA, F.a3, 4, lane, wave);
      } else if (L == 3) {
        stage(sE, e3, 4, lane, wave);
        stage(sA, F.a2, 4, lane, wave);
      } else if (L == 2) {
        stage(sE, e2, 4, lane, wave);
        stage(sA, F.a1, 4, lane, wave);
      } else {
        stage(sE, e1, 4, lane, wave);
        const int col = 16 * wave + (lane & 15);
        if (g < 2) {
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const int k = 8 * g + q;
            const float v = k < kIn ? f[k] : (k == kIn ? 1.0f : 0.0f);
            sA[k * kChunk + col] = bf16_bits(v);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < kTiles; t++) {
        int TL, mb, nb;
        tile_info(t, &TL, &mb, &nb);
        if (TL == L && (t & 3) == wave) acc[t >> 2] = wgrad_tile(sE, sA, mb, nb, lane, acc[t >> 2]);
      }
    }
    __syncthreads();
  }
  // write this workgroup's gradient row
  float* row = partial + (size_t)blockIdx.x * kMlpParams;
  const int c = lane & 15;
#pragma unroll
  for (int t = 0; t < kTiles; t++) {
    if ((t & 3) != wave) continue;
    int L, mb, nb;
    tile_info(t, &L, &mb, &nb);
    const int nin = L == 1 ? kIn : kHid, nout = L == 4 ? kOut : kHid;
    const int wbase = L == 1 ? kW1 : (L == 2 ? kW2 : (L == 3 ? kW3 : kW4));
    const int bbase = L == 1 ? kB1 : (L == 2 ? kB2 : (L == 3 ? kB3 : kB4));
    const int in = 16 * nb + c;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int out = 16 * mb + 4 * g + r;
      if (out >= nout) continue;
      if (in < nin)
        row[wbase + out * nin + in] = acc[t >> 2][r];
      else if (in == nin)
        row[bbase + out] = acc[t >> 2][r];
    }
  }
  for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d, 64);
  if (lane == 0) sloss[wave] = lsum;
  __syncthreads();
  if (threadIdx.x == 0)
    partial_loss[blockIdx.x] = ((sloss[0] + sloss[1]) + sloss[2]) + sloss[3];
}

// Sums the workgroups' gradient rows in order; optionally writes the gradient and/or
// applies Adam to the master weights.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ p,
                                                       float* __restrict__ m,
                                                       float* __restrict__ v,
                                                       const float* __restrict__ partial,
                                                       uint32_t rows, float* __restrict__ grad_out,
                                                       int update, float lr, float c1, float c2) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (uint32_t)kMlpParams) return;
  float gsum = 0.0f;
  for (uint32_t r = 0; r < rows; r++) gsum += partial[(size_t)r * kMlpParams + k];
  if (grad_out) grad_out[k] = gsum;
  if (!update) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m[k] = b1 * m[k] + (1.0f - b1) * gsum;
  v[k] = b2 * v[k] + (1.0f - b2) * gsum * gsum;
  p[k] -= lr * (m[k] / c1) / (sqrtf(v[k] / c2) + eps);
}

__global__ void mlp_loss_sum_kernel(const float* __restrict__ partial_loss, uint32_t rows,
                                    float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (uint32_t r = 0; r < rows; r++) s += partial_loss[r];
  *out = s;
}

}  // namespace

size_t mlp_packed_bytes() { return (size_t)kBlocks * 64 * 16; }

void mlp_pack(splat_ctx* ctx, const float* params, void* packed) {
  launch(ctx, "mlp_pack", mlp_pack_kernel, dim3(blocks_for(kBlocks * 64, 256)), dim3(256), 0,
         params, (bf16x8*)packed);
}

void mlp_forward(splat_ctx* ctx, const void* packed, const float* params, const float* feats,
                 size_t n, float* logits) {
  const unsigned blocks = std::min<unsigned>(blocks_for(n, 64), 4096u);
  launch(ctx, "mlp_forward", mlp_forward_kernel, dim3(blocks), dim3(256), 0,
         (const bf16x8*)packed, params, feats, (uint32_t)n, logits);
}

unsigned mlp_grad_rows(size_t count) {
  return std::max<unsigned>(1u, std::min<unsigned>(blocks_for(count, kChunk), 256u));
}

void mlp_grad(splat_ctx* ctx, const void* packed, const float* params, const float* feats,
              const uint8_t* labels, const float* weights, const uint32_t* order, size_t count,
              float inv_batch, float* partial, float* partial_loss) {
  launch(ctx, "mlp_grad", mlp_grad_kernel, dim3(mlp_grad_rows(count)), dim3(256), 0,
         (const bf16x8*)packed, params, feats, labels, weights, order, (uint32_t)count,
         inv_batch, partial, partial_loss);
}

void mlp_reduce_update(splat_ctx* ctx, float* params, float* m, float* v, const float* partial,
                       unsigned rows, float* grad_out, bool update, float lr, float c1,
                       float c2) {
  launch(ctx, "mlp_adam", mlp_adam_kernel, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0,
         params, m, v, partial, rows, grad_out, update ? 1 : 0, lr, c1, c2);
}

void mlp_loss_sum(splat_ctx* ctx, const float* partial_loss, unsigned rows, float* out) {
  launch(ctx, "mlp_loss_sum", mlp_loss_sum_kernel, dim3(1), dim3(64), 0, partial_loss, rows, out);
}

}  // namespace splat
