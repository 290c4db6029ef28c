/*
 * cs_mlp.h -- the learned closure's network on CDNA4 matrix cores.
 *
 * A 30-64-64-1 perceptron with GELU units (inputs padded to 32), bfloat16
 * weights and activations, FP32 accumulation, evaluated 32 samples per
 * wavefront with v_mfma_f32_32x32x16_bf16.  The reference has no learned
 * part; this is the sub-grid closure of the coarse-grid method.
 *
 * Layouts.  Every product is computed transposed, D[feature][sample] =
 * W^T[feature][k] X^T[k][sample]: a 32x32 block of D gives lane l the
 * sample l%32 and, in its 16 accumulators r, the features
 * F(g, r) = 8 (r/4) + 4 g + r%4 of the block (g = l/32).  The next
 * layer takes its B operand (lane l: sample l%32, eight consecutive k of
 * lane group g) straight from those accumulators, registers 8q..8q+7 for
 * the q-th chunk of 16, so its k index runs over the block's features in
 * the order F(g, 8q + j); the weights are stored pre-permuted to match
 * (MlpDev::opnd), which leaves the sum unchanged.
 */
#ifndef CS_MLP_H
#define CS_MLP_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "csparks.h"

namespace {

constexpr int MLP_IN = 30, MLP_PAD = 32, MLP_H = 64;
constexpr int MLP_NPARAM = MLP_PAD * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;
/* offsets of the packed FP32 parameter vector (W1 [in][out], b1, W2
 * [in][out], b2, w3, b3) */
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_PAD * MLP_H, MLP_OFF_W2 = MLP_OFF_B1 + MLP_H,
              MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H, MLP_OFF_W3 = MLP_OFF_B2 + MLP_H,
              MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

/* operand image segments: layer 1 (block m, chunk c), layer 2 forward
 * (out block m, in block mb, chunk q), layer 2 backward (in block mi, out
 * block mo, chunk q) */
constexpr int SEG_L1 = 0, SEG_L2F = 4, SEG_L2B = 12, NSEG = 20;

struct alignas(16) MlpDev {
  uint16_t opnd[NSEG][64][8];  /* bfloat16 bits of the A operands        */
  float b1[MLP_H], b2[MLP_H];
  float w3[MLP_H];             /* bfloat16-rounded output weights         */
  float b3;
  /* feature normalisation: x = (f - mean) / std, with f the logarithm of
   * a concentration (features 0..27) or a count (28, 29); the output is
   * log(c) = y * out_std + out_mean; lo/hi bound the normalised inputs
   * seen in training (the range monitor) */
  float in_mean[MLP_IN], in_std[MLP_IN], in_lo[MLP_IN], in_hi[MLP_IN];
  float out_mean, out_std, margin;
  float pad[1];
};

__host__ __device__ inline int mlp_feat(int g, int r) { return 8 * (r / 4) + 4 * g + (r % 4); }

__device__ inline bf16x8 mlp_opnd(const MlpDev* __restrict__ P, int seg, int lane)
{
  return *reinterpret_cast<const bf16x8*>(&P->opnd[seg][lane][0]);
}

/* erf by Abramowitz and Stegun, Handbook of Mathematical Functions,
 * formula 7.1.26: erf(x) = 1 - (a1 t + .. + a5 t^5) exp(-x^2) with
 * t = 1 / (1 + p x) for x >= 0, odd in x, |error| <= 1.5e-7 (plus FP32
 * rounding).  Branch-free and about a third of the instructions of the
 * library erff, which the two GELU layers evaluate 64 times per sample. */
__device__ inline float mlp_erf(float x)
{
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
  float p = 1.061405429f;
  p = fmaf(p, t, -1.453152027f);
  p = fmaf(p, t, 1.421413741f);
  p = fmaf(p, t, -0.284496736f);
  p = fmaf(p, t, 0.254829592f);
  const float e = 1.0f - p * t * __expf(-ax * ax);
  return copysignf(e, x);
}

__device__ inline float mlp_gelu(float z) { return 0.5f * z * (1.0f + mlp_erf(z * 0.70710678118654752f)); }

__device__ inline float mlp_gelu_grad(float z)
{
  return 0.5f * (1.0f + mlp_erf(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ inline __bf16 to_bf16(float v) { return (__bf16)v; }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

/* Bits of element j of a bfloat16 vector.  Taken through the vector's
 * 32-bit words: indexing the __bf16 vector itself (v[j] with j odd or
 * above 1) was compiled to a 16-bit store of element 0's half-word. */
__device__ inline uint16_t bf16_bits(const bf16x8& v, int j)
{
  const u32x4 w = __builtin_bit_cast(u32x4, v);
  return (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
}

/* The 16 values of a 64-entry per-feature table (bias, output weight) at
 * the features lane group g holds in block m: four runs of four
 * consecutive features, so four 16-byte loads. */
__device__ inline void feat_vec(const float* __restrict__ tab, int m, int g, float (&v)[16])
{
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const float4 q = *reinterpret_cast<const float4*>(tab + 32 * m + 8 * a + 4 * g);
    v[4 * a] = q.x;
    v[4 * a + 1] = q.y;
    v[4 * a + 2] = q.z;
    v[4 * a + 3] = q.w;
  }
}

/* Forward pass of the wavefront's 32 samples.  x[c] is the B operand of
 * input chunk c (lane l: sample l%32, inputs 16c + 8g .. + 7); opnd(seg)
 * returns the lane's A operand of image segment seg.  Returns y for sample
 * l%32 (in both lane groups) and keeps what the backward pass needs: the
 * pre-activations z1, z2 and the bfloat16 layer-1 outputs h1b;
 * on_h2(m, r, v) receives each bfloat16-rounded layer-2 output. */
template <typename Opnd, typename OnH2>
__device__ inline float mlp_forward(const MlpDev* __restrict__ P, Opnd opnd, const bf16x8 (&x)[2], f32x16 (&z1)[2],
                                    f32x16 (&z2)[2], bf16x8 (&h1b)[2][2], OnH2 on_h2)
{
  const int lane = threadIdx.x & 63, g = lane >> 5;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    float b[16];
    feat_vec(P->b1, m, g, b);
#pragma unroll
    for (int r = 0; r < 16; r++) z1[m][r] = b[r];
#pragma unroll
    for (int c = 0; c < 2; c++)
      z1[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(opnd(SEG_L1 + 2 * m + c), x[c], z1[m], 0, 0, 0);
  }
#pragma unroll
  for (int mb = 0; mb < 2; mb++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int j = 0; j < 8; j++) h1b[mb][q][j] = to_bf16(mlp_gelu(z1[mb][8 * q + j]));
  float part = 0.0f;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    float b[16], w[16];
    feat_vec(P->b2, m, g, b);
#pragma unroll
    for (int r = 0; r < 16; r++) z2[m][r] = b[r];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
      for (int q = 0; q < 2; q++)
        z2[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(opnd(SEG_L2F + 4 * m + 2 * mb + q), h1b[mb][q], z2[m], 0, 0, 0);
    feat_vec(P->w3, m, g, w);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float h2 = (float)to_bf16(mlp_gelu(z2[m][r]));
      on_h2(m, r, h2);
      part += w[r] * h2;
    }
  }
  return part + __shfl_xor(part, 32, 64) + P->b3;
}

}  // namespace

/* Device image (MlpDev) of a network, defined in cs_closure.hip. */
const void* cs_internal_mlp_device(const cs_mlp* m);

#endif /* CS_MLP_H */
