/*
 * cs_closure.hip -- training and inference of the learned closure's
 * network (cs_mlp_*): forward and backward passes on bfloat16 matrix-core
 * products with FP32 accumulation, fixed-order gradient reductions, Adam.
 *
 * The reference has no learned part.  The network corrects the
 * coarse-grid calcium that channel gating reads (evaluate_RyR_
 * stochastically, boundaryfluxes.c:295-359, reads the cleft voxel's
 * concentration); its inputs and labels come from paired fine/coarse
 * runs of the same dyads, and its use inside an ensemble's gating step
 * lives in cs_batch.hip.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <vector>

#include "csparks.h"
#include "cs_device.h"
#include "cs_mlp.h"

namespace {

constexpr int LROW = 40;                   /* LDS row of 32 samples, padded */

/* FP32 master parameters -> the device image: pre-permuted bfloat16
 * operands of both layers, biases, bfloat16-rounded output weights.  One
 * block; thread k fills element k of the image. */
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ p, MlpDev* __restrict__ P)
{
  for (int k = threadIdx.x; k < NSEG * 64 * 8; k += blockDim.x) {
    const int seg = k / 512, lane = (k / 8) % 64, j = k % 8, g = lane / 32, l = lane % 32;
    float w;
    if (seg < SEG_L2F) {
      const int m = (seg - SEG_L1) / 2, c = (seg - SEG_L1) % 2;
      w = p[MLP_OFF_W1 + (16 * c + 8 * g + j) * MLP_H + 32 * m + l];
    } else if (seg < SEG_L2B) {
      const int s = seg - SEG_L2F, m = s / 4, mb = (s / 2) % 2, q = s % 2;
      w = p[MLP_OFF_W2 + (32 * mb + mlp_feat(g, 8 * q + j)) * MLP_H + 32 * m + l];
    } else {
      const int s = seg - SEG_L2B, mi = s / 4, mo = (s / 2) % 2, q = s % 2;
      w = p[MLP_OFF_W2 + (32 * mi + l) * MLP_H + 32 * mo + mlp_feat(g, 8 * q + j)];
    }
    P->opnd[seg][lane][j] = __builtin_bit_cast(uint16_t, to_bf16(w));
  }
  for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) {
    P->b1[k] = p[MLP_OFF_B1 + k];
    P->b2[k] = p[MLP_OFF_B2 + k];
    P->w3[k] = (float)to_bf16(p[MLP_OFF_W3 + k]);
  }
  if (threadIdx.x == 0) P->b3 = p[MLP_OFF_B3];
}

/* Normalised FP32 inputs [n][30] -> bfloat16 rows [n][32] (zero padded). */
__global__ void k_mlp_to_bf16(const float* __restrict__ x, uint64_t n, uint16_t* __restrict__ xb)
{
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n * MLP_PAD;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = k / MLP_PAD;
    const int f = (int)(k % MLP_PAD);
    xb[k] = __builtin_bit_cast(uint16_t, to_bf16(f < MLP_IN ? x[s * MLP_IN + f] : 0.0f));
  }
}

/* B operands of input chunk c for sample s of a bfloat16 row array. */
__device__ inline void load_x(const uint16_t* __restrict__ xb, uint64_t s, bool valid, bf16x8 (&x)[2])
{
  const int g = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    if (valid) x[c] = *reinterpret_cast<const bf16x8*>(xb + s * MLP_PAD + 16 * c + 8 * g);
    else
#pragma unroll
      for (int j = 0; j < 8; j++) x[c][j] = to_bf16(0.0f);
  }
}

/* Inference: y[s] for n samples of bfloat16 rows, 32 per wavefront. */
__global__ void __launch_bounds__(256)
k_mlp_forward(const MlpDev* __restrict__ P, const uint16_t* __restrict__ xb, uint64_t n, float* __restrict__ y)
{
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t w0 = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  for (uint64_t t = w0; t * 32 < n; t += waves) {
    /* reload the weights every round instead of holding them all in
     * registers across the loop */
    asm volatile("" ::: "memory");
    const uint64_t s = t * 32 + (lane & 31);
    bf16x8 x[2];
    load_x(xb, s, s < n, x);
    f32x16 z1[2], z2[2];
    bf16x8 h1b[2][2];
    const float out = mlp_forward(P, [&](int seg) { return mlp_opnd(P, seg, lane); }, x, z1, z2, h1b,
                                  [](int, int, float) {});
    if (lane < 32 && s < n) y[s] = out;
  }
}

/* Forward and backward passes over n samples (rows idx[k] of xb when idx
 * is given, else rows k), one wavefront per block, writing each
 * wavefront's gradient of L = sum (y - t)^2 / norm into
 * part[wave][MLP_NPARAM] and its sum of squared errors into sse[wave].
 * The weight operands sit in LDS; the weight gradients are products over
 * the 32 samples of a round, formed on the matrix cores from activations
 * and back-propagated gradients transposed through LDS. */
__global__ void __launch_bounds__(64)
k_mlp_grad(const MlpDev* __restrict__ P, const uint16_t* __restrict__ xb, const float* __restrict__ tv,
           const uint32_t* __restrict__ idx, uint64_t n, float norm, float* __restrict__ part,
           double* __restrict__ sse)
{
  __shared__ alignas(16) uint16_t Lop[NSEG][64][8];
  /* x, h1, h2, g_z2, g_z1 as [feature][sample] bfloat16; g_y per sample */
  __shared__ alignas(16) uint16_t Lx[MLP_PAD][LROW];
  __shared__ alignas(16) uint16_t Lh1[MLP_H][LROW];
  __shared__ alignas(16) uint16_t Lh2[MLP_H][LROW];
  __shared__ alignas(16) uint16_t Lz2[MLP_H][LROW];
  __shared__ alignas(16) uint16_t Lz1[MLP_H][LROW];
  __shared__ float Lgy[32];
  const int lane = threadIdx.x & 63, g = lane >> 5, l = lane & 31;
  {
    const uint4* src = reinterpret_cast<const uint4*>(&P->opnd[0][0][0]);
    uint4* dst = reinterpret_cast<uint4*>(&Lop[0][0][0]);
    for (int k = lane; k < NSEG * 64; k += 64) dst[k] = src[k];
  }
  auto lop = [&](int seg) { return *reinterpret_cast<const bf16x8*>(&Lop[seg][lane][0]); };
  const uint64_t waves = gridDim.x;
  const uint64_t wid = blockIdx.x;
  const uint64_t tiles = (n + 31) / 32;

  f32x16 gw1[2], gw2[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++) {
    gw1[a] = f32x16{};
    gw2[a][0] = f32x16{};
    gw2[a][1] = f32x16{};
  }
  float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f;
  double err2 = 0.0;
  __syncthreads();

  for (uint64_t t = wid; t < tiles; t += waves) {
    /* reload the weights every round instead of holding them all in
     * registers across the loop */
    asm volatile("" ::: "memory");
    const uint64_t k = t * 32 + l;
    const bool valid = k < n;
    const uint64_t s = valid ? (idx ? (uint64_t)idx[k] : k) : 0;
    bf16x8 x[2];
    load_x(xb, s, valid, x);
    f32x16 z1[2], z2[2];
    bf16x8 h1b[2][2];
    const float y = mlp_forward(P, lop, x, z1, z2, h1b, [&](int m, int r, float h2) {
      Lh2[32 * m + mlp_feat(g, r)][l] = __builtin_bit_cast(uint16_t, to_bf16(h2));
    });
    const float gy = valid ? 2.0f * (y - tv[s]) / norm : 0.0f;
    if (g == 0) {
      Lgy[l] = gy;
      gb3 += gy;
      if (valid) err2 += (double)(y - tv[s]) * (double)(y - tv[s]);
    }
    /* g_z2 = w3 g_y gelu'(z2), bfloat16 */
    bf16x8 gz2b[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
      float w[16];
      feat_vec(P->w3, m, g, w);
#pragma unroll
      for (int r = 0; r < 16; r++) gz2b[m][r / 8][r % 8] = to_bf16(w[r] * gy * mlp_gelu_grad(z2[m][r]));
    }
    /* transposes for the weight gradients */
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int j = 0; j < 8; j++) Lx[16 * c + 8 * g + j][l] = bf16_bits(x[c], j);
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int f = 32 * m + mlp_feat(g, r);
        Lh1[f][l] = bf16_bits(h1b[m][r / 8], r % 8);
        Lz2[f][l] = bf16_bits(gz2b[m][r / 8], r % 8);
      }
    /* g_h1 = W2 g_z2 on the matrix cores, then g_z1 = g_h1 gelu'(z1) */
#pragma unroll
    for (int mi = 0; mi < 2; mi++) {
      f32x16 gh1 = f32x16{};
#pragma unroll
      for (int mo = 0; mo < 2; mo++)
#pragma unroll
        for (int q = 0; q < 2; q++)
          gh1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lop(SEG_L2B + 4 * mi + 2 * mo + q), gz2b[mo][q], gh1, 0, 0, 0);
      /* z1 of this block again (two products), rather than holding it
       * in registers through the layer-2 backward pass */
      f32x16 z = f32x16{};
      {
        float b[16];
        feat_vec(P->b1, mi, g, b);
#pragma unroll
        for (int r = 0; r < 16; r++) z[r] = b[r];
      }
#pragma unroll
      for (int c = 0; c < 2; c++)
        z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lop(SEG_L1 + 2 * mi + c), x[c], z, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; r++)
        Lz1[32 * mi + mlp_feat(g, r)][l] = __builtin_bit_cast(uint16_t, to_bf16(gh1[r] * mlp_gelu_grad(z[r])));
    }
    __syncthreads();
    /* dW^T[o][i] += sum over the round's samples of g_z[o][s] a[i][s] */
    auto row = [&](const uint16_t (*L)[LROW], int r, int c) {
      return *reinterpret_cast<const bf16x8*>(&L[r][16 * c + 8 * g]);
    };
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int mo = 0; mo < 2; mo++) {
        const bf16x8 az2 = row(Lz2, 32 * mo + l, c);
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
          gw2[mo][mi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az2, row(Lh1, 32 * mi + l, c), gw2[mo][mi], 0, 0, 0);
        gw1[mo] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(row(Lz1, 32 * mo + l, c), row(Lx, l, c), gw1[mo], 0, 0, 0);
      }
    /* lane f: bias gradients and the output weight's gradient of feature f */
    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    for (int j = 0; j < 32; j++) {
      s1 += (float)__builtin_bit_cast(__bf16, Lz1[lane][j]);
      s2 += (float)__builtin_bit_cast(__bf16, Lz2[lane][j]);
      s3 += (float)__builtin_bit_cast(__bf16, Lh2[lane][j]) * Lgy[j];
    }
    gb1 += s1;
    gb2 += s2;
    gw3 += s3;
    __syncthreads();   /* reads done before the next round writes */
  }

  /* this wavefront's gradient, in the packed parameter order */
  float* o = part + wid * (uint64_t)MLP_NPARAM;
#pragma unroll
  for (int mo = 0; mo < 2; mo++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int of = 32 * mo + mlp_feat(g, r);
      o[MLP_OFF_W1 + l * MLP_H + of] = gw1[mo][r];
#pragma unroll
      for (int mi = 0; mi < 2; mi++) o[MLP_OFF_W2 + (32 * mi + l) * MLP_H + of] = gw2[mo][mi][r];
    }
  o[MLP_OFF_B1 + lane] = gb1;
  o[MLP_OFF_B2 + lane] = gb2;
  o[MLP_OFF_W3 + lane] = gw3;
  for (int off = 16; off > 0; off >>= 1) {
    gb3 += __shfl_xor(gb3, off, 64);
    err2 += __shfl_xor(err2, off, 64);
  }
  if (lane == 0) {
    o[MLP_OFF_B3] = gb3;
    sse[wid] = err2;
  }
}

/* grad[k] = sum over the wavefront partials, in wavefront order. */
__global__ void k_mlp_reduce(const float* __restrict__ part, uint32_t nwaves, float* __restrict__ grad,
                             const double* __restrict__ sse, double* __restrict__ sse_out)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < MLP_NPARAM) {
    double acc = 0.0;
    for (uint32_t w = 0; w < nwaves; w++) acc += part[(uint64_t)w * MLP_NPARAM + k];
    grad[k] = (float)acc;
  }
  if (k == 0) {
    double e = 0.0;
    for (uint32_t w = 0; w < nwaves; w++) e += sse[w];
    *sse_out = e;
  }
}

/* One Adam step (Kingma & Ba 2015) on the FP32 master parameters. */
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= MLP_NPARAM) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

}  // namespace

struct cs_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  MlpDev* d_dev = nullptr;
  float* d_p = nullptr;       /* FP32 master parameters */
  float* d_m = nullptr;       /* Adam moments */
  float* d_v = nullptr;
  float* d_grad = nullptr;
  float* d_part = nullptr;
  double* d_sse = nullptr;
  double* d_sse_out = nullptr;
  uint32_t grad_waves = 0;    /* wavefronts of a gradient launch */
  uint64_t adam_step = 0;
  /* training set */
  uint16_t* d_x = nullptr;
  float* d_t = nullptr;
  uint32_t* d_idx = nullptr;
  uint64_t ndata = 0, cursor = 0;
  std::vector<uint32_t> order;
  uint64_t shuffle_state = 0;
  /* scratch of cs_mlp_forward / cs_mlp_grad */
  float* d_xs = nullptr;
  uint16_t* d_xbs = nullptr;
  float* d_ts = nullptr;
  float* d_ys = nullptr;
  uint64_t scratch_n = 0;
  MlpDev host_norm{};         /* normalisation fields (host copy) */
};

namespace {

void free_mlp(cs_mlp* m)
{
  if (!m) return;
  (void)hipSetDevice(m->device);
  void* ptrs[] = {m->d_dev, m->d_p, m->d_m, m->d_v, m->d_grad, m->d_part, m->d_sse, m->d_sse_out, m->d_x, m->d_t,
                  m->d_idx, m->d_xs, m->d_xbs, m->d_ts, m->d_ys};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int pack(cs_mlp* m)
{
  hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, m->stream, m->d_p, m->d_dev);
  HIP_TRY(hipGetLastError());
  return CS_OK;
}

int ensure_scratch(cs_mlp* m, uint64_t n)
{
  if (n <= m->scratch_n && m->d_xs) return CS_OK;
  void** ptrs[] = {(void**)&m->d_xs, (void**)&m->d_xbs, (void**)&m->d_ts, (void**)&m->d_ys};
  for (void** p : ptrs)
    if (*p) { HIP_TRY(hipFree(*p)); *p = nullptr; }
  HIP_TRY(hipMalloc(&m->d_xs, sizeof(float) * n * MLP_IN));
  HIP_TRY(hipMalloc(&m->d_xbs, sizeof(uint16_t) * n * MLP_PAD));
  HIP_TRY(hipMalloc(&m->d_ts, sizeof(float) * n));
  HIP_TRY(hipMalloc(&m->d_ys, sizeof(float) * n));
  m->scratch_n = n;
  return CS_OK;
}

int to_bf16_rows(cs_mlp* m, const float* d_x, uint64_t n, uint16_t* d_xb)
{
  const unsigned blocks = (unsigned)std::min<uint64_t>(4096, (n * MLP_PAD + 255) / 256 + 1);
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3(blocks), dim3(256), 0, m->stream, d_x, n, d_xb);
  HIP_TRY(hipGetLastError());
  return CS_OK;
}

/* Gradient of sum (y - t)^2 / norm over n samples into d_grad and the sum
 * of squared errors into d_sse_out. */
int run_grad(cs_mlp* m, const uint16_t* d_xb, const float* d_t, const uint32_t* d_idx, uint64_t n, float norm)
{
  hipLaunchKernelGGL(k_mlp_grad, dim3(m->grad_waves), dim3(64), 0, m->stream, m->d_dev,
                     d_xb, d_t, d_idx, n, norm, m->d_part, m->d_sse);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->d_part,
                     m->grad_waves, m->d_grad, m->d_sse, m->d_sse_out);
  HIP_TRY(hipGetLastError());
  return CS_OK;
}

}  // namespace

extern "C" {

int cs_mlp_create(const float* params, cs_mlp** out)
{
  if (!out || !params) return set_err(CS_EINVAL, "NULL argument");
  *out = nullptr;
  cs_mlp* m = new cs_mlp();
  auto bail = [&](int code) { free_mlp(m); return code; };
  if (hipGetDevice(&m->device) != hipSuccess) return bail(set_err(CS_EDEVICE, "no HIP device"));
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "hipStreamCreate failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, m->device) != hipSuccess) return bail(set_err(CS_EDEVICE, "device query failed"));
  m->grad_waves = (uint32_t)std::max(1, prop.multiProcessorCount) * 3;   /* three blocks per CU */
  if (hipMalloc(&m->d_dev, sizeof(MlpDev)) != hipSuccess || hipMalloc(&m->d_p, sizeof(float) * MLP_NPARAM) != hipSuccess ||
      hipMalloc(&m->d_m, sizeof(float) * MLP_NPARAM) != hipSuccess ||
      hipMalloc(&m->d_v, sizeof(float) * MLP_NPARAM) != hipSuccess ||
      hipMalloc(&m->d_grad, sizeof(float) * MLP_NPARAM) != hipSuccess ||
      hipMalloc(&m->d_part, sizeof(float) * MLP_NPARAM * (size_t)m->grad_waves) != hipSuccess ||
      hipMalloc(&m->d_sse, sizeof(double) * m->grad_waves) != hipSuccess ||
      hipMalloc(&m->d_sse_out, sizeof(double)) != hipSuccess)
    return bail(set_err(CS_ENOMEM, "device allocation of the network failed"));
  /* identity normalisation until cs_mlp_set_normalisation */
  MlpDev& h = m->host_norm;
  for (int k = 0; k < MLP_IN; k++) {
    h.in_mean[k] = 0.0f;
    h.in_std[k] = 1.0f;
    h.in_lo[k] = -INFINITY;
    h.in_hi[k] = INFINITY;
  }
  h.out_mean = 0.0f;
  h.out_std = 1.0f;
  h.margin = 0.0f;
  if (hipMemcpyAsync(m->d_dev, &h, sizeof(MlpDev), hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(m->d_p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemsetAsync(m->d_m, 0, sizeof(float) * MLP_NPARAM, m->stream) != hipSuccess ||
      hipMemsetAsync(m->d_v, 0, sizeof(float) * MLP_NPARAM, m->stream) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "upload of the network failed"));
  int rc = pack(m);
  if (rc != CS_OK) return bail(rc);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return bail(set_err(CS_EDEVICE, "network setup failed"));
  *out = m;
  return CS_OK;
}

int cs_mlp_destroy(cs_mlp* m)
{
  if (!m) return CS_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_mlp(m);
  return CS_OK;
}

int cs_mlp_num_params(void) { return MLP_NPARAM; }

int cs_mlp_get_params(cs_mlp* m, float* params)
{
  if (!m || !params) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpyAsync(params, m->d_p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return CS_OK;
}

int cs_mlp_set_params(cs_mlp* m, const float* params)
{
  if (!m || !params) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpyAsync(m->d_p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
  int rc = pack(m);
  if (rc != CS_OK) return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return CS_OK;
}

int cs_mlp_set_normalisation(cs_mlp* m, const float* in_mean, const float* in_std, const float* in_lo,
                             const float* in_hi, float out_mean, float out_std, float margin)
{
  if (!m || !in_mean || !in_std || !in_lo || !in_hi) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(m->device));
  MlpDev& h = m->host_norm;
  for (int k = 0; k < MLP_IN; k++) {
    if (!(in_std[k] > 0)) return set_err(CS_EINVAL, "in_std[%d] must be positive", k);
    h.in_mean[k] = in_mean[k];
    h.in_std[k] = in_std[k];
    h.in_lo[k] = in_lo[k];
    h.in_hi[k] = in_hi[k];
  }
  if (!(out_std > 0)) return set_err(CS_EINVAL, "out_std must be positive");
  h.out_mean = out_mean;
  h.out_std = out_std;
  h.margin = margin;
  const size_t off = offsetof(MlpDev, in_mean);
  HIP_TRY(hipMemcpyAsync((char*)m->d_dev + off, (const char*)&h + off, sizeof(MlpDev) - off, hipMemcpyHostToDevice,
                         m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return CS_OK;
}

int cs_mlp_forward(cs_mlp* m, const float* x, uint64_t n, float* y)
{
  if (!m || (!x && n) || (!y && n)) return set_err(CS_EINVAL, "NULL argument");
  if (n == 0) return CS_OK;
  HIP_TRY(hipSetDevice(m->device));
  int rc = ensure_scratch(m, n);
  if (rc != CS_OK) return rc;
  HIP_TRY(hipMemcpyAsync(m->d_xs, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, m->stream));
  if ((rc = to_bf16_rows(m, m->d_xs, n, m->d_xbs)) != CS_OK) return rc;
  const unsigned blocks = (unsigned)std::min<uint64_t>(2048, (n + 127) / 128);
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, m->stream, m->d_dev, m->d_xbs, n, m->d_ys);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(y, m->d_ys, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return CS_OK;
}

int cs_mlp_grad(cs_mlp* m, const float* x, const float* t, uint64_t n, float* grad, double* loss)
{
  if (!m || !x || !t || !grad) return set_err(CS_EINVAL, "NULL argument");
  if (n == 0) return set_err(CS_EINVAL, "no samples");
  HIP_TRY(hipSetDevice(m->device));
  int rc = ensure_scratch(m, n);
  if (rc != CS_OK) return rc;
  HIP_TRY(hipMemcpyAsync(m->d_xs, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->d_ts, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  if ((rc = to_bf16_rows(m, m->d_xs, n, m->d_xbs)) != CS_OK) return rc;
  if ((rc = run_grad(m, m->d_xbs, m->d_ts, nullptr, n, (float)n)) != CS_OK) return rc;
  double sse = 0.0;
  HIP_TRY(hipMemcpyAsync(grad, m->d_grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipMemcpyAsync(&sse, m->d_sse_out, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (loss) *loss = sse / (double)n;
  return CS_OK;
}

int cs_mlp_set_data(cs_mlp* m, const float* x, const float* t, uint64_t n, uint64_t shuffle_seed)
{
  if (!m || !x || !t) return set_err(CS_EINVAL, "NULL argument");
  if (n == 0 || n > 0xFFFFFFFFull) return set_err(CS_ERANGE, "training set size out of range");
  HIP_TRY(hipSetDevice(m->device));
  void** ptrs[] = {(void**)&m->d_x, (void**)&m->d_t, (void**)&m->d_idx};
  for (void** p : ptrs)
    if (*p) { HIP_TRY(hipFree(*p)); *p = nullptr; }
  float* d_xf = nullptr;
  HIP_TRY(hipMalloc(&d_xf, sizeof(float) * n * MLP_IN));
  HIP_TRY(hipMalloc(&m->d_x, sizeof(uint16_t) * n * MLP_PAD));
  HIP_TRY(hipMalloc(&m->d_t, sizeof(float) * n));
  HIP_TRY(hipMalloc(&m->d_idx, sizeof(uint32_t) * n));
  HIP_TRY(hipMemcpyAsync(d_xf, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->d_t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  int rc = to_bf16_rows(m, d_xf, n, m->d_x);
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipFree(d_xf));
  if (rc != CS_OK) return rc;
  m->ndata = n;
  m->cursor = n;   /* shuffle before the first batch */
  m->order.resize(n);
  for (uint64_t k = 0; k < n; k++) m->order[k] = (uint32_t)k;
  m->shuffle_state = shuffle_seed * 0x9E3779B97F4A7C15ull + 1;
  return CS_OK;
}

int cs_mlp_train(cs_mlp* m, uint32_t steps, uint32_t batch, float lr, double* last_loss)
{
  if (!m) return set_err(CS_EINVAL, "NULL argument");
  if (!m->ndata) return set_err(CS_EINVAL, "no training set: call cs_mlp_set_data first");
  if (batch < 32 || batch > m->ndata) return set_err(CS_ERANGE, "batch must be in 32..%llu", (unsigned long long)m->ndata);
  HIP_TRY(hipSetDevice(m->device));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t st = 0; st < steps; st++) {
    if (m->cursor + batch > m->ndata) {
      /* a new epoch: Fisher-Yates with a 64-bit LCG, uploaded once */
      for (uint64_t k = m->ndata - 1; k > 0; k--) {
        m->shuffle_state = m->shuffle_state * 6364136223846793005ull + 1442695040888963407ull;
        const uint64_t j = (m->shuffle_state >> 33) % (k + 1);
        std::swap(m->order[k], m->order[j]);
      }
      HIP_TRY(hipMemcpyAsync(m->d_idx, m->order.data(), sizeof(uint32_t) * m->ndata, hipMemcpyHostToDevice, m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      m->cursor = 0;
    }
    int rc = run_grad(m, m->d_x, m->d_t, m->d_idx + m->cursor, batch, (float)batch);
    if (rc != CS_OK) return rc;
    m->cursor += batch;
    m->adam_step++;
    const float c1 = 1.0f - powf(b1, (float)m->adam_step), c2 = 1.0f - powf(b2, (float)m->adam_step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->d_p, m->d_m, m->d_v,
                       m->d_grad, lr, b1, b2, eps, c1, c2);
    if ((rc = pack(m)) != CS_OK) return rc;
  }
  double sse = 0.0;
  HIP_TRY(hipMemcpyAsync(&sse, m->d_sse_out, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (last_loss) *last_loss = sse / (double)batch;
  return CS_OK;
}

}  // extern "C"

/* The device image of a network, for the ensemble's closure gating. */
const void* cs_internal_mlp_device(const cs_mlp* m) { return m ? m->d_dev : nullptr; }
