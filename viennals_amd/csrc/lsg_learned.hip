// Training and batch-inference kernels of the learned weno5 weights.
//
// The network (5 -> 32 -> 32 -> 3, GELU, softmax) maps the normalised
// one-sided differences of a weno5 stencil to the weights of its three
// candidate polynomials.  A training sample is 8 floats (5 inputs, the 3
// candidate polynomials divided by max|d|) plus a target (the fine-grid
// derivative times dx / max|d|); the loss is the mean squared error of
// sum_k w_k p_k.
//
// kMlpGrad: one wave per workgroup, 64 samples per tile (lane = sample).
// The 32x32 hidden layer runs on the matrix cores
// (v_mfma_f32_32x32x16_bf16: bf16 operands, fp32 accumulate) for the
// forward product H1 W2^T, the input gradient dZ2 W2 and the weight
// gradient dZ2^T H1 (accumulated over all tiles of the wave in the MFMA
// accumulators).  The 5->32 and 32->3 layers, GELU, softmax and loss run on
// the vector ALUs in fp32, with the inputs and the layer-2 activations
// rounded to bf16 where the inference's matrix products take them.  Per-wave partial gradients go to a buffer that a
// second kernel sums in a fixed order (deterministic), then Adam updates the
// fp32 master parameters.
//
// MFMA lane layouts (32x32x16 bf16): A operand lane l holds row l%32,
// k = 8*(l/32) .. +7; B operand lane l holds column l%32, same k; the fp32
// accumulator element r of lane l is row 8*(r/4) + 4*(l/32) + r%4,
// column l%32.
#include <hip/hip_runtime.h>

#include "lsg_device.hpp"
#include "lsg_items.hpp"
#include "lsg_learned.hpp"

namespace lsg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// GELU x sigmoid(2u) (mlpGelu) and its derivative
// sigmoid(2u) + 2 x sigmoid(2u) (1 - sigmoid(2u)) u'.
__device__ inline float geluD(float x, float &dg) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float x2 = x * x;
  const float u = c * (x + a * x2 * x);
  const float sg = mlpSigmoid2(u);
  dg = sg + 2.0f * x * sg * (1.0f - sg) * c * (1.0f + 3.0f * a * x2);
  return x * sg;
}

__device__ inline bf16x8 packBf16(const float *v) {
  i16x8 s;
  s[0] = (short)bf16Bits(v[0]);
  s[1] = (short)bf16Bits(v[1]);
  s[2] = (short)bf16Bits(v[2]);
  s[3] = (short)bf16Bits(v[3]);
  s[4] = (short)bf16Bits(v[4]);
  s[5] = (short)bf16Bits(v[5]);
  s[6] = (short)bf16Bits(v[6]);
  s[7] = (short)bf16Bits(v[7]);
  return __builtin_bit_cast(bf16x8, s);
}
__device__ inline bf16x8 packBits(const uint16_t *v) {
  i16x8 s;
  for (int t = 0; t < 8; ++t) s[t] = (short)v[t];
  return __builtin_bit_cast(bf16x8, s);
}

constexpr int HP = 40;  // padded row length (bf16) of the bf16 tiles
constexpr int FP = 33;  // padded row length (fp32) of the fp32 tiles

struct MlpLds {
  uint16_t w2[MLP_H][HP];   // W2 as bf16, row j (output), column i (input)
  uint16_t w2t[MLP_H][HP];  // W2 transposed
  float w1[MLP_H * MLP_IN], b1[MLP_H], b2[MLP_H], w3[MLP_OUT * MLP_H], b3[4];
  uint16_t h1b[64][HP];     // bf16 hidden activations of layer 1, row = sample
  float buf1[64][FP];       // z2, then dz2 (fp32)
  float buf2[64][FP];       // h2
  float buf3[64][FP];       // dh1, then dz1
  float xs[64][MLP_IN];
  float dz3[64][4];
};

// one-wave workgroups: the barrier orders the LDS accesses of the wave
__device__ inline void waveSync() { __syncthreads(); }

// Partial gradient of one wave over its tiles of samples [0, n): row
// partials[blockIdx.x][0..MLP_NPARAM) and the summed squared residual at
// [MLP_NPARAM].  invCount = 1 / (number of samples of the loss).
__global__ __launch_bounds__(64) void kMlpGrad(const float *X, const float *T, uint32_t n, float invCount,
                                              const float *P, float *partials) {
  __shared__ MlpLds L;
  const int l = threadIdx.x;
  // stage the parameters (weights as their bf16 values)
  for (int e = l; e < MLP_H * MLP_H; e += 64) {
    const int j = e / MLP_H, i = e % MLP_H;
    const uint16_t b = bf16Bits(P[MLP_W2 + e]);
    L.w2[j][i] = b;
    L.w2t[i][j] = b;
  }
  for (int e = l; e < MLP_H * MLP_IN; e += 64) L.w1[e] = bf16Float(bf16Bits(P[MLP_W1 + e]));
  for (int e = l; e < MLP_OUT * MLP_H; e += 64) L.w3[e] = bf16Float(bf16Bits(P[MLP_W3 + e]));
  if (l < MLP_H) {
    L.b1[l] = P[MLP_B1 + l];
    L.b2[l] = P[MLP_B2 + l];
  }
  if (l < MLP_OUT) L.b3[l] = P[MLP_B3 + l];
  waveSync();

  const int col = l & 31;
  const int half = l >> 5;
  f32x16 accW2;
  for (int r = 0; r < 16; ++r) accW2[r] = 0.f;
  float aW1[MLP_IN] = {0, 0, 0, 0, 0}, aB1 = 0, aB2 = 0, aW3[MLP_OUT] = {0, 0, 0}, aB3 = 0;
  float lossAcc = 0;

  for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const uint32_t s = base + l;
    float x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float t = 0;
    if (s < n) {
      for (int q = 0; q < 8; ++q) x[q] = X[(size_t)s * 8 + q];
      t = T[s];
    }
    // the inputs enter layer 1 as bf16 (as in the matrix-core inference)
    for (int i = 0; i < MLP_IN; ++i) x[i] = bf16Float(bf16Bits(x[i]));
    // layer 1 (VALU), keep gelu'
    float d1[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
      float z = L.b1[j];
      for (int i = 0; i < MLP_IN; ++i) z += L.w1[j * MLP_IN + i] * x[i];
      const float h = geluD(z, d1[j]);
      L.h1b[l][j] = bf16Bits(h);
    }
    for (int i = 0; i < MLP_IN; ++i) L.xs[l][i] = x[i];
    waveSync();
    // layer 2 forward on the matrix cores: Z2[s][j] = sum_i H1[s][i] W2[j][i]
    for (int rb = 0; rb < 2; ++rb) {
      f32x16 acc;
      for (int r = 0; r < 16; ++r) acc[r] = 0.f;
      for (int kb = 0; kb < MLP_H; kb += 16) {
        const bf16x8 a = packBits(&L.h1b[rb * 32 + col][kb + 8 * half]);
        const bf16x8 b = packBits(&L.w2[col][kb + 8 * half]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
      for (int r = 0; r < 16; ++r) L.buf1[rb * 32 + 8 * (r >> 2) + 4 * half + (r & 3)][col] = acc[r];
    }
    waveSync();
    // layer 2 activation, layer 3, softmax, loss (lane = sample)
    float d2[MLP_H];
    float z3[MLP_OUT] = {L.b3[0], L.b3[1], L.b3[2]};
    for (int j = 0; j < MLP_H; ++j) {
      const float z = L.buf1[l][j] + L.b2[j];
      const float h = bf16Float(bf16Bits(geluD(z, d2[j])));  // enters layer 3 as bf16
      L.buf2[l][j] = h;
      for (int k = 0; k < MLP_OUT; ++k) z3[k] += L.w3[k * MLP_H + j] * h;
    }
    const float m = fmaxf(z3[0], fmaxf(z3[1], z3[2]));
    float w[MLP_OUT], se = 0;
    for (int k = 0; k < MLP_OUT; ++k) {
      w[k] = mlpExp(z3[k] - m);
      se += w[k];
    }
    for (int k = 0; k < MLP_OUT; ++k) w[k] = w[k] / se;
    const float pred = w[0] * x[5] + w[1] * x[6] + w[2] * x[7];
    const float res = (s < n) ? pred - t : 0.f;
    lossAcc += res * res;
    const float dpred = 2.f * res * invCount;
    float dw[MLP_OUT], dz3[MLP_OUT];
    for (int k = 0; k < MLP_OUT; ++k) dw[k] = dpred * x[5 + k];
    const float wdw = w[0] * dw[0] + w[1] * dw[1] + w[2] * dw[2];
    for (int k = 0; k < MLP_OUT; ++k) {
      dz3[k] = w[k] * (dw[k] - wdw);
      L.dz3[l][k] = dz3[k];
    }
    waveSync();  // every lane has read its z2 row before it is overwritten
    for (int i = 0; i < MLP_H; ++i) {
      float dh = 0;
      for (int k = 0; k < MLP_OUT; ++k) dh += L.w3[k * MLP_H + i] * dz3[k];
      L.buf1[l][i] = dh * d2[i];  // dz2
    }
    waveSync();
    // layer 2 backward on the matrix cores
    // dH1[s][i] = sum_j dZ2[s][j] W2[j][i]
    for (int rb = 0; rb < 2; ++rb) {
      f32x16 acc;
      for (int r = 0; r < 16; ++r) acc[r] = 0.f;
      for (int kb = 0; kb < MLP_H; kb += 16) {
        const bf16x8 a = packBf16(&L.buf1[rb * 32 + col][kb + 8 * half]);
        const bf16x8 b = packBits(&L.w2t[col][kb + 8 * half]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
      for (int r = 0; r < 16; ++r) L.buf3[rb * 32 + 8 * (r >> 2) + 4 * half + (r & 3)][col] = acc[r];
    }
    // dW2[j][i] += sum_s dZ2[s][j] H1[s][i]  (k = sample)
    for (int kb = 0; kb < 64; kb += 16) {
      float av[8];
      uint16_t bv[8];
      for (int q = 0; q < 8; ++q) {
        av[q] = L.buf1[kb + 8 * half + q][col];
        bv[q] = L.h1b[kb + 8 * half + q][col];
      }
      accW2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(packBf16(av), packBits(bv), accW2, 0, 0, 0);
    }
    waveSync();
    for (int i = 0; i < MLP_H; ++i) L.buf3[l][i] = L.buf3[l][i] * d1[i];  // dz1
    waveSync();
    // column sums (lane: column col, samples half*32 .. +31)
    for (int q = 0; q < 32; ++q) {
      const int sm = half * 32 + q;
      const float dz1 = L.buf3[sm][col];
      for (int i = 0; i < MLP_IN; ++i) aW1[i] += dz1 * L.xs[sm][i];
      aB1 += dz1;
      aB2 += L.buf1[sm][col];
      const float h2 = L.buf2[sm][col];
      for (int k = 0; k < MLP_OUT; ++k) aW3[k] += L.dz3[sm][k] * h2;
      if (col < MLP_OUT) aB3 += L.dz3[sm][col];
    }
    waveSync();
  }
  // combine the two halves and write the partial row
  float *out = partials + (size_t)blockIdx.x * MLP_PSTRIDE;
  for (int i = 0; i < MLP_IN; ++i) aW1[i] += __shfl_down(aW1[i], 32);
  aB1 += __shfl_down(aB1, 32);
  aB2 += __shfl_down(aB2, 32);
  for (int k = 0; k < MLP_OUT; ++k) aW3[k] += __shfl_down(aW3[k], 32);
  aB3 += __shfl_down(aB3, 32);
  if (l < 32) {
    for (int i = 0; i < MLP_IN; ++i) out[MLP_W1 + col * MLP_IN + i] = aW1[i];
    out[MLP_B1 + col] = aB1;
    out[MLP_B2 + col] = aB2;
    for (int k = 0; k < MLP_OUT; ++k) out[MLP_W3 + k * MLP_H + col] = aW3[k];
    if (col < MLP_OUT) out[MLP_B3 + col] = aB3;
  }
  for (int r = 0; r < 16; ++r) out[MLP_W2 + (8 * (r >> 2) + 4 * half + (r & 3)) * MLP_H + col] = accW2[r];
  for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
  if (l == 0) out[MLP_NPARAM] = lossAcc;
}

// Sum of the partial rows in block order (deterministic).
__global__ void kMlpReduce(const float *partials, int rows, float *grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_PSTRIDE) return;
  float s = 0;
  for (int r = 0; r < rows; ++r) s += partials[(size_t)r * MLP_PSTRIDE + p];
  grad[p] = s;
}

// Adam on the fp32 master parameters (Kingma & Ba 2015), step t >= 1.
__global__ void kMlpAdam(float *P, float *M, float *V, const float *G, int t, float lr, float b1, float b2,
                         float eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  const float mh = m / (1.f - powf(b1, (float)t));
  const float vh = v / (1.f - powf(b2, (float)t));
  P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Learned weno5 on n stencils of 7 values (test and inspection entry).
__global__ void kWenoLearned(uint32_t n, const double *x, double dx, int plus, const uint16_t *prm, double *out) {
  __shared__ float W[MLP_NPARAM];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) W[i] = bf16Float(prm[i]);
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double st[7];
    for (int k = 0; k < 7; ++k) st[k] = x[(size_t)i * 7 + k];
    out[i] = weno5Learned(st, dx, plus != 0, W);
  }
}

void launchWenoLearned(hipStream_t s, uint32_t n, const double *x, double dx, int plus, const uint16_t *prm,
                       double *out) {
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  kWenoLearned<<<blocks, 256, 0, s>>>(n, x, dx, plus, prm, out);
}

void launchMlpGrad(hipStream_t s, const float *X, const float *T, uint32_t n, const float *P, float *partials,
                   int rows, float *grad) {
  kMlpGrad<<<rows, 64, 0, s>>>(X, T, n, 1.0f / (float)n, P, partials);
  kMlpReduce<<<(MLP_PSTRIDE + 255) / 256, 256, 0, s>>>(partials, rows, grad);
}

void launchMlpAdam(hipStream_t s, float *P, float *M, float *V, const float *G, int t, float lr) {
  kMlpAdam<<<(MLP_NPARAM + 255) / 256, 256, 0, s>>>(P, M, V, G, t, lr, 0.9f, 0.999f, 1e-8f);
}

}  // namespace lsg
