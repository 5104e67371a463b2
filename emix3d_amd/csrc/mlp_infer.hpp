// Inference of the learned eddy-viscosity closure on the matrix cores: the
// 10 -> 64 -> 64 -> 64 -> 1 network (SiLU hidden, softplus output) for 64
// grid points per wavefront, one point per lane.
//
// Each hidden layer is computed transposed, H^T[64 neurons x 16 points] =
// W[64 x K] X^T[K x 16], with v_mfma_f32_16x16x32_bf16 (BF16 operands,
// FP32 accumulation): four 16-neuron tiles (mt) per layer and, per
// wavefront, four 16-point tiles (t) side by side, whose sixteen
// independent products per layer keep the matrix core busy while the
// activations of the previous tile are formed.  The 16x16 result a lane
// holds (neurons 16mt + 4g + i of point n, g = lane/16, n = lane%16) is
// already that lane's B operand of the next layer once the K order of the
// next layer's weights is permuted to match (perm_in): hidden activations
// never leave the lane.  The features go in through a wave-private LDS
// transpose (10 x 64 floats), and the output neuron, also a matrix-core
// product (its weights in every row), leaves lane l with the value of
// point l of the group.
#pragma once
#include "compact.hpp"

namespace emx {

constexpr int MLP_IN = 10, MLP_H = 64;
constexpr int MW1 = 0, MB1 = MW1 + MLP_H * MLP_IN, MW2 = MB1 + MLP_H,
              MB2 = MW2 + MLP_H * MLP_H, MW3 = MB2 + MLP_H, MB3 = MW3 + MLP_H * MLP_H,
              MW4 = MB3 + MLP_H, MB4 = MW4 + MLP_H, MNP = MB4 + 1;
static_assert(MNP == EMX_MLP_NPARAM, "parameter count");

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

// Packed A operands (k_mlp_pack): A1[mt][lane][8] (layer 1, features
// padded to K = 32), A2[mt][s][lane][8], A3[mt][s][lane][8], A4[s][lane][8]
// (the output neuron, its weights in all 16 rows); lane l holds row
// m = l%16 (neuron 16mt+m), K = 8(l/16) + j.
constexpr int MLP_PACK_A1 = 0, MLP_PACK_A2 = 4 * 64 * 8, MLP_PACK_A3 = MLP_PACK_A2 + 8 * 64 * 8,
              MLP_PACK_A4 = MLP_PACK_A3 + 8 * 64 * 8, MLP_PACK_ELEMS = MLP_PACK_A4 + 2 * 64 * 8;

// Input neuron of K index k = 8g + j (lane group g, element j) in K-step s
// of layers 2 and 3: the lane's own result elements of output tiles 2s and
// 2s+1 (neurons 16t + 4g + i).
__host__ __device__ __forceinline__ int perm_in(int s, int g, int j) {
  return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// The hidden layers run in a scaled form that leaves one exponential, one
// reciprocal, one add and one multiply per activation: with s = -log2(e),
// each layer forms y = s z (its bias scaled by s, and layer 1's weights),
// so that e^-z = 2^y, and keeps h~ = y / (1 + 2^y) = s silu(z) as the next
// layer's input; the next layer's y' = s (W h + b) = W h~ + s b needs no
// other change, and the output layer's weights are divided by s.  The
// reciprocal is the hardware one (1 ulp; the BF16 operands carry 8 bits).
constexpr float MLP_YS = -1.44269504088896340736f;
__device__ __forceinline__ float silu_scaled(float y) {
  return y * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y));
}
// softplus(z) = max(z, 0) + log1p(e), e = exp(-|z|) in (0, 1], with the
// hardware base-2 exponential and logarithm: log1p(e) = ln(u) * e / (u - 1)
// for u = 1 + e rounded (Goldberg's correction: a few ulp), and e itself
// where u rounds to 1.  One point per lane, 14 instructions instead of the
// ~100 of the library's log1pf.
__device__ __forceinline__ float softplus_fast(float z) {
  const float e = __builtin_amdgcn_exp2f(-1.44269504088896340736f * __builtin_fabsf(z));
  const float u = 1.0f + e;
  const float d = u - 1.0f;
  const float lp = d == 0.0f ? e
                             : 0.69314718055994530942f * __builtin_amdgcn_logf(u) *
                                   (e * __builtin_amdgcn_rcpf(d));
  return __builtin_fmaxf(z, 0.0f) + lp;
}

// The packed A operands staged in LDS by a workgroup (22 KB, MLP_PACK_ELEMS
// BF16 values in units of bf16x8: layer 1 at 0, layer 2 at 256, layer 3 at
// 768, the output neuron at 1280).  A wavefront reads a layer's operands
// from LDS for each point tile, so weights hold few registers (three
// wavefronts per SIMD fit).
constexpr int MLP_LW = MLP_PACK_ELEMS / 8;
__device__ __forceinline__ void mlp_stage_weights(bf16x8* lw, const __bf16* A) {
  const bf16x8* src = reinterpret_cast<const bf16x8*>(A);
  for (int t = threadIdx.x; t < MLP_LW; t += blockDim.x) lw[t] = src[t];
}

// LDS copy of the vectors of the network, shared by a workgroup, in the
// scaled form: s b1[64], s b2[64], s b3[64], w4[64] / s (unused by the
// matrix-core output layer), then b4 and out_scale.
constexpr int MLP_VEC = 4 * MLP_H + 2;
__device__ __forceinline__ void mlp_stage_vectors(float* vec, const float* P, const float* norm) {
  for (int t = threadIdx.x; t < MLP_VEC; t += blockDim.x) {
    float v;
    if (t < MLP_H) v = MLP_YS * P[MB1 + t];
    else if (t < 2 * MLP_H) v = MLP_YS * P[MB2 + t - MLP_H];
    else if (t < 3 * MLP_H) v = MLP_YS * P[MB3 + t - 2 * MLP_H];
    else if (t < 4 * MLP_H) v = P[MW4 + t - 3 * MLP_H] / MLP_YS;
    else if (t == 4 * MLP_H) v = P[MB4];
    else v = norm[2 * MLP_IN];
    vec[t] = v;
  }
}

// Bias (or w4) of neurons 16t + 4g + 0..3 of vector row r.
__device__ __forceinline__ f32x4 mlp_vec4(const float* vec, int r, int t, int g) {
  return *reinterpret_cast<const f32x4*>(vec + r * MLP_H + 16 * t + 4 * g);
}

// nu_t (out_scale * softplus) of the lane's own point of a 64-point group.
// lw: the staged weights; f: the lane's normalised features; xs: this
// wave's LDS [MLP_IN][64]; vec: the staged vectors.  Every lane of the wave
// must call it.
__device__ __forceinline__ float mlp_group64(const bf16x8* lw, const float f[MLP_IN],
                                             float* xs, const float* vec, int l) {
  const int g = l >> 4, n = l & 15;
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) xs[k * 64 + l] = f[k];
  wave_sync();
  // activations, already the next layer's B operands: hb[t][s] holds
  // neurons 16(2s) + 4g + 0..3 and 16(2s+1) + 4g + 0..3 of point 16t + n
  bf16x8 hb[4][2];
  // layer 1: B operand of tile t = features 8g..8g+7 of point 16t+n (K
  // padded with zeros to 32)
#pragma unroll
  for (int t = 0; t < 4; t++) {
    bf16x8 a1[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a1[mt] = lw[mt * 64 + l];
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = 8 * g + j;
      const float v = xs[(k < MLP_IN ? k : MLP_IN - 1) * 64 + 16 * t + n];
      x[j] = (__bf16)(k < MLP_IN ? v : 0.0f);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      // the (scaled) bias is the accumulator's initial value
      f32x4 acc = mlp_vec4(vec, 0, mt, g);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[mt], x, acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) hb[t][mt >> 1][4 * (mt & 1) + i] = (__bf16)silu_scaled(acc[i]);
    }
  }
  wave_sync();  // xs is rewritten by the next group
  // layers 2 and 3
#pragma unroll
  for (int layer = 0; layer < 2; layer++) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      bf16x8 a[4][2];
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int s = 0; s < 2; s++) a[mt][s] = lw[256 + layer * 512 + (mt * 2 + s) * 64 + l];
      const bf16x8 bop[2] = {hb[t][0], hb[t][1]};
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 acc = mlp_vec4(vec, 1 + layer, mt, g);
#pragma unroll
        for (int s = 0; s < 2; s++)
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][s], bop[s], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
          hb[t][mt >> 1][4 * (mt & 1) + i] = (__bf16)silu_scaled(acc[i]);
      }
    }
  }
  // output neuron on the matrix cores: with w4 / s in every row of A, every
  // row of the 16x16 result holds z of the tile's 16 points, so the lane of
  // point 16t + n holds it for t = its own lane group g: lane l keeps point l
  float zs = 0.0f;
  const bf16x8 a4[2] = {lw[1280 + l], lw[1280 + 64 + l]};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; s++)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a4[s], hb[t][s], acc, 0, 0, 0);
    zs = g == t ? acc[0] : zs;
  }
  return vec[4 * MLP_H + 1] * softplus_fast(zs + vec[4 * MLP_H]);
}

// The ten normalised features of point p: du_a/dx_b = jbn(1+4b) *
// dvel(3b+a) at index 3a+b, then delta; x = (f - shift) * scale.
__device__ __forceinline__ void mlp_features(const double* dvel, const double* jbn,
                                             const double* delta, long long N, long long p,
                                             const float* norm, float x[MLP_IN]) {
  const double m[3] = {jbn[p], jbn[4 * N + p], jbn[8 * N + p]};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
      x[3 * a + b] = (float)((m[b] * dvel[(3 * b + a) * N + p] - norm[3 * a + b]) *
                             norm[MLP_IN + 3 * a + b]);
  x[9] = (float)((delta[p] - norm[9]) * norm[MLP_IN + 9]);
}

}  // namespace emx
