// Learned eddy-viscosity closure: a 10 -> 64 -> 64 -> 64 -> 1 network
// (SiLU hidden activations, softplus output) of the physical velocity
// gradient and the filter width.  The reference has no counterpart.
//
// Parameters (FP32, EMX_MLP_NPARAM values):
//   W1[64][10] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
// Normalisation (FP32, EMX_MLP_NNORM): in_shift[10] in_scale[10] out_scale,
//   x_k = (f_k - in_shift_k) * in_scale_k,  nu = out_scale * softplus(z4).
// Features f of a grid point: du_a/dx_b = jbn(1+4b) * dvel(3b+a) at index
// 3a+b, then delta.
//
// Inference (k_mlp_infer) runs the three hidden layers on the matrix cores:
// v_mfma_f32_16x16x32_bf16, BF16 operands, FP32 accumulation.  Each layer
// is computed transposed, H^T[64 x 16 points] = W[64 x K] X^T[K x 16], so
// the 16x16 result tile a lane holds (4 neurons of one point) is already
// the B operand of the next layer once the K order of that layer's weights
// is permuted to match (perm_in below): no data moves between lanes.
// Training (FP32 vector code) stores pre-activations, back-propagates
// per point and reduces the weight gradients per block, then over blocks.
#include <algorithm>

#include "mlp_infer.hpp"
#include "ops_common.hpp"

namespace emx {

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float softplus_f(float z) {
  return z > 20.0f ? z : log1pf(__expf(z));
}

// Physical gradient and filter width of point p (the 10 features).
__device__ __forceinline__ void point_features(const double* dvel, const double* jbn,
                                               const double* delta, long long N,
                                               long long p, double f[MLP_IN]) {
  const double m[3] = {jbn[p], jbn[4 * N + p], jbn[8 * N + p]};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) f[3 * a + b] = m[b] * dvel[(3 * b + a) * N + p];
  f[9] = delta[p];
}

// Packs the parameters into per-lane BF16 A operands, one work-item per
// packed value (MLP_PACK_ELEMS of them, in units of 64 lanes x 8):
//   A1[mt][lane][8]      layer 1 (features padded to K = 32)
//   A2[mt][s][lane][8], A3[mt][s][lane][8]   layers 2, 3
//   A4[s][lane][8]       the output neuron's weights (divided by the scale
//                        of the hidden layers) in every row
// lane l holds row m = l%16 (output neuron 16mt+m), K = 8(l/16) + j.
__global__ void k_mlp_pack(const float* P, __bf16* A) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_PACK_ELEMS) return;
  const int j = e & 7, l = (e >> 3) & 63, unit = e >> 9;
  const int m = l % 16, g = l / 16;
  float v;
  if (e < MLP_PACK_A2) {
    // layer 1 in the scaled form of mlp_group64: y = MLP_YS * z
    const int k = 8 * g + j;
    v = k < MLP_IN ? MLP_YS * P[MW1 + (16 * unit + m) * MLP_IN + k] : 0.0f;
  } else if (e < MLP_PACK_A4) {
    const int layer = e < MLP_PACK_A3 ? 0 : 1;
    const int r = unit - 4 - 8 * layer, mt = r >> 1, sk = r & 1;
    v = P[(layer == 0 ? MW2 : MW3) + (16 * mt + m) * MLP_H + perm_in(sk, g, j)];
  } else {
    v = P[MW4 + perm_in(unit - 20, g, j)] / MLP_YS;
  }
  A[e] = (__bf16)v;
}

// nu_t of every grid point: persistent wavefronts, each taking groups of
// 64 consecutive points (one per lane, loads coalesced).  Three wavefronts
// per SIMD (at most 168 registers, no spills) hide one group's loads behind
// the others' networks; keeping the next group's inputs in registers
// instead allows only two and was measured slower (0.97 against 0.88 ms at
// 256^3).  Grids hold at most 2^28 points (emx_create), so every access is
// a uniform field base plus a 32-bit byte offset: no 64-bit address
// arithmetic per load.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_mlp_infer(const double* dvel, const double* jbn, const double* delta, long long N,
            const float* P, const float* norm, const __bf16* A, double* nut, double c4,
            double* vis) {
  __shared__ __attribute__((aligned(16))) float vec[MLP_VEC + 2];
  __shared__ double nrm[2 * MLP_IN];  // in_shift, in_scale
  __shared__ float xs[4][MLP_IN * 64];
  __shared__ bf16x8 lw[MLP_LW];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  mlp_stage_weights(lw, A);
  mlp_stage_vectors(vec, P, norm);
  for (int t = threadIdx.x; t < 2 * MLP_IN; t += blockDim.x) nrm[t] = (double)norm[t];
  __syncthreads();
  const long long ngroup = (N + 63) / 64;
  const long long nwave = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long grp = (long long)blockIdx.x * (blockDim.x >> 6) + w; grp < ngroup;
       grp += nwave) {
    const long long p = grp * 64 + l;
    const unsigned off = (unsigned)(p < N ? p : N - 1) * 8u;
    auto at = [off](const double* base) -> const double& {
      return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + off);
    };
    // raw inputs of the point: dvel 1..9, jbn 1, 5, 9, delta
    double r[13];
#pragma unroll
    for (int c = 0; c < 9; c++) r[c] = at(dvel + c * N);
    r[9] = at(jbn);
    r[10] = at(jbn + 4 * N);
    r[11] = at(jbn + 8 * N);
    r[12] = at(delta);
    float x[MLP_IN];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
        x[3 * a + b] =
            (float)((r[9 + b] * r[3 * b + a] - nrm[3 * a + b]) * nrm[MLP_IN + 3 * a + b]);
    x[9] = (float)((r[12] - nrm[9]) * nrm[MLP_IN + 9]);
    const float nu = mlp_group64(lw, x, xs[w], vec, l);
    if (p < N) {
      double* const np = reinterpret_cast<double*>(reinterpret_cast<char*>(nut) + off);
      *np = (double)nu;
      if (vis) {
        double* const vp = reinterpret_cast<double*>(reinterpret_cast<char*>(vis) + off);
        *vp = c4 * 1.0 + (double)nu;  // viscosidad (solver/viscosidad_incomp.f90:15-17)
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Training kernels (FP32 vector code).

// Features of the listed points, [n][10] (indices clamped to [0, N)).
__global__ void k_mlp_gather(const double* dvel, const double* jbn, const double* delta,
                             long long N, const long long* idx, int n, float* feat) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    double f[MLP_IN];
    const long long p = idx[q] < 0 ? 0 : (idx[q] >= N ? N - 1 : idx[q]);
    point_features(dvel, jbn, delta, N, p, f);
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) feat[(long long)q * MLP_IN + k] = (float)f[k];
  }
}

// act[q] = z1[64] z2[64] z3[64] z4 (pre-activations), nu[q].
constexpr int MLP_ACT = 3 * MLP_H + 1;

__global__ void __launch_bounds__(256)
k_mlp_forward(const float* P, const float* norm, const float* feat, int n, float* act,
              float* nu) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    float x[MLP_IN], h[MLP_H], hn[MLP_H];
    float* a = act + (long long)q * MLP_ACT;
#pragma unroll
    for (int k = 0; k < MLP_IN; k++)
      x[k] = (feat[(long long)q * MLP_IN + k] - norm[k]) * norm[MLP_IN + k];
    for (int o = 0; o < MLP_H; o++) {
      float s = P[MB1 + o];
#pragma unroll
      for (int k = 0; k < MLP_IN; k++) s += P[MW1 + o * MLP_IN + k] * x[k];
      a[o] = s;
      h[o] = silu_f(s);
    }
    for (int o = 0; o < MLP_H; o++) {
      float s = P[MB2 + o];
#pragma unroll 16
      for (int k = 0; k < MLP_H; k++) s += P[MW2 + o * MLP_H + k] * h[k];
      a[MLP_H + o] = s;
      hn[o] = silu_f(s);
    }
    for (int o = 0; o < MLP_H; o++) {
      float s = P[MB3 + o];
#pragma unroll 16
      for (int k = 0; k < MLP_H; k++) s += P[MW3 + o * MLP_H + k] * hn[k];
      a[2 * MLP_H + o] = s;
      h[o] = silu_f(s);
    }
    float s = P[MB4];
#pragma unroll 16
    for (int k = 0; k < MLP_H; k++) s += P[MW4 + k] * h[k];
    a[3 * MLP_H] = s;
    nu[q] = norm[2 * MLP_IN] * softplus_f(s);
  }
}

// Per-point back-propagation of dnu: dl[q] = d1[64] d2[64] d3[64] d4
// (derivatives with respect to the pre-activations).
__global__ void __launch_bounds__(256)
k_mlp_backward(const float* P, const float* norm, const float* act, const float* dnu,
               int n, float* dl) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const float* a = act + (long long)q * MLP_ACT;
    float* d = dl + (long long)q * MLP_ACT;
    const float z4 = a[3 * MLP_H];
    const float d4 = dnu[q] * norm[2 * MLP_IN] / (1.0f + __expf(-z4));
    d[3 * MLP_H] = d4;
    float d3[MLP_H], d2[MLP_H];
    for (int k = 0; k < MLP_H; k++) {
      d3[k] = d4 * P[MW4 + k] * dsilu_f(a[2 * MLP_H + k]);
      d[2 * MLP_H + k] = d3[k];
    }
    for (int k = 0; k < MLP_H; k++) {
      float s = 0.0f;
#pragma unroll 16
      for (int o = 0; o < MLP_H; o++) s += P[MW3 + o * MLP_H + k] * d3[o];
      d2[k] = s * dsilu_f(a[MLP_H + k]);
      d[MLP_H + k] = d2[k];
    }
    for (int k = 0; k < MLP_H; k++) {
      float s = 0.0f;
#pragma unroll 16
      for (int o = 0; o < MLP_H; o++) s += P[MW2 + o * MLP_H + k] * d2[o];
      d[k] = s * dsilu_f(a[k]);
    }
  }
}

// Weight gradients, partial per block: block b sums points
// [b*chunk, (b+1)*chunk) into part[b][MNP].  Points are staged through LDS
// in groups of 16: inputs x (10), activations h1..h3 (3x64), deltas
// d1..d3 (3x64) and d4.
constexpr int WG_PTS = 16;
__global__ void __launch_bounds__(256)
k_mlp_wgrad(const float* norm, const float* feat, const float* act, const float* dl, int n,
            int chunk, float* part) {
  __shared__ float sx[WG_PTS][MLP_IN];
  __shared__ float sh[WG_PTS][3][MLP_H];  // h1, h2, h3
  __shared__ float sd[WG_PTS][MLP_ACT];   // d1, d2, d3, d4
  const int q0 = blockIdx.x * chunk;
  const int q1 = min(q0 + chunk, n);
  float acc[(MNP + 255) / 256];
#pragma unroll
  for (int r = 0; r < (MNP + 255) / 256; r++) acc[r] = 0.0f;
  for (int qb = q0; qb < q1; qb += WG_PTS) {
    __syncthreads();
    for (int t = threadIdx.x; t < WG_PTS * MLP_ACT; t += blockDim.x) {
      const int pp = t / MLP_ACT, c = t % MLP_ACT;
      const int q = qb + pp;
      float dv = 0.0f, hv = 0.0f;
      if (q < q1) {
        dv = dl[(long long)q * MLP_ACT + c];
        if (c < 3 * MLP_H) hv = silu_f(act[(long long)q * MLP_ACT + c]);
      }
      sd[pp][c] = dv;
      if (c < 3 * MLP_H) sh[pp][c / MLP_H][c % MLP_H] = hv;
    }
    for (int t = threadIdx.x; t < WG_PTS * MLP_IN; t += blockDim.x) {
      const int pp = t / MLP_IN, k = t % MLP_IN;
      const int q = qb + pp;
      sx[pp][k] = q < q1 ? (feat[(long long)q * MLP_IN + k] - norm[k]) * norm[MLP_IN + k] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < (MNP + 255) / 256; r++) {
      const int w = r * 256 + threadIdx.x;
      if (w >= MNP) continue;
      float s = 0.0f;
      if (w < MB1) {
        const int o = w / MLP_IN, k = w % MLP_IN;
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][o] * sx[pp][k];
      } else if (w < MW2) {
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][w - MB1];
      } else if (w < MB2) {
        const int o = (w - MW2) / MLP_H, k = (w - MW2) % MLP_H;
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][MLP_H + o] * sh[pp][0][k];
      } else if (w < MW3) {
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][MLP_H + w - MB2];
      } else if (w < MB3) {
        const int o = (w - MW3) / MLP_H, k = (w - MW3) % MLP_H;
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][2 * MLP_H + o] * sh[pp][1][k];
      } else if (w < MW4) {
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][2 * MLP_H + w - MB3];
      } else if (w < MB4) {
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][3 * MLP_H] * sh[pp][2][w - MW4];
      } else {
        for (int pp = 0; pp < WG_PTS; pp++) s += sd[pp][3 * MLP_H];
      }
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < (MNP + 255) / 256; r++) {
    const int w = r * 256 + threadIdx.x;
    if (w < MNP) part[(long long)blockIdx.x * MNP + w] = acc[r];
  }
}

__global__ void k_mlp_reduce(const float* part, int nblk, float* grad) {
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < MNP; w += gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += part[(long long)b * MNP + w];
    grad[w] = s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ---------------------------------------------------------------------------

size_t mlp_pack_elems() { return (size_t)MLP_PACK_ELEMS; }

int op_mlp_pack(emx_ctx* ctx, const float* params, const float* norm) {
  __bf16* A = reinterpret_cast<__bf16*>(ctx->mlp_pack);
  EMX_TIMED(ctx, "mlp_pack", 0.0,
            hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_PACK_ELEMS + 255) / 256), dim3(256), 0,
                               ctx->stream, params, A));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_mlp_nut(emx_ctx* ctx, const float* params, const float* norm, const double* dvel,
               const double* jbn, const double* delta, double* nut, double c4, double* vis) {
  const long long N = ctx->d.N();
  if (N > (1LL << 29)) {  // k_mlp_infer's 32-bit byte offsets
    set_error("emx_mlp_nut: %lld points, at most 2^29", N);
    return EMX_ERR_SIZE;
  }
  __bf16* A = reinterpret_cast<__bf16*>(ctx->mlp_pack);
  EMX_TRY(op_mlp_pack(ctx, params, norm));
  static const int resident = resident_blocks(k_mlp_infer, 0);
  const long long groups = (N + 63) / 64;
  long long blocks = (groups + 3) / 4;
  if (blocks > resident) blocks = resident;
  // reads dvel (9), jbn (3), delta; writes nut (and vis)
  EMX_TIMED(ctx, "mlp_infer", (vis ? 15.0 : 14.0) * 8.0 * N,
            hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                               dvel, jbn, delta, N, params, norm, A, nut, c4, vis));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_mlp_gather(emx_ctx* ctx, const double* dvel, const double* jbn, const double* delta,
                  const long long* idx, int n, float* feat) {
  hipLaunchKernelGGL(k_mlp_gather, dim3(std::min(4096, (n + 255) / 256)), dim3(256), 0,
                     ctx->stream, dvel, jbn, delta, ctx->d.N(), idx, n, feat);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx

extern "C" {

int emx_mlp_forward_f32(const float* params, const float* norm, const float* feat, int n,
                        float* act, float* nu, void* stream) {
  if (!params || !norm || !feat || !act || !nu || n <= 0) {
    emx::set_error("emx_mlp_forward_f32: bad arguments");
    return EMX_ERR_ARG;
  }
  hipLaunchKernelGGL(emx::k_mlp_forward, dim3(std::min(8192, (n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, params, norm, feat, n, act, nu);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

size_t emx_mlp_backward_work(int n) {
  const int chunk = 256;
  const size_t nblk = (size_t)((n + chunk - 1) / chunk);
  return (size_t)n * emx::MLP_ACT + nblk * emx::MNP;
}

int emx_mlp_backward_f32(const float* params, const float* norm, const float* feat,
                         const float* act, const float* dnu, int n, float* grad, float* work,
                         size_t work_floats, void* stream) {
  if (!params || !norm || !feat || !act || !dnu || !grad || !work || n <= 0) {
    emx::set_error("emx_mlp_backward_f32: bad arguments");
    return EMX_ERR_ARG;
  }
  if (work_floats < emx_mlp_backward_work(n)) {
    emx::set_error("emx_mlp_backward_f32: work holds %zu floats, %zu needed", work_floats,
                   emx_mlp_backward_work(n));
    return EMX_ERR_SIZE;
  }
  const int chunk = 256;
  const int nblk = (n + chunk - 1) / chunk;
  float* dl = work;
  float* part = work + (size_t)n * emx::MLP_ACT;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(emx::k_mlp_backward, dim3(std::min(8192, (n + 255) / 256)), dim3(256), 0,
                     st, params, norm, act, dnu, n, dl);
  hipLaunchKernelGGL(emx::k_mlp_wgrad, dim3(nblk), dim3(256), 0, st, norm, feat, act, dl, n,
                     chunk, part);
  hipLaunchKernelGGL(emx::k_mlp_reduce, dim3((emx::MNP + 255) / 256), dim3(256), 0, st, part,
                     nblk, grad);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int emx_adam_f32(float* params, const float* grad, float* m, float* v, int n, float lr,
                 float beta1, float beta2, float eps, int step, void* stream) {
  if (!params || !grad || !m || !v || n <= 0 || step < 1) {
    emx::set_error("emx_adam_f32: bad arguments");
    return EMX_ERR_ARG;
  }
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(emx::k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     params, grad, m, v, n, lr, beta1, beta2, eps, c1, c2);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // extern "C"
