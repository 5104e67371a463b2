// Structure-function eddy viscosity (INISGDM geometry, viscosidad_t,
// viscosidad).
#include "ops_common.hpp"

namespace emx {

// INISGDM spacings and filter factors, solver/sgdm_incomp.f90:20-108.
// blockmax receives one max spacing per block.
__global__ void k_sgs_geometry(const double* x, Dims d, double* sgd,
                               double* dmins, double* blockmax) {
  __shared__ double red[256];
  const long long N = d.N();
  double mymax = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(p % d.nx);
    const int j = (int)((p / d.nx) % d.ny);
    const int k = (int)(p / ((long long)d.nx * d.ny));
    const long long sx = 1, sy = d.nx, sz = (long long)d.nx * d.ny;
    const double* X = x;
    const double* Y = x + N;
    const double* Z = x + 2 * N;
    double dxp, dxm, dyp, dym, dzp, dzm;
    if (i == 0) { dxp = X[p + sx] - X[p]; dxm = dxp; }
    else if (i == d.nx - 1) { dxm = X[p] - X[p - sx]; dxp = dxm; }
    else { dxm = X[p] - X[p - sx]; dxp = X[p + sx] - X[p]; }
    if (j == 0) { dyp = Y[p + sy] - Y[p]; dym = dyp; }
    else if (j == d.ny - 1) { dym = Y[p] - Y[p - sy]; dyp = dym; }
    else { dym = Y[p] - Y[p - sy]; dyp = Y[p + sy] - Y[p]; }
    if (k == 0) { dzp = Z[p + sz] - Z[p]; dzm = dzp; }
    else if (k == d.nz - 1) { dzm = Z[p] - Z[p - sz]; dzp = dzm; }
    else { dzm = Z[p] - Z[p - sz]; dzp = Z[p + sz] - Z[p]; }
    const double dzmax = fmax(dzm, dzp), dymax = fmax(dym, dyp),
                 dxmax = fmax(dxm, dxp);
    const double dzmin = fmin(dzm, dzp), dymin = fmin(dym, dyp),
                 dxmin = fmin(dxm, dxp);
    mymax = fmax(fmax(dxmax, dymax), fmax(dzmax, mymax));
    if (dmins) dmins[p] = fmin(fmin(dzmin, dymin), dxmin);
    const double dxyz = pow(dxmax * dymax * dzmax, 1. / 3.);
    sgd[p] = dxyz;
    sgd[N + p] = pow(dxyz / dxp, 1.0 / 3.0);
    sgd[2 * N + p] = pow(dxyz / dxm, 1.0 / 3.0);
    sgd[3 * N + p] = pow(dxyz / dyp, 1.0 / 3.0);
    sgd[4 * N + p] = pow(dxyz / dym, 1.0 / 3.0);
    sgd[5 * N + p] = pow(dxyz / dzp, 1.0 / 3.0);
    sgd[6 * N + p] = pow(dxyz / dzm, 1.0 / 3.0);
  }
  red[threadIdx.x] = mymax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

// Structure function F2 and the pre-selection eddy viscosity plus the
// vorticity used by SELECTIVO: solver/sgdm_incomp.f90:129-137 (f_structura
// :180-298 and amu0*sqrt(F2)*dxyzsgd) and VORT :330-341.
// Threads own (i,j) columns and march in k over a slab of kz planes.
__global__ void __launch_bounds__(256)
k_sgs_structure(const double* u, const double* dvel, const double* jbn,
                const double* sgd, Dims d, int kz, double* amut, double* w) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const long long N = d.N();
  const long long sx = 1, sy = d.nx, sz = (long long)d.nx * d.ny;
  const double amu0 = 0.1781;
  const int k0 = blockIdx.z * kz, k1 = min(k0 + kz, d.nz);
  const double* U1 = u;
  const double* U2 = u + N;
  const double* U3 = u + 2 * N;
  for (int k = k0; k < k1; k++) {
    const long long p = i * sx + j * sy + k * sz;
    long long a, b;
    double s, du, dv, dw, fac;
#define TERM(A, B, FAC)                                           \
  du = (U1[A] - U1[B]) * (FAC);                                   \
  dv = (U2[A] - U2[B]) * (FAC);                                   \
  dw = (U3[A] - U3[B]) * (FAC);
    // x+
    if (i < d.nx - 1) { a = p + sx; b = p; } else { a = p; b = p - sx; }
    fac = sgd[N + p];
    TERM(a, b, fac);
    s = du * du + dv * dv + dw * dw;
    // x-
    if (i > 0) { a = p - sx; b = p; } else { a = p + sx; b = p; }
    fac = sgd[2 * N + p];
    TERM(a, b, fac);
    s = s + du * du + dv * dv + dw * dw;
    // y+
    if (j < d.ny - 1) { a = p + sy; b = p; } else { a = p; b = p - sy; }
    fac = sgd[3 * N + p];
    TERM(a, b, fac);
    s = s + du * du + dv * dv + dw * dw;
    // y-
    if (j > 0) { a = p - sy; b = p; } else { a = p; b = p + sy; }
    fac = sgd[4 * N + p];
    TERM(a, b, fac);
    s = s + du * du + dv * dv + dw * dw;
    // z+
    if (k < d.nz - 1) { a = p + sz; b = p; } else { a = p; b = p - sz; }
    fac = sgd[5 * N + p];
    TERM(a, b, fac);
    s = s + du * du + dv * dv + dw * dw;
    // z-
    if (k > 0) { a = p - sz; b = p; } else { a = p; b = p + sz; }
    fac = sgd[6 * N + p];
    TERM(a, b, fac);
    s = s + du * du + dv * dv + dw * dw;
#undef TERM
    const double f2 = s / 6.e0;
    amut[p] = amu0 * sqrt(f2) * sgd[p];
    const double j1 = jbn[p], j5 = jbn[4 * N + p], j9 = jbn[8 * N + p];
    w[p] = j5 * dvel[7 * N + p] - j9 * dvel[5 * N + p];
    w[N + p] = j9 * dvel[2 * N + p] - j1 * dvel[6 * N + p];
    w[2 * N + p] = j1 * dvel[3 * N + p] - j5 * dvel[N + p];
  }
}

// SELECTIVO (solver/sgdm_incomp.f90:360-392) and viscosidad
// (solver/viscosidad_incomp.f90:12-20).
__global__ void __launch_bounds__(256)
k_sgs_select(const double* w, Dims d, int kz, double c4, double* amut,
             double* vis, double* vis5) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const long long N = d.N();
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  const double cutoff = cos(20.0 * atan(1.0) / 45.0);
  const int k0 = blockIdx.z * kz, k1 = min(k0 + kz, d.nz);
  const double* wx = w;
  const double* wy = w + N;
  const double* wz = w + 2 * N;
  for (int k = k0; k < k1; k++) {
    const long long p = i + j * sy + k * sz;
    const long long n1 = (i == d.nx - 1) ? p : p + 1;
    const long long n2 = (i == 0) ? p : p - 1;
    const long long n3 = (k == d.nz - 1) ? p : p + sz;
    const long long n4 = (k == 0) ? p : p - sz;
    const long long n5 = (j == d.ny - 1) ? p : p + sy;
    const long long n6 = (j == 0) ? p : p - sy;
    const double wxn = wx[n1] + wx[n2] + wx[n3] + wx[n4] + wx[n5] + wx[n6];
    const double wyn = wy[n1] + wy[n2] + wy[n3] + wy[n4] + wy[n5] + wy[n6];
    const double wzn = wz[n1] + wz[n2] + wz[n3] + wz[n4] + wz[n5] + wz[n6];
    const double a = wx[p], b = wy[p], c = wz[p];
    double wn = a * a + b * b + c * c;
    wn = sqrt(wn * (wxn * wxn + wyn * wyn + wzn * wzn));
    const double s1 = (a * wxn + b * wyn + c * wzn) / wn;
    double m = amut[p];
    if (s1 > cutoff) m = 0.0;
    amut[p] = m;
    if (vis5) vis5[p] = m / 0.7;
    vis[p] = c4 * 1.0 + m;
  }
}

int op_sgs_geometry(emx_ctx* ctx, const double* x, double* sgd,
                    double* dmins, double* blockmax, int nblocks) {
  hipLaunchKernelGGL(k_sgs_geometry, dim3(nblocks), dim3(256), 0, ctx->stream,
                     x, ctx->d, sgd, dmins, blockmax);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_sgdm(emx_ctx* ctx, const double* u, const double* dvel,
            const double* jbn, const double* sgd, double c4, double* amut,
            double* vis, double* vis5) {
  const Dims d = ctx->d;
  const int kz = 8;
  dim3 grid((d.nx + 63) / 64, (d.ny + 3) / 4, (d.nz + kz - 1) / kz);
  dim3 block(64, 4);
  double* w = ctx->sgs_tmp;
  EMX_TIMED(ctx, "sgdm_structure", 184.0 * d.N(),
          hipLaunchKernelGGL(k_sgs_structure, grid, block, 0, ctx->stream, u, dvel,
                     jbn, sgd, d, kz, amut, w));
  EMX_TRY_HIP(hipGetLastError());
  EMX_TIMED(ctx, "sgdm_select", (vis5 ? 56.0 : 48.0) * d.N(),
          hipLaunchKernelGGL(k_sgs_select, grid, block, 0, ctx->stream, w, d, kz, c4,
                     amut, vis, vis5));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

// ---------------------------------------------------------------------------
// Dynamic Smagorinsky eddy viscosity (Germano et al. 1991 identity, Lilly
// 1992 least squares).  The reference has no such model (its sgdm is the
// structure-function model above); this is the target of the learned
// closure.  Physical gradient on the diagonal-metric grid:
// du_a/dx_b = m_b * dvel(3b+a) with m = jbn 1, 5, 9.  Test filter: the box
// filter of ratio 2, weights (1/4, 1/2, 1/4) per direction with clamped
// indices at the walls, summed x innermost, then y, then z (the separable
// passes x, y, z in that order).

__device__ __forceinline__ void sym_strain(const double* dvel, const double* jbn,
                                           long long N, long long p, double S[6],
                                           double* aS) {
  const double m[3] = {jbn[p], jbn[4 * N + p], jbn[8 * N + p]};
  double g[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) g[3 * a + b] = m[b] * dvel[(3 * b + a) * N + p];
  // pairs (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  double s2 = 0.0;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    S[q] = 0.5 * (g[3 * ia[q] + ib[q]] + g[3 * ib[q] + ia[q]]);
    s2 += (q < 3 ? 1.0 : 2.0) * S[q] * S[q];
  }
  *aS = sqrt(2.0 * s2);
}

// Box filter of NQ quantities formed per point by f(p, out[NQ]).
template <int NQ, class F>
__device__ __forceinline__ void box27(const Dims& d, int i, int j, int k, F f, double out[NQ]) {
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  const int xs[3] = {i > 0 ? i - 1 : 0, i, i < d.nx - 1 ? i + 1 : d.nx - 1};
  const int ys[3] = {j > 0 ? j - 1 : 0, j, j < d.ny - 1 ? j + 1 : d.ny - 1};
  const int zs[3] = {k > 0 ? k - 1 : 0, k, k < d.nz - 1 ? k + 1 : d.nz - 1};
  const double w[3] = {0.25, 0.5, 0.25};
  double Z[NQ];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double Y[NQ];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double v0[NQ], v1[NQ], v2[NQ];
      const long long row = ys[b] * sy + zs[c] * sz;
      f(row + xs[0], v0);
      f(row + xs[1], v1);
      f(row + xs[2], v2);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const double X = 0.25 * v0[q] + 0.5 * v1[q] + 0.25 * v2[q];
        Y[q] = b == 0 ? w[0] * X : Y[q] + w[b] * X;
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) Z[q] = c == 0 ? w[0] * Y[q] : Z[q] + w[c] * Y[q];
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) out[q] = Z[q];
}

// Germano-Lilly contractions at each point: with hat() the box filter,
//   L_ij = hat(u_i u_j) - hat(u_i) hat(u_j),
//   M_ij = 2 delta^2 (hat(|S| S_ij) - 4 |hat S| hat S_ij),
//   LM = L:M, MM = M:M (off-diagonal pairs counted twice).
__global__ void __launch_bounds__(256)
k_dynsmag_lm(const double* u, const double* dvel, const double* jbn, const double* delta,
             Dims d, double* LM, double* MM) {
  const long long N = d.N();
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(p % d.nx), j = (int)((p / d.nx) % d.ny),
              k = (int)(p / ((long long)d.nx * d.ny));
    double hS[6], hSS[6], huu[6], hu[3];
    box27<6>(d, i, j, k, [&](long long q, double* v) {
      double a;
      sym_strain(dvel, jbn, N, q, v, &a);
    }, hS);
    box27<6>(d, i, j, k, [&](long long q, double* v) {
      double S[6], a;
      sym_strain(dvel, jbn, N, q, S, &a);
#pragma unroll
      for (int t = 0; t < 6; t++) v[t] = a * S[t];
    }, hSS);
    box27<6>(d, i, j, k, [&](long long q, double* v) {
#pragma unroll
      for (int t = 0; t < 6; t++) v[t] = u[ia[t] * N + q] * u[ib[t] * N + q];
    }, huu);
    box27<3>(d, i, j, k, [&](long long q, double* v) {
#pragma unroll
      for (int t = 0; t < 3; t++) v[t] = u[t * N + q];
    }, hu);
    double s2 = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) s2 += (q < 3 ? 1.0 : 2.0) * hS[q] * hS[q];
    const double ahS = sqrt(2.0 * s2);
    const double d2 = delta[p] * delta[p];
    double lm = 0.0, mm = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const double L = huu[q] - hu[ia[q]] * hu[ib[q]];
      const double M = 2.0 * d2 * (hSS[q] - 4.0 * ahS * hS[q]);
      const double wq = q < 3 ? 1.0 : 2.0;
      lm += wq * L * M;
      mm += wq * M * M;
    }
    LM[p] = lm;
    MM[p] = mm;
  }
}

// nu_t = max(0, C delta^2 |S|) with C = hat(LM) / hat(MM) (0 where
// hat(MM) <= 0).
__global__ void __launch_bounds__(256)
k_dynsmag_nut(const double* LM, const double* MM, const double* dvel, const double* jbn,
              const double* delta, Dims d, double* nut, double c4, double* vis) {
  const long long N = d.N();
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(p % d.nx), j = (int)((p / d.nx) % d.ny),
              k = (int)(p / ((long long)d.nx * d.ny));
    double h[2];
    box27<2>(d, i, j, k, [&](long long q, double* v) {
      v[0] = LM[q];
      v[1] = MM[q];
    }, h);
    double S[6], aS;
    sym_strain(dvel, jbn, N, p, S, &aS);
    const double C = h[1] > 0.0 ? h[0] / h[1] : 0.0;
    const double v = C * delta[p] * delta[p] * aS;
    const double nu = v > 0.0 ? v : 0.0;
    nut[p] = nu;
    if (vis) vis[p] = c4 * 1.0 + nu;  // viscosidad (solver/viscosidad_incomp.f90:15-17)
  }
}

int op_dynsmag(emx_ctx* ctx, const double* u, const double* dvel, const double* jbn,
               const double* delta, double* nut, double c4, double* vis) {
  const Dims d = ctx->d;
  const long long N = d.N();
  double* LM = ctx->sgs_tmp;
  double* MM = ctx->sgs_tmp + N;
  // reads u(3), dvel(9), jbn(3), delta; writes LM, MM / reads LM, MM,
  // dvel(9), jbn(3), delta; writes nut
  EMX_TIMED(ctx, "dynsmag_lm", 18.0 * 8.0 * N,
            hipLaunchKernelGGL(k_dynsmag_lm, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, u, dvel, jbn, delta, d, LM, MM));
  EMX_TIMED(ctx, "dynsmag_nut", 16.0 * 8.0 * N,
            hipLaunchKernelGGL(k_dynsmag_nut, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, LM, MM, dvel, jbn, delta, d, nut, c4, vis));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx
