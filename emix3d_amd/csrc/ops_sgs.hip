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

}  // namespace emx
