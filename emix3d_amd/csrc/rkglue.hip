// The RUNGEK glue around the spatial operator: the SSP-RK3 update, the
// boundary values of the conservative variables, the conservative <->
// primitive conversions and the velocity clip (solver/rungek_ico.f90,
// src/MSVSingle/frontera_{x,y,z}.f90, herramientas_incomp2.f90).  All are
// pointwise or boundary-plane kernels (HBM-bound, no reuse).
#include "ops_common.hpp"

namespace emx {

// SSP-RK3 update of the conservative variables after the flux divergence
// of sub-step `sub` (solver/rungek_ico.f90:87-91, 152-153, 216-219), in the
// reference's evaluation order:
//   1: um1 = um0 + dt*(rs + rsv)
//   2: um1 = (0.75*um0 + 0.25*um) + (0.25*dt)*(rs + rsv)
//   3: um1 = ((1/3)*um0 + (2/3)*um) + ((2/3)*dt)*(rs + rsv)
__global__ void k_rk_update(int sub, double* um1, const double* um0, const double* um,
                           const double* rs, const double* rsv, double dt, long long n) {
  const double third = 1.0 / 3.0, two3 = 2.0 / 3.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const double r = rs[p] + rsv[p];
    double v;
    if (sub == 1) v = um0[p] + dt * r;
    else if (sub == 2) v = 0.75 * um0[p] + 0.25 * um[p] + 0.25 * dt * r;
    else v = third * um0[p] + two3 * um[p] + two3 * dt * r;
    um1[p] = v;
  }
}

int op_rk_update(emx_ctx* ctx, int sub, double* um1, const double* um0, const double* um,
                 const double* rs, const double* rsv, double dt) {
  const long long n = 5 * ctx->d.N();
  EMX_TIMED(ctx, "rk_update", (sub == 1 ? 32.0 : 40.0) * n,
            hipLaunchKernelGGL(k_rk_update, dim3(pointwise_blocks(n)), dim3(256), 0, ctx->stream,
                               sub, um1, um0, um, rs, rsv, dt, n));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

// Boundary values of um1, one work-item per boundary column, statements in
// the Fortran order (so faces that share an edge see the reference's
// sequence).  z faces: frontera_z.f90:219-262 (k = 1 copies k = 2 with
// zero w-momentum; k = nz, nz-1: density 1, the rest copied from nz-2, zero
// w-momentum).
__global__ void k_frontera_z(double* um1, Dims d) {
  const long long nx = d.nx, ny = d.ny, nz = d.nz, N = nx * ny * nz;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  double* q = um1 + c;  // column (i, j); point k at q[k*nx*ny], field m at +m*N
  const long long s = nx * ny;
  auto at = [&](long long k, int m) -> double& { return q[k * s + m * N]; };
  at(0, 0) = at(1, 0);
  at(0, 1) = at(1, 1);
  at(0, 2) = at(1, 2);
  at(0, 3) = 0.0;
  at(0, 4) = at(1, 4);
  at(nz - 1, 0) = 1.0;
  at(nz - 2, 0) = 1.0;
  at(nz - 1, 1) = at(nz - 3, 1);
  at(nz - 2, 1) = at(nz - 3, 1);
  at(nz - 1, 2) = at(nz - 3, 2);
  at(nz - 2, 2) = at(nz - 3, 2);
  at(nz - 1, 3) = 0.0;
  at(nz - 2, 3) = 0.0;
  at(nz - 1, 4) = at(nz - 3, 4);
  at(nz - 2, 4) = at(nz - 3, 4);
}

// y faces (frontera_y.f90:220-247): density and scalar copied from the
// neighbour, momentum zero.
__global__ void k_frontera_y(double* um1, Dims d) {
  const long long nx = d.nx, ny = d.ny, nz = d.nz, N = nx * ny * nz;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * nz) return;
  const long long i = c % nx, k = c / nx;
  double* q = um1 + i + k * nx * ny;  // point j at q[j*nx]
  auto at = [&](long long j, int m) -> double& { return q[j * nx + m * N]; };
  at(0, 0) = at(1, 0);
  at(ny - 1, 0) = at(ny - 2, 0);
  at(0, 1) = 0.0;
  at(ny - 1, 1) = 0.0;
  at(0, 2) = 0.0;
  at(ny - 1, 2) = 0.0;
  at(0, 3) = 0.0;
  at(ny - 1, 3) = 0.0;
  at(0, 4) = at(1, 4);
  at(ny - 1, 4) = at(ny - 2, 4);
}

// x faces (frontera_x.f90:223-256): momentum zero, scalar and density
// copied from the neighbour.
__global__ void k_frontera_x(double* um1, Dims d) {
  const long long nx = d.nx, ny = d.ny, nz = d.nz, N = nx * ny * nz;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ny * nz) return;
  double* q = um1 + c * nx;  // row (j, k); point i at q[i]
  auto at = [&](long long i, int m) -> double& { return q[i + m * N]; };
  at(0, 1) = 0.0;
  at(0, 2) = 0.0;
  at(0, 3) = 0.0;
  at(0, 4) = at(1, 4);
  at(0, 0) = at(1, 0);
  at(nx - 1, 1) = 0.0;
  at(nx - 1, 2) = 0.0;
  at(nx - 1, 3) = 0.0;
  at(nx - 1, 4) = at(nx - 2, 4);
  at(nx - 1, 0) = at(nx - 2, 0);
}

int op_frontera_val(emx_ctx* ctx, double* um1) {
  const Dims& d = ctx->d;
  const long long nz_cols = (long long)d.nx * d.ny, ny_cols = (long long)d.nx * d.nz,
                  nx_rows = (long long)d.ny * d.nz;
  const double N = (double)d.N();
  // bytes: the boundary values written and the neighbours read (5 fields)
  EMX_TIMED(ctx, "frontera_z", 8.0 * 5 * 5 * nz_cols,
            hipLaunchKernelGGL(k_frontera_z, dim3((unsigned)((nz_cols + 255) / 256)), dim3(256),
                               0, ctx->stream, um1, d));
  EMX_TRY_HIP(hipGetLastError());
  EMX_TIMED(ctx, "frontera_y", 8.0 * 5 * 4 * ny_cols,
            hipLaunchKernelGGL(k_frontera_y, dim3((unsigned)((ny_cols + 255) / 256)), dim3(256),
                               0, ctx->stream, um1, d));
  EMX_TRY_HIP(hipGetLastError());
  EMX_TIMED(ctx, "frontera_x", 8.0 * 5 * 4 * nx_rows,
            hipLaunchKernelGGL(k_frontera_x, dim3((unsigned)((nx_rows + 255) / 256)), dim3(256),
                               0, ctx->stream, um1, d));
  EMX_TRY_HIP(hipGetLastError());
  (void)N;
  return EMX_OK;
}

// varc_to_varnc (herramientas_incomp2.f90:2-30) at one point.
__device__ __forceinline__ void varc_point(const double* um1, long long N, long long p,
                                           double mach, double* u, double* pres,
                                           double* temp, double* conc) {
  const double r = um1[p], s = um1[4 * N + p];
  if (conc) conc[p] = s;
  temp[p] = s;
  pres[p] = r / (mach * mach);
#pragma unroll
  for (int c = 0; c < 3; c++) u[c * N + p] = um1[(1 + c) * N + p] / r;
}

__global__ void k_varc_to_varnc(const double* um1, double mach, double* u, double* pres,
                                double* temp, double* conc, long long N) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x)
    varc_point(um1, N, p, mach, u, pres, temp, conc);
}

// varnc_to_varc (herramientas_incomp2.f90:32-58).
__global__ void k_varnc_to_varc(const double* u, const double* pres, const double* temp,
                                double mach, double* um0, double* pres0, long long N) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const double pr = pres[p];
    const double r = pr * mach * mach;
    um0[p] = r;
    um0[4 * N + p] = temp[p];
#pragma unroll
    for (int c = 0; c < 3; c++) um0[(1 + c) * N + p] = u[c * N + p] * r;
    if (pres0) pres0[p] = pr;
  }
}

// The velocity clip of RUNGEK (rungek_ico.f90:102-123): varc_to_varnc, u
// clipped to [-6, 6] (min then max), momentum = u*um1(1)*mask, and
// varc_to_varnc of the result, fused per point.
__global__ void k_rk_clip(double* um1, const double* mask, double mach, double* u, double* pres,
                          double* temp, double* conc, long long N) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const double r = um1[p], s = um1[4 * N + p], mk = mask[p];
    if (conc) conc[p] = s;
    temp[p] = s;
    pres[p] = r / (mach * mach);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double v = um1[(1 + c) * N + p] / r;
      v = v < 6.0 ? v : 6.0;
      v = v > -6.0 ? v : -6.0;
      const double m = v * r * mk;
      um1[(1 + c) * N + p] = m;
      u[c * N + p] = m / r;
    }
  }
}

// um1 rebuilt from the primitive variables after the level-set step
// (rungek_ico.f90:126-135).
__global__ void k_rk_rebuild(double* um1, const double* u, const double* pres, double mach,
                             long long N) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const double r = pres[p] * mach * mach;
    um1[p] = r;
#pragma unroll
    for (int c = 0; c < 3; c++) um1[(1 + c) * N + p] = u[c * N + p] * r;
  }
}

int op_varc_to_varnc(emx_ctx* ctx, const double* um1, double mach, double* u, double* pres,
                     double* temp, double* conc) {
  const long long N = ctx->d.N();
  EMX_TIMED(ctx, "varc_to_varnc", 8.0 * (5 + 5 + (conc ? 1 : 0)) * N,
            hipLaunchKernelGGL(k_varc_to_varnc, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, um1, mach, u, pres, temp, conc, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_varnc_to_varc(emx_ctx* ctx, const double* u, const double* pres, const double* temp,
                     double mach, double* um0, double* pres0) {
  const long long N = ctx->d.N();
  EMX_TIMED(ctx, "varnc_to_varc", 8.0 * (5 + 5 + (pres0 ? 1 : 0)) * N,
            hipLaunchKernelGGL(k_varnc_to_varc, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, u, pres, temp, mach, um0, pres0, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_rk_clip(emx_ctx* ctx, double* um1, const double* mask, double mach, double* u,
               double* pres, double* temp, double* conc) {
  const long long N = ctx->d.N();
  // reads um1 1..5 and mask, writes um1 2..4, u, pres, temp (and conc)
  EMX_TIMED(ctx, "rk_clip", 8.0 * (6 + 8 + (conc ? 1 : 0)) * N,
            hipLaunchKernelGGL(k_rk_clip, dim3(pointwise_blocks(N)), dim3(256), 0, ctx->stream,
                               um1, mask, mach, u, pres, temp, conc, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_rk_rebuild(emx_ctx* ctx, double* um1, const double* u, const double* pres, double mach) {
  const long long N = ctx->d.N();
  EMX_TIMED(ctx, "rk_rebuild", 8.0 * (4 + 4) * N,
            hipLaunchKernelGGL(k_rk_rebuild, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, um1, u, pres, mach, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx
