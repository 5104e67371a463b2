// Fluxes (flujos), flux divergence (divergencia) and their fused forms.
#include "ops_common.hpp"

namespace emx {

__global__ void k_fill(double* a, double v, long long n) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x)
    a[p] = v;
}

__global__ void k_flujos(FluxIn F, int neq, double* flux) {
  const long long N = F.N;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    double e[6];
#pragma unroll
    for (int d = 0; d < 3; d++) e[d] = conv_flux(F, neq, d, p);
    if (neq >= 2 && neq <= 4) {
#pragma unroll
      for (int d = 0; d < 3; d++) e[3 + d] = visc_flux(F, neq, d, p);
    } else {
      e[3] = e[4] = e[5] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 6; q++) flux[q * N + p] = e[q];
  }
}

// The eight projected fluxes J*(metric_dir*flux) of direction dir (slots
// of ProbDivAll: convective 1..5, viscous 2..4), the arithmetic of
// ProbDivAll::load.
__global__ void k_flux_project(FluxIn F, int dir, double* out) {
  const long long N = F.N;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const double det = F.jbn[9 * N + p];
    const double met = F.jbn[(4 * dir) * N + p];
#pragma unroll
    for (int slot = 0; slot < 8; slot++) {
      const double f = slot < 5 ? conv_flux(F, slot + 1, dir, p) : visc_flux(F, slot - 3, dir, p);
      out[slot * N + p] = det * (met * f);
    }
  }
}

int op_flux(emx_ctx* ctx, int neq, const FluxPtrs& fp, double* flux) {
  FluxIn F{fp.um, fp.u, fp.pres, fp.vis, fp.jbn, fp.dvel, ctx->d.N()};
  // inputs: um 2..4 [neq 1]; um 2..4, u(neq-1), pres, vis, jbn 1/5/9 and
  // five dvel [neq 2..4]; u 1..3, um5 [neq 5]; outputs: 6 fields
  EMX_TIMED(ctx, "flujos", 8.0 * (6 + (neq == 1 ? 3 : (neq == 5 ? 4 : 14))) * F.N,
          hipLaunchKernelGGL(k_flujos, dim3(pointwise_blocks(F.N)), dim3(256), 0,
                     ctx->stream, F, neq, flux));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

template <class Src, bool VISC>
static int div_passes(emx_ctx* ctx, const Src& src, const double* jbn,
                      const double ih[3], double* rs, double* rsv) {
  const long long N = ctx->d.N();
  ProbDiv<Src, VISC> p;
  p.src = src;
  p.jbn = jbn;
  p.N = N;
  p.acc = ctx->acc;
  p.accv = ctx->acc + N;
  p.rs = rs;
  p.rsv = rsv;
  for (int dir = 0; dir < 3; dir++) {
    p.dir = dir;
    p.pass = dir;
    EMX_TRY(launch_lines(ctx, dir, p, ih[dir]));
  }
  return EMX_OK;
}

int op_divergence(emx_ctx* ctx, const double* flux, const double* jbn,
                  const double ih[3], double* rs, double* rsv) {
  FluxArr src{flux, ctx->d.N()};
  return div_passes<FluxArr, true>(ctx, src, jbn, ih, rs, rsv);
}

int op_flux_div(emx_ctx* ctx, int neq, const FluxPtrs& fp, const double ih[3],
                double* rs, double* rsv) {
  const long long N = ctx->d.N();
  FluxLive src{{fp.um, fp.u, fp.pres, fp.vis, fp.jbn, fp.dvel, N}, neq};
  if (neq >= 2 && neq <= 4)
    return div_passes<FluxLive, true>(ctx, src, fp.jbn, ih, rs, rsv);
  EMX_TRY((div_passes<FluxLive, false>(ctx, src, fp.jbn, ih, rs, rsv)));
  // the viscous fluxes of the continuity and scalar equations are zero
  // (flujos_compact.f90:32-40,187-214), so their divergence is zero
  EMX_TIMED(ctx, "fill", 8.0 * N,
          hipLaunchKernelGGL(k_fill, dim3(pointwise_blocks(N)), dim3(256), 0,
                     ctx->stream, rsv, 0.0, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_flux_div_all(emx_ctx* ctx, const FluxPtrs& fp, const double ih[3],
                    double* rs, double* rsv) {
  const long long N = ctx->d.N();
  const FluxIn in{fp.um, fp.u, fp.pres, fp.vis, fp.jbn, fp.dvel, N};
  for (int dir = 0; dir < 3; dir++) {
    const DirPlan& pl = ctx->plan[PLAN_FLUX][dir];
    if (pl.kind == KIND_MATERIALIZED) {
      // projected fluxes into scratch (16 inputs read, 8 written), then one
      // single-field solve per flux on the derivative plan
      EMX_TIMED(ctx, "flux_project", 24.0 * 8.0 * N,
                hipLaunchKernelGGL(k_flux_project, dim3(pointwise_blocks(N)), dim3(256), 0,
                                   ctx->stream, in, dir, ctx->fluxbuf));
      EMX_TRY_HIP(hipGetLastError());
      for (int slot = 0; slot < 8; slot++) {
        double* out = slot < 5 ? rs + slot * N : rsv + (slot - 4) * N;
        ProbAccum p{ctx->fluxbuf + slot * N, ctx->acc + slot * N, out, fp.jbn + 10 * N, dir};
        EMX_TRY(launch_lines(ctx, dir, p, ih[dir]));
      }
      continue;
    }
    // one launch of all eight solves on the wavefront kind with short
    // chunks (8*C formed values per work-item), else two launches of four
    if (pl.kind == KIND_WAVE && pl.C <= 4) {
      ProbDivAll<2> p{in, dir, dir, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p, ih[dir]));
    } else {
      ProbDivAll<0> p0{in, dir, dir, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p0, ih[dir]));
      ProbDivAll<1> p1{in, dir, dir, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p1, ih[dir]));
    }
  }
  // zero viscous divergence of the continuity and scalar equations
  EMX_TIMED(ctx, "fill", 8.0 * N,
          hipLaunchKernelGGL(k_fill, dim3(pointwise_blocks(N)), dim3(256), 0,
                     ctx->stream, rsv, 0.0, N));
  EMX_TIMED(ctx, "fill", 8.0 * N,
          hipLaunchKernelGGL(k_fill, dim3(pointwise_blocks(N)), dim3(256), 0,
                     ctx->stream, rsv + 4 * N, 0.0, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

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

}  // namespace emx
