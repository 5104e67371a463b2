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

// The eight projected fluxes J*(metric_dir*flux) (slots of ProbDivAll:
// convective 1..5, viscous 2..4) of every direction in MASK, the
// arithmetic of ProbDivAll::load.  The inputs are read once per point for
// all directions; block k of eight output fields holds the k-th direction
// of MASK in ascending order.
template <int MASK>
__global__ void k_flux_project(FluxIn F, double* out) {
  const long long N = F.N;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const FluxPoint P(F, p);
    const double det = P.J(10, p);
    int k = 0;
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
      if (!(MASK & (1 << dir))) continue;
      const double met = P.J(1 + 4 * dir, p);
#pragma unroll
      for (int slot = 0; slot < 8; slot++) {
        const double f = slot < 5 ? conv_flux(P, slot + 1, dir, p) : visc_flux(P, slot - 3, dir, p);
        out[(8 * k + slot) * N + p] = det * (met * f);
      }
      k++;
    }
  }
}

// Distinct input fields of the projected fluxes of the directions in mask:
// um(2+dir) and um5, u 1..3, pres, vis, J and the metrics 1/5/9, and the
// dvel components of the viscous fluxes (diagonal 1+4a, off-diagonal pairs
// (0,1): 2,4; (0,2): 3,7; (1,2): 6,8).
static int project_reads(int mask) {
  int um = 1, dv = 0;
  for (int dir = 0; dir < 3; dir++) {
    if (!(mask & (1 << dir))) continue;
    um |= 2 << dir;
    for (int a = 0; a < 3; a++) {
      if (a == dir) { dv |= 1 << (4 * a); continue; }
      const int lo = a < dir ? a : dir, hi = a < dir ? dir : a;
      if (lo == 0 && hi == 1) dv |= (1 << 1) | (1 << 3);
      else if (lo == 0 && hi == 2) dv |= (1 << 2) | (1 << 6);
      else dv |= (1 << 7) | (1 << 5);
    }
  }
  return __builtin_popcount(um) + 3 + 2 + 4 + __builtin_popcount(dv);
}

static int launch_project(emx_ctx* ctx, const FluxIn& in, int mask) {
  const long long N = in.N;
  const double bytes = 8.0 * N * (project_reads(mask) + 8 * __builtin_popcount(mask));
  const dim3 g(pointwise_blocks(N)), b(256);
  switch (mask) {
#define EMX_PROJECT(M)                                                               \
  case M:                                                                            \
    EMX_TIMED(ctx, "flux_project", bytes,                                           \
              hipLaunchKernelGGL(k_flux_project<M>, g, b, 0, ctx->stream, in, ctx->fluxbuf)); \
    break;
    EMX_PROJECT(1) EMX_PROJECT(2) EMX_PROJECT(3) EMX_PROJECT(4) EMX_PROJECT(5)
    EMX_PROJECT(6) EMX_PROJECT(7)
#undef EMX_PROJECT
    default:
      return EMX_OK;
  }
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
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
  // Directions on the materialized kind go first: their projected fluxes
  // come from one pass over the inputs, and each flux then takes one
  // single-field solve on the derivative plan.  The fused directions follow,
  // so that the last pass, which also reads 1/J, reads the inputs and 1/J
  // once for all eight fluxes.  (The three directional terms are summed in
  // pass order.)
  int mat = 0;
  for (int dir = 0; dir < 3; dir++)
    if (ctx->plan[PLAN_FLUX][dir].kind == KIND_MATERIALIZED) mat |= 1 << dir;
  if (__builtin_popcount(mat) > ctx->fluxbuf_dirs) {
    set_error("flux divergence: projected-flux scratch holds %d directions, plan needs %d",
              ctx->fluxbuf_dirs, __builtin_popcount(mat));
    return EMX_ERR_NOMEM;
  }
  EMX_TRY(launch_project(ctx, in, mat));
  int pass = 0, k = 0;
  for (int dir = 0; dir < 3; dir++) {
    if (!(mat & (1 << dir))) continue;
    // slots (0,1) (2,3) (4,5) (6,7): outputs rs 1..5 then rsv 2..4
    for (int slot = 0; slot < 8; slot += 2) {
      double* o0 = slot < 5 ? rs + slot * N : rsv + (slot - 4) * N;
      double* o1 = slot + 1 < 5 ? rs + (slot + 1) * N : rsv + (slot - 3) * N;
      ProbAccum p{ctx->fluxbuf + (8 * k + slot) * N, ctx->acc + slot * N, o0, o1,
                  fp.jbn + 10 * N, N, pass};
      EMX_TRY(launch_lines(ctx, dir, p, ih[dir]));
    }
    k++;
    pass++;
  }
  for (int dir = 0; dir < 3; dir++) {
    if (mat & (1 << dir)) continue;
    const DirPlan& pl = ctx->plan[PLAN_FLUX][dir];
    // one launch of all eight solves on the wavefront kind with short
    // chunks (8*C formed values per work-item), else two launches of four
    if (pl.kind == KIND_WAVE && pl.C <= 4) {
      ProbDivAll<2> p{in, dir, pass, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p, ih[dir]));
    } else {
      ProbDivAll<0> p0{in, dir, pass, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p0, ih[dir]));
      ProbDivAll<1> p1{in, dir, pass, ctx->acc, rs, rsv};
      EMX_TRY(launch_lines(ctx, dir, p1, ih[dir]));
    }
    pass++;
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


}  // namespace emx
