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

// The projection fused with the xi solves, for plans with xi on the
// materialized kind and the wavefront kind for xi derivatives: one
// workgroup per xi line (persistent over lines).  Its work-items form the
// projected fluxes of their points of the line from the 22 inputs, as
// k_flux_project: the eta/zeta ones (directions in yz, block k of eight
// scratch fields for the k-th materialized direction of mat) go to scratch,
// the eight xi ones to LDS.  Then each of the four wavefronts solves two of
// the eight xi lines with the wavefront kind's solve and writes them as
// the first-pass accumulators (ProbAccum pass 0).  The xi fluxes never go
// to memory.
//
// (Forming the eddy viscosity in this pass instead of reading vis was
// measured slower at 256^3.  The structure-function model: 2.7 ms against
// 1.96 ms for the sgdm kernel and this pass, its stencil needing the
// velocity and vorticity of the four neighbour lines along eta and zeta,
// whose loads could not be kept in flight.  The learned closure: 2.1-2.9 ms
// against 2.2 ms, its matrix-core network and SiLU arithmetic serialised
// with the streaming at the one or two waves per SIMD its registers allow.)
template <int S, int C>
__global__ void __launch_bounds__(256)
k_flux_project_xi(FluxIn F, LineTab T, long long nlines, int nx, int ny, int yz, int mat,
                  double ih, double* fluxbuf, double* acc) {
  constexpr int TS = 64 * C;
  extern __shared__ double lds[];
  // tables | the line's eight xi fluxes.  Wave w solves slots 2w, 2w+1 and
  // uses their rows (2*TS >= 64*(C+1) doubles) as its transpose buffer once
  // it has read them: no other wave touches them before the next line.
  const WaveTabs wt{lds, lds + TS, lds + 2 * TS, lds + 3 * TS, lds + 4 * TS, lds + 5 * TS,
                    lds + 5 * TS + 384};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double* lineF = lds + 5 * TS + 768;  // [8][TS]
  double* xb = lineF + 2 * wv * TS;
  // lines of at most 256 points: one point per work-item, whose next-line
  // inputs are loaded while the current line is solved
  constexpr bool PRE = C <= 4;
  const long long N = F.N;
  const int n = T.n, P = T.P;
  auto line_base = [&](long long Lx) { return (Lx % ny) * nx + (Lx / ny) * (long long)nx * ny; };
  stage_wave_tabs<C>(T, lds);
  FluxPoint Pn;
  if (PRE && blockIdx.x < nlines && threadIdx.x < nx) Pn = FluxPoint(F, line_base(blockIdx.x) + threadIdx.x);
  for (long long L = blockIdx.x; L < nlines; L += gridDim.x) {
    const long long base = line_base(L);
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
      const long long p = base + x;
      const FluxPoint Pt = PRE ? Pn : FluxPoint(F, p);
      const double det = Pt.J(10, p);
#pragma unroll
      for (int dir = 0; dir < 3; dir++) {
        if (dir > 0 && !(yz & (1 << dir))) continue;
        const double met = Pt.J(1 + 4 * dir, p);
        const int k = __builtin_popcount(mat & ((1 << dir) - 1));
#pragma unroll
        for (int slot = 0; slot < 8; slot++) {
          const double f =
              slot < 5 ? conv_flux(Pt, slot + 1, dir, p) : visc_flux(Pt, slot - 3, dir, p);
          const double val = det * (met * f);
          if (dir == 0) lineF[slot * TS + x] = val;
          else fluxbuf[(8 * k + slot) * N + p] = val;
        }
      }
    }
    if (PRE && L + gridDim.x < nlines && threadIdx.x < nx)
      Pn = FluxPoint(F, line_base(L + gridDim.x) + threadIdx.x);
    lds_barrier();  // the line's xi fluxes (and, the first time, the tables)
    double v[C][2];
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      const int idc = idx < n ? idx : n - 1;
#pragma unroll
      for (int a = 0; a < 2; a++) v[q][a] = lineF[(2 * wv + a) * TS + idc];
    }
    wave_solve<S, C, 2>(v, wt, xb, n, P, ih, lane);
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      if (idx < n) {
#pragma unroll
        for (int a = 0; a < 2; a++) acc[(2 * wv + a) * N + base + idx] = v[q][a];
      }
    }
    lds_barrier();  // lineF is rewritten for the next line
  }
}

template <int S, int C>
static hipError_t launch_project_xi_c(emx_ctx* ctx, const FluxIn& in, const LineTab& T, int yz,
                                      int mat, double ih) {
  const size_t lds = (size_t)(5 * 64 * C + 768 + 8 * 64 * C) * sizeof(double);
  auto kern = k_flux_project_xi<S, C>;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  static const int resident = resident_blocks(kern, lds);
  const long long nlines = (long long)ctx->d.ny * ctx->d.nz;
  const long long blocks = nlines < resident ? nlines : resident;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, in, T, nlines,
                     ctx->d.nx, ctx->d.ny, yz, mat, ih, ctx->fluxbuf, ctx->acc);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_project_xi_s(emx_ctx* ctx, const FluxIn& in, const LineTab& T, int yz,
                                      int mat, double ih) {
  switch (T.C) {
    case 1: return launch_project_xi_c<S, 1>(ctx, in, T, yz, mat, ih);
    case 2: return launch_project_xi_c<S, 2>(ctx, in, T, yz, mat, ih);
    case 4: return launch_project_xi_c<S, 4>(ctx, in, T, yz, mat, ih);
    case 8: return launch_project_xi_c<S, 8>(ctx, in, T, yz, mat, ih);
    default: return launch_project_xi_c<S, 16>(ctx, in, T, yz, mat, ih);
  }
}

static int launch_project_xi(emx_ctx* ctx, const FluxIn& in, int mat, double ih) {
  const long long N = in.N;
  const int yz = mat & 6;
  const LineTab& T = ctx->plan[PLAN_DERIV][0].tab;
  const double bytes = 8.0 * N * (project_reads(mat) + 8 * __builtin_popcount(yz) + 8);
  LaunchTimer timer(ctx, "flux_project_xi", 0, bytes);
  hipError_t e;
  switch (ctx->scheme) {
    case EMX_SCHEME_C4: e = launch_project_xi_s<EMX_SCHEME_C4>(ctx, in, T, yz, mat, ih); break;
    case EMX_SCHEME_C6F: e = launch_project_xi_s<EMX_SCHEME_C6F>(ctx, in, T, yz, mat, ih); break;
    default: e = launch_project_xi_s<EMX_SCHEME_C6C>(ctx, in, T, yz, mat, ih); break;
  }
  if (e != hipSuccess) {
    set_error("fused projection and xi solve launch failed: %s", hipGetErrorString(e));
    return EMX_ERR_HIP;
  }
  return EMX_OK;
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
  // xi materialized with the wavefront kind for its solves: the projection
  // and the xi solves in one pass (the xi fluxes stay in LDS)
  const bool fused_xi = (mat & 1) && ctx->plan[PLAN_DERIV][0].kind == KIND_WAVE;
  if (fused_xi) EMX_TRY(launch_project_xi(ctx, in, mat, ih[0]));
  else EMX_TRY(launch_project(ctx, in, mat));
  int pass = fused_xi ? 1 : 0, k = fused_xi ? 1 : 0;
  for (int dir = fused_xi ? 1 : 0; dir < 3; dir++) {
    if (!(mat & (1 << dir))) continue;
    // slots (0,1) (2,3) (4,5) (6,7): outputs rs 1..5 then rsv 2..4.  (Four
    // slots per launch, which reads 1/J once per four fluxes in the last
    // pass, measured 1.6x slower per flux on the 32-lines-per-wave kind:
    // twice the values per work-item.)
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
