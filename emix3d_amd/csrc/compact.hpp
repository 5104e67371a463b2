// Device side of the compact first derivative: the right-hand side rows of
// each scheme and two line-solver kernels built on the chunked Thomas
// algorithm of LineTab (emx_internal.hpp).
//
// Kernel shapes (CDNA4, 64-wide wavefronts):
//  * k_lines_wave    -- one wavefront per line; lane c owns C consecutive
//    points.  Used for xi lines, which are contiguous in memory: a wave
//    reads one line with unit-stride loads.  Chunk carries travel through
//    six shuffle steps of a weighted prefix scan.
//  * k_lines_strided -- 64 neighbouring lines (consecutive i) per
//    workgroup, the line split over up to 16 waves, C points per lane.
//    Used for eta/zeta lines: every load instruction of a wave reads 64
//    consecutive doubles (512 B), so the strided directions need no
//    transpose (the reference transposes twice per eta/zeta derivative,
//    src/Schemes/derivCompact.cu:327-335,389-398).  Carries travel through
//    LDS as a dot product with precomputed weights.
// Both read every input once from HBM (the 3-point halo of a chunk comes
// from cache) and write every output once.
//
// A "problem" object P supplies NV values per point (load) and consumes NV
// derivatives per point (store), so flux formation, metric projection and
// the accumulation of the divergence are fused into the solve.
#pragma once
#include "emx_internal.hpp"

namespace emx {

// RHS of row l from w[d] = f_{l+d}, d in [-3,3], times the inverse spacing.
//  C4 : solver/derivada.f90:104-109 (consrs: :83-89)
//  C6F: src/MSVSingle/derivada_incomp.f90:178-185 (consrs: :81-86)
//  C6C: src/Schemes/derivCompact.cu:90-140
template <int S>
__device__ __forceinline__ double rhs_row(const double* w, int l, int n,
                                          double ih) {
  if constexpr (S == EMX_SCHEME_C4) {
    const double ars = 3. / 4., ars1 = -17. / 6., brs1 = 3. / 2., crs1 = 1.5;
    if (l == 0)
      return ih * (ars1 * w[0] + brs1 * w[1] + crs1 * w[2] - 1. / 6. * w[3]);
    if (l == n - 1)
      return ih * (-ars1 * w[0] - brs1 * w[-1] - crs1 * w[-2] +
                   1. / 6. * w[-3]);
    return ih * (ars * (w[1] - w[-1]));
  } else if constexpr (S == EMX_SCHEME_C6F) {
    const double ars = 7. / 9., brs = 1. / 36., ars1 = -2.5, brs1 = 2.,
                 crs1 = 0.5, ars2 = 0.75;
    if (l == 0) return ih * (ars1 * w[0] + brs1 * w[1] + crs1 * w[2]);
    if (l == n - 1) return ih * (-ars1 * w[0] - brs1 * w[-1] - crs1 * w[-2]);
    if (l == 1 || l == n - 2) return ih * (ars2 * (w[1] - w[-1]));
    return ih * (ars * (w[1] - w[-1]) + brs * (w[2] - w[-2]));
  } else {
    const double ars = 14.0 / (9.0 * 2.0), brs = 1.0 / (9.0 * 4.0),
                 ars1 = -17.0 / 6.0, brs1 = 3.0 / 2.0, crs1 = 3.0 / 2.0,
                 drs1 = -1.0 / 6.0, ars2 = -3.0 / 4.0, brs2 = 0.0,
                 crs2 = 3.0 / 4.0, drs2 = 0.0;
    if (l == 0)
      return ih * (ars1 * w[0] + brs1 * w[1] + crs1 * w[2] + drs1 * w[3]);
    if (l == 1)
      return ih * (ars2 * w[-1] + brs2 * w[0] + crs2 * w[1] + drs2 * w[2]);
    if (l == n - 1)
      return ih * (-ars1 * w[0] - brs1 * w[-1] - crs1 * w[-2] - drs1 * w[-3]);
    if (l == n - 2)
      return ih * (-ars2 * w[1] - brs2 * w[0] - crs2 * w[-1] - drs2 * w[-2]);
    return ih * (ars * (w[1] - w[-1]) + (brs * (w[2] - w[-2])));
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// ---------------------------------------------------------------------------
// One wavefront per line.  Lines are numbered L = j + ny*k; the line's
// first point is at L*line_base_stride... expressed through (nyl, sj, sk):
// base(L) = (L % nyl)*sj + (L / nyl)*sk, point l at base + l*sl.
template <int S, int C, class Prob>
__global__ void __launch_bounds__(256)
k_lines_wave(Prob prob, LineTab T, long long nlines, int nyl, long long sj,
             long long sk, long long sl, double ih) {
  constexpr int NV = Prob::NV;
  constexpr int W = C + 6;
  const int lane = threadIdx.x & 63;
  const long long L = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (L >= nlines) return;  // wave-uniform: no block barrier follows
  const long long base = (L % nyl) * sj + (L / nyl) * sk;
  const int n = T.n;
  const int s = lane * C;
  const bool act = lane < T.P;

  double w[NV][W];
  if constexpr (C >= 3) {
    // own points from memory, the 3-point halos from the neighbour lanes
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      double v[NV];
      if (act && l < n) {
        prob.load(base + (long long)l * sl, v);
      } else {
#pragma unroll
        for (int a = 0; a < NV; a++) v[a] = 0.0;
      }
#pragma unroll
      for (int a = 0; a < NV; a++) w[a][q + 3] = v[a];
    }
#pragma unroll
    for (int a = 0; a < NV; a++) {
#pragma unroll
      for (int h = 0; h < 3; h++) {
        w[a][h] = __shfl_up(w[a][C + h], 1, 64);
        w[a][C + 3 + h] = __shfl_down(w[a][3 + h], 1, 64);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < W; q++) {
      const int l = clampi(s - 3 + q, 0, n - 1);
      double v[NV];
      if (act) {
        prob.load(base + (long long)l * sl, v);
      } else {
#pragma unroll
        for (int a = 0; a < NV; a++) v[a] = 0.0;
      }
#pragma unroll
      for (int a = 0; a < NV; a++) w[a][q] = v[a];
    }
  }

  // local forward sweep from a zero carry
  double y[NV][C];
  double yp[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) yp[a] = 0.0;
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (act && l < n) {
      const double ibl = T.ib[l], ml = T.m[l];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        const double r = rhs_row<S>(&w[a][q + 3], l, n, ih);
        yp[a] = r * ibl - ml * yp[a];
        y[a][q] = yp[a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < NV; a++) y[a][q] = 0.0;
    }
  }
  // forward carries: inclusive weighted scan over lanes
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double B = act ? yp[a] : 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int d = 1 << k;
      const double up = __shfl_up(B, d, 64);
      if (lane >= d) B += T.Af[k * 64 + lane] * up;
    }
    double Y = __shfl_up(B, 1, 64);
    if (lane == 0) Y = 0.0;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (act && l < n) y[a][q] += T.H[l] * Y;
    }
  }
  // local backward sweep from a zero carry
  double xs[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) xs[a] = 0.0;
#pragma unroll
  for (int q = C - 1; q >= 0; q--) {
    const int l = s + q;
    if (act && l < n) {
      const double g1 = T.gam[l + 1];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        xs[a] = y[a][q] - g1 * xs[a];
        y[a][q] = xs[a];
      }
    }
  }
  // backward carries: weighted scan from the right
  double X[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double B = act ? xs[a] : 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int d = 1 << k;
      const double dn = __shfl_down(B, d, 64);
      if (lane + d < 64) B += T.Ab[k * 64 + lane] * dn;
    }
    X[a] = __shfl_down(B, 1, 64);
    if (lane == 63) X[a] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (act && l < n) {
      const double Gl = T.G[l];
      double v[NV];
#pragma unroll
      for (int a = 0; a < NV; a++) v[a] = y[a][q] + Gl * X[a];
      prob.store(base + (long long)l * sl, v);
    }
  }
}

// ---------------------------------------------------------------------------
// 64 lines per workgroup (lane = line), P waves along the line.
// Line of lane i and outer index o starts at i*si + o*so; point l at +l*sl.
template <int S, int C, class Prob>
__global__ void __launch_bounds__(1024)
k_lines_strided(Prob prob, LineTab T, int nlanes, long long si, long long so,
                long long sl, double ih) {
  constexpr int NV = Prob::NV;
  constexpr int W = C + 6;
  __shared__ double sh[NV][16][64];
  const int lane = threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int n = T.n, P = T.P;
  const int i = blockIdx.x * 64 + lane;
  const bool act = i < nlanes;
  const long long base = (long long)i * si + (long long)blockIdx.y * so;
  const int s = c * C;

  double w[NV][W];
#pragma unroll
  for (int q = 0; q < W; q++) {
    const int l = clampi(s - 3 + q, 0, n - 1);
    double v[NV];
    if (act) {
      prob.load(base + (long long)l * sl, v);
    } else {
#pragma unroll
      for (int a = 0; a < NV; a++) v[a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < NV; a++) w[a][q] = v[a];
  }

  double y[NV][C];
  double yp[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) yp[a] = 0.0;
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (l < n) {
      const double ibl = T.ib[l], ml = T.m[l];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        const double r = rhs_row<S>(&w[a][q + 3], l, n, ih);
        yp[a] = r * ibl - ml * yp[a];
        y[a][q] = yp[a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < NV; a++) y[a][q] = 0.0;
    }
  }
#pragma unroll
  for (int a = 0; a < NV; a++) sh[a][c][lane] = yp[a];
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double Y = 0.0;
    for (int t = 0; t < c; t++) Y += T.Wf[c * P + t] * sh[a][t][lane];
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (l < n) y[a][q] += T.H[l] * Y;
    }
  }
  __syncthreads();

  double xs[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) xs[a] = 0.0;
#pragma unroll
  for (int q = C - 1; q >= 0; q--) {
    const int l = s + q;
    if (l < n) {
      const double g1 = T.gam[l + 1];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        xs[a] = y[a][q] - g1 * xs[a];
        y[a][q] = xs[a];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NV; a++) sh[a][c][lane] = xs[a];
  __syncthreads();
  double X[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double acc = 0.0;
    for (int t = c + 1; t < P; t++) acc += T.Wb[c * P + t] * sh[a][t][lane];
    X[a] = acc;
  }
  if (!act) return;  // after the last barrier
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (l < n) {
      const double Gl = T.G[l];
      double v[NV];
#pragma unroll
      for (int a = 0; a < NV; a++) v[a] = y[a][q] + Gl * X[a];
      prob.store(base + (long long)l * sl, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Host-side dispatch over (scheme, kind, C) for one direction of a grid.

// Geometry of direction dir of an nx*ny*nz grid.
struct DirGeom {
  int n;                 // line length
  // wave kind
  long long nlines; int nyl; long long sj, sk;
  // strided kind
  int nlanes; int nouter; long long si, so;
  long long sl;          // stride along the line
};

inline DirGeom dir_geom(const Dims& d, int dir) {
  const long long nx = d.nx, ny = d.ny, nz = d.nz;
  DirGeom g{};
  if (dir == 0) {
    g.n = d.nx; g.sl = 1;
    g.nlines = ny * nz; g.nyl = d.ny; g.sj = nx; g.sk = nx * ny;
    g.nlanes = d.ny; g.nouter = d.nz; g.si = nx; g.so = nx * ny;
  } else if (dir == 1) {
    g.n = d.ny; g.sl = nx;
    g.nlines = nx * nz; g.nyl = d.nx; g.sj = 1; g.sk = nx * ny;
    g.nlanes = d.nx; g.nouter = d.nz; g.si = 1; g.so = nx * ny;
  } else {
    g.n = d.nz; g.sl = nx * ny;
    g.nlines = nx * ny; g.nyl = d.nx; g.sj = 1; g.sk = nx;
    g.nlanes = d.nx; g.nouter = d.ny; g.si = 1; g.so = nx;
  }
  return g;
}

template <int S, int C, class Prob>
hipError_t launch_wave_c(const Prob& prob, const LineTab& T, const DirGeom& g,
                         double ih, hipStream_t st) {
  const long long blocks = (g.nlines + 3) / 4;
  hipLaunchKernelGGL((k_lines_wave<S, C, Prob>), dim3((unsigned)blocks),
                     dim3(256), 0, st, prob, T, g.nlines, g.nyl, g.sj, g.sk,
                     g.sl, ih);
  return hipGetLastError();
}

template <int S, int C, class Prob>
hipError_t launch_strided_c(const Prob& prob, const LineTab& T,
                            const DirGeom& g, double ih, hipStream_t st) {
  dim3 grid((g.nlanes + 63) / 64, g.nouter);
  dim3 block(64, T.P);
  hipLaunchKernelGGL((k_lines_strided<S, C, Prob>), grid, block, 0, st, prob,
                     T, g.nlanes, g.si, g.so, g.sl, ih);
  return hipGetLastError();
}

template <int S, class Prob>
hipError_t launch_lines_s(const Prob& prob, const LineTab& T, const DirGeom& g,
                          double ih, hipStream_t st) {
  if (T.kind == KIND_WAVE) {
    switch (T.C) {
      case 1: return launch_wave_c<S, 1>(prob, T, g, ih, st);
      case 2: return launch_wave_c<S, 2>(prob, T, g, ih, st);
      case 4: return launch_wave_c<S, 4>(prob, T, g, ih, st);
      case 8: return launch_wave_c<S, 8>(prob, T, g, ih, st);
      case 16: return launch_wave_c<S, 16>(prob, T, g, ih, st);
    }
  } else {
    switch (T.C) {
      case 8: return launch_strided_c<S, 8>(prob, T, g, ih, st);
      case 16: return launch_strided_c<S, 16>(prob, T, g, ih, st);
      case 32: return launch_strided_c<S, 32>(prob, T, g, ih, st);
    }
  }
  return hipErrorInvalidValue;
}

// Solve all lines of direction dir with the context plan.
template <class Prob>
int launch_lines(emx_ctx* ctx, int dir, const Prob& prob, double ih) {
  const DirGeom g = dir_geom(ctx->d, dir);
  const LineTab& T = ctx->plan[dir].tab;
  hipError_t e;
  switch (ctx->scheme) {
    case EMX_SCHEME_C4: e = launch_lines_s<EMX_SCHEME_C4>(prob, T, g, ih, ctx->stream); break;
    case EMX_SCHEME_C6F: e = launch_lines_s<EMX_SCHEME_C6F>(prob, T, g, ih, ctx->stream); break;
    default: e = launch_lines_s<EMX_SCHEME_C6C>(prob, T, g, ih, ctx->stream); break;
  }
  if (e != hipSuccess) {
    set_error("line-solver launch (dir %d, kind %d, C %d) failed: %s", dir,
              T.kind, T.C, hipGetErrorString(e));
    return EMX_ERR_HIP;
  }
  return EMX_OK;
}

// Valid chunk sizes per kind.
inline bool valid_chunk(int kind, int C) {
  if (kind == KIND_WAVE) return C == 1 || C == 2 || C == 4 || C == 8 || C == 16;
  return C == 8 || C == 16 || C == 32;
}

// Default plan: xi lines one wave per line with the smallest C that fits
// 64 lanes; eta/zeta lines strided with the smallest C giving <= 16 waves.
inline void default_plan(int dir, int n, int* kind, int* C) {
  if (dir == 0) {
    *kind = KIND_WAVE;
    int c = 1;
    while (c * 64 < n) c *= 2;
    *C = c;
  } else {
    *kind = KIND_STRIDED;
    int c = 8;
    while (c * 16 < n && c < 32) c *= 2;
    *C = c;
  }
}

}  // namespace emx
