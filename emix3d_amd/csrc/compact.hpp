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
// Also evaluated on the host (with unit vectors) to tabulate B^T for the
// transposed solve.
template <int S>
__host__ __device__ __forceinline__ double rhs_row(const double* w, int l, int n,
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

// Compact-filter RHS of row l (FILTRONPER, src/MSVSingle/filtro_incomp.f90:
// 318-337, constants :140-145): w[d] = u_{l+d}, d in [-5,5]; 0-based rows
// 5..n-6 take the 11-point explicit filter, the five rows at each end u_l.
// cfilt is the filter parameter; sums left to right as the Fortran.
__host__ __device__ __forceinline__ double filter_rhs_row(const double* w, int l, int n,
                                                          double cfilt) {
  if (l < 5 || l > n - 6) return w[0];
  const double arsf = (193. + 126. * cfilt) / 256.;
  const double brsf = (105. + 302. * cfilt) / 256.;
  const double crsf = (-15. + 30. * cfilt) / 64.;
  const double drsf = (45. - 90. * cfilt) / 512.;
  const double ersf = (-5. + 10. * cfilt) / 256.;
  const double frsf = (1. - 2 * cfilt) / 512.;
  return 0.5 * (arsf * (w[0] + w[0]) + brsf * (w[1] + w[-1]) + crsf * (w[2] + w[-2]) +
                drsf * (w[3] + w[-3]) + ersf * (w[4] + w[-4]) + frsf * (w[5] + w[-5]));
}

// Right-hand side of line scheme S at row l; for the filter schemes the
// last argument is cfilt, for the derivatives the inverse spacing.
template <int S>
__host__ __device__ __forceinline__ double line_rhs(const double* w, int l, int n, double p) {
  if constexpr (S == FILTER_NP || S == FILTER_P) return filter_rhs_row(w, l, n, p);
  else return rhs_row<S>(w, l, n, p);
}

// Half-width of the right-hand-side stencil of line scheme S.
template <int S>
struct scheme_halo { static constexpr int value = (S == FILTER_NP || S == FILTER_P) ? 5 : 3; };

// Problem traits: TILE (many inputs per point, formed once), PREFETCH
// (load() is plain loads of few fields, so the wavefront kind may keep the
// next line's loads in flight while it solves the current one).
template <class Prob, class = void>
struct uses_tile { static constexpr bool value = false; };
template <class Prob>
struct uses_tile<Prob, decltype(void(Prob::TILE))> {
  static constexpr bool value = Prob::TILE;
};
template <class Prob, class = void>
struct prefetches { static constexpr bool value = false; };
template <class Prob>
struct prefetches<Prob, decltype(void(Prob::PREFETCH))> {
  static constexpr bool value = Prob::PREFETCH;
};

// BATCH_STORE: the problem stores Q points of a line at once
// (store_batch<Q>), reading what its stores need for all Q first.
template <class Prob, class = void>
struct batch_store { static constexpr bool value = false; };
template <class Prob>
struct batch_store<Prob, decltype(void(Prob::BATCH_STORE))> {
  static constexpr bool value = Prob::BATCH_STORE;
};

// CHUNK_STORE: the problem scatters a work-item's whole chunk at once
// (store_w<C>), used where the chunk is contiguous in memory.
template <class Prob, class = void>
struct chunk_store { static constexpr bool value = false; };
template <class Prob>
struct chunk_store<Prob, decltype(void(Prob::CHUNK_STORE))> {
  static constexpr bool value = Prob::CHUNK_STORE;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Orders the LDS traffic of one wavefront: its writes are visible to its
// own later reads (other lanes included).  The wavefront kind exchanges data
// only inside a wave, so it needs no workgroup barrier per field.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS traffic only: global loads issued
// before it stay in flight (a full __syncthreads also waits for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Line group of workgroup bxp (of gx along the lines) for LG lines per
// wavefront.  A group of 8 doubles (64 bytes) shares each 128-byte cache
// line with its neighbour.  Workgroups go to the eight XCDs round robin,
// each XCD with its own L2, so the two halves of a line go to workgroups
// whose ids differ by 8: the same XCD, dispatched together.
template <int LG>
__device__ __forceinline__ int line_group(int bxp, int gx) {
  if constexpr (LG >= 16) {
    return bxp;
  } else {
    const int base = bxp & ~15;
    if (base + 16 > gx) return bxp;  // the last, incomplete set of 16
    const int r = bxp & 15;
    return base + 2 * (r & 7) + (r >> 3);
  }
}

// Thomas factors and carry multipliers of the wavefront kind, staged in
// LDS in the lanes' chunk order (see lines_wave_body).
struct WaveTabs {
  const double *ib, *m, *gam1, *H, *G, *af, *ab;
};

// Stages the wavefront kind's tables into lds (5*64*C + 768 doubles) in the
// lanes' chunk order.  blob: ib[n] m[n] gamma[n+1] H[n] G[n]; table tb of
// point l at tb*n + (tb >= 2) + l (gamma shifted by one), then Af, Ab.
template <int C>
__device__ __forceinline__ void stage_wave_tabs(const LineTab& T, double* lds) {
  constexpr int TS = 64 * C;
  const int n = T.n;
  for (int t = threadIdx.x; t < 5 * TS + 768; t += blockDim.x) {
    double val;
    if (t < 5 * TS) {
      const int tb = t / TS, r = t % TS, l = (r % 64) * C + r / 64;
      val = l < n ? gref(T.blob + tb * n + (tb >= 2 ? 1 : 0) + l) : 0.0;
    } else {
      val = gref(T.Af + (t - 5 * TS));
    }
    lds[t] = val;
  }
}

// Solves NV fields of one line held by a wavefront in the coalesced order
// (v[q][a]: point q*64 + lane of field a), in place: each field moves to
// the chunk order through the wave's LDS buffer xb, takes the local
// forward and backward sweeps and the carry scans, and moves back.
template <int S, int C, int NV>
__device__ __forceinline__ void wave_solve(double (&v)[C][NV], const WaveTabs& wt, double* xb,
                                           int n, int P, double ih, int lane) {
  constexpr int H = scheme_halo<S>::value;
  constexpr int W = C + 2 * H;
  const int s = lane * C;
  const bool act = lane < P;
#pragma unroll
  for (int a = 0; a < NV; a++) {
    // to the chunk order
    wave_sync();
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      xb[idx + idx / C] = v[q][a];
    }
    wave_sync();
    double w[W];
#pragma unroll
    for (int q = 0; q < W; q++) {
      const int l = clampi(s - H + q, 0, n - 1);
      w[q] = xb[l + l / C];
    }
    // local forward sweep from a zero carry
    double y[C];
    double yp = 0.0;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (act && l < n) {
        const double r = line_rhs<S>(&w[q + H], l, n, ih);
        yp = r * wt.ib[q * 64 + lane] - wt.m[q * 64 + lane] * yp;
        y[q] = yp;
      } else {
        y[q] = 0.0;
      }
    }
    // forward carries: inclusive weighted scan over lanes
    {
      double B = act ? yp : 0.0;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const int d = 1 << k;
        const double up = __shfl_up(B, d, 64);
        if (lane >= d) B += wt.af[k * 64 + lane] * up;
      }
      double Y = __shfl_up(B, 1, 64);
      if (lane == 0) Y = 0.0;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (act && l < n) y[q] += wt.H[q * 64 + lane] * Y;
      }
    }
    // local backward sweep from a zero carry
    double xs = 0.0;
#pragma unroll
    for (int q = C - 1; q >= 0; q--) {
      const int l = s + q;
      if (act && l < n) {
        xs = y[q] - wt.gam1[q * 64 + lane] * xs;
        y[q] = xs;
      }
    }
    // backward carries: weighted scan from the right, then the result
    {
      double B = act ? xs : 0.0;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const int d = 1 << k;
        const double dn = __shfl_down(B, d, 64);
        if (lane + d < 64) B += wt.ab[k * 64 + lane] * dn;
      }
      double X = __shfl_down(B, 1, 64);
      if (lane == 63) X = 0.0;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (act && l < n) y[q] += wt.G[q * 64 + lane] * X;
      }
    }
    // back to the coalesced order
    wave_sync();
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (act && l < n) xb[l + l / C] = y[q];
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      v[q][a] = xb[idx + idx / C];
    }
  }
}

// ---------------------------------------------------------------------------
// One wavefront per line (xi lines, contiguous in memory).  Lines are
// numbered L = j + nyl*k; base(L) = (L % nyl)*sj + (L / nyl)*sk, point l at
// base + l*sl.  Lane c owns the chunk [cC, cC+C).  Every global access is
// coalesced (lane t touches points t + 64q); the NV values of a point are
// formed once, then each field moves to the chunk order through a per-wave
// LDS buffer and is solved on its own (few registers for many fields).
// The Thomas factors are staged in LDS once per workgroup (4 lines), in the
// chunk order of the lanes: the factor of point l = lane*C + q sits at
// q*64 + lane, so the 64 lanes read 64 consecutive doubles (no bank
// conflicts), and the backward factor is stored shifted (entry l holds
// gamma[l+1]).
template <int S, int C, class Prob>
__device__ __forceinline__ void lines_wave_body(const Prob& prob, const LineTab& T,
                                                long long nlines, int nyl, long long sj,
                                                long long sk, long long sl, double ih,
                                                long long blk, long long nblk) {
  constexpr int NV = Prob::NV;
  constexpr bool PRE = prefetches<Prob>::value;
  constexpr int STR = C + 1;  // chunk stride in the transpose buffer
  constexpr int TS = 64 * C;  // per-point table stride (chunk order)
  extern __shared__ double lds[];
  const int n = T.n, P = T.P;
  double* s_ib = lds;
  double* s_m = s_ib + TS;
  double* s_gam1 = s_m + TS;
  double* s_H = s_gam1 + TS;
  double* s_G = s_H + TS;
  double* s_af = s_G + TS;    // Af[384] Ab[384]
  double* s_ab = s_af + 384;
  const WaveTabs wt{s_ib, s_m, s_gam1, s_H, s_G, s_af, s_ab};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double* xb = s_ab + 384 + wv * 64 * STR;
  const long long stride = (long long)nblk * 4;
  long long L = blk * 4 + wv;
  auto line_base = [&](long long Lx) { return (Lx % nyl) * sj + (Lx / nyl) * sk; };
  // coalesced loads of one line (all issued before anything waits on them)
  auto load_line = [&](long long Lx, double (&vv)[C][NV]) {
    const long long bx = line_base(Lx);
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      const int idc = idx < n ? idx : n - 1;
      prob.load(bx + (long long)idc * sl, vv[q]);
    }
  };
  double v[C][NV];
  if (L < nlines) load_line(L, v);
  stage_wave_tabs<C>(T, lds);
  __syncthreads();  // the tables; past here each wave only touches its own xb
  if (L >= nlines) return;

  // The wave's lines L, L + stride, ...; with PRE the next line's loads are
  // in flight while the current one is solved.
  for (;;) {
    const long long Ln = L + stride;
    const bool more = Ln < nlines;
    double vn[C][NV];
    if constexpr (PRE) {
      if (more) load_line(Ln, vn);
    }
    wave_solve<S, C, NV>(v, wt, xb, n, P, ih, lane);
    const long long bs = line_base(L);
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int idx = q * 64 + lane;
      if (idx < n) prob.store(bs + (long long)idx * sl, v[q]);
    }
    if (!more) break;
    L = Ln;
    if constexpr (PRE) {
#pragma unroll
      for (int q = 0; q < C; q++)
#pragma unroll
        for (int a = 0; a < NV; a++) v[q][a] = vn[q][a];
    } else {
      load_line(L, v);
    }
  }
}

template <int S, int C, class Prob>
__global__ void __launch_bounds__(256)
k_lines_wave(Prob prob, LineTab T, long long nlines, int nyl, long long sj,
             long long sk, long long sl, double ih) {
  lines_wave_body<S, C>(prob, T, nlines, nyl, sj, sk, sl, ih, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// LG neighbouring lines per wavefront (lane = line within the group), the
// 64/LG lane groups of a wave and the waves of the workgroup taking
// consecutive chunks of C points along the line.  With LG = 64 every load
// instruction of a wave reads 512 contiguous bytes; LG = 32 / 16 still read
// whole 128-byte lines and let a line be cut into more, shorter chunks
// (fewer registers per work-item).  Line of lane-in-group i and outer index o
// starts at i*si + o*so; point l at +l*sl.  The Thomas factors and the banded
// chunk-carry weights (weights below 1e-24 dropped) are staged in LDS while
// the data loads are in flight; carries travel through LDS.
// The body takes the workgroup's line group bx and outer index by, so that
// the batched kernel (many grids in one launch) can share it.
template <int S, int C, int LG, class Prob>
__device__ __forceinline__ void lines_strided_body(const Prob& prob, const LineTab& T,
                                                   int nlanes, long long si, long long so,
                                                   long long sl, double ih, int bx, int by) {
  constexpr int NV = Prob::NV;
  constexpr int H = scheme_halo<S>::value;
  constexpr int W = C + 2 * H;
  extern __shared__ double lds[];
  const int n = T.n, P = T.P, K = T.K, KB = T.KB;
  const int ntab = 5 * n + 1;
  double* s_ib = lds;
  double* s_m = s_ib + n;
  double* s_gam = s_m + n;
  double* s_H = s_gam + n + 1;
  double* s_G = s_H + n;
  double* s_wf = s_G + n;       // [P*KB]
  double* s_wb = s_wf + P * KB;  // [P*KB]
  double* sh = s_wb + P * KB;   // [NV][P][LG] chunk carries
  const int tid = threadIdx.x;
  const int li = tid % LG;
  const int c = tid / LG;  // chunk index
  const int i = bx * LG + li;
  const bool act = i < nlanes && c < P;
  const int ic = i < nlanes ? i : nlanes - 1;
  const int cc = c < P ? c : P - 1;
  const long long base = (long long)ic * si + (long long)by * so;
  const int s = c * C;

  // window loads, unconditional (clamped addresses), all in flight at once
  double w[NV][W];
#pragma unroll
  for (int q = 0; q < W; q++) {
    const int l = clampi(cc * C - H + q, 0, n - 1);
    double v[NV];
    prob.load(base + (long long)l * sl, v);
#pragma unroll
    for (int a = 0; a < NV; a++) w[a][q] = v[a];
  }
  {
    const int nstage = ntab + 2 * P * KB;
    const int woff = (int)(T.WfB - T.blob);
    for (int t = tid; t < nstage; t += blockDim.x)
      lds[t] = gref(T.blob + (t < ntab ? t : t - ntab + woff));
  }
  __syncthreads();

  double y[NV][C];
  double yp[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) yp[a] = 0.0;
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (act && l < n) {
      const double ibl = s_ib[l], ml = s_m[l];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        const double r = line_rhs<S>(&w[a][q + H], l, n, ih);
        yp[a] = r * ibl - ml * yp[a];
        y[a][q] = yp[a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < NV; a++) y[a][q] = 0.0;
    }
  }
  // work-items past the last chunk (c >= P, in the last wavefront) own no
  // slot
  if (c < P) {
#pragma unroll
    for (int a = 0; a < NV; a++) sh[(a * P + c) * LG + li] = act ? yp[a] : 0.0;
  }
  __syncthreads();
  if (act) {
    const int kk = c < K ? c : K;
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double Y = 0.0;
      for (int k = 0; k < kk; k++)
        Y += s_wf[c * KB + k] * sh[(a * P + c - 1 - k) * LG + li];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) y[a][q] += s_H[l] * Y;
      }
    }
  }
  __syncthreads();

  double xs[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) xs[a] = 0.0;
#pragma unroll
  for (int q = C - 1; q >= 0; q--) {
    const int l = s + q;
    if (act && l < n) {
      const double g1 = s_gam[l + 1];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        xs[a] = y[a][q] - g1 * xs[a];
        y[a][q] = xs[a];
      }
    }
  }
  if (c < P) {
#pragma unroll
    for (int a = 0; a < NV; a++) sh[(a * P + c) * LG + li] = act ? xs[a] : 0.0;
  }
  __syncthreads();
  if (!act) return;  // after the last barrier
  const int kb = P - 1 - c < K ? P - 1 - c : K;
  double X[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double acc = 0.0;
    for (int k = 0; k < kb; k++)
      acc += s_wb[c * KB + k] * sh[(a * P + c + 1 + k) * LG + li];
    X[a] = acc;
  }
  if constexpr (batch_store<Prob>::value && C % 4 == 0) {
    // groups of four points, each stored by one store_batch where the whole
    // group lies on the line
#pragma unroll
    for (int q0 = 0; q0 < C; q0 += 4) {
      if (s + q0 + 4 <= n) {
        double v[4][NV];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const double Gl = s_G[s + q0 + i];
#pragma unroll
          for (int a = 0; a < NV; a++) v[i][a] = y[a][q0 + i] + Gl * X[a];
        }
        prob.template store_batch<4>(base + (long long)(s + q0) * sl, sl, v);
      } else {
#pragma unroll
        for (int q = q0; q < q0 + 4; q++) {
          const int l = s + q;
          if (l < n) {
            const double Gl = s_G[l];
            double v[NV];
#pragma unroll
            for (int a = 0; a < NV; a++) v[a] = y[a][q] + Gl * X[a];
            prob.store(base + (long long)l * sl, v);
          }
        }
      }
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (l < n) {
      const double Gl = s_G[l];
      double v[NV];
#pragma unroll
      for (int a = 0; a < NV; a++) v[a] = y[a][q] + Gl * X[a];
      prob.store(base + (long long)l * sl, v);
    }
  }
}

template <int S, int C, int LG, class Prob>
__global__ void __launch_bounds__(1024)
k_lines_strided(Prob prob, LineTab T, int nlanes, long long si, long long so,
                long long sl, double ih) {
  lines_strided_body<S, C, LG>(prob, T, nlanes, si, so, sl, ih,
                               line_group<LG>(blockIdx.x, gridDim.x), blockIdx.y);
}

// ---------------------------------------------------------------------------
// Many grids in one launch: workgroup b belongs to the grid g with
// blk_start[g] <= b < blk_start[g+1] and solves that grid's line group
// (b - blk_start[g]) % gx, outer index (b - blk_start[g]) / gx.  Each grid
// has its own table (line length), geometry and problem data.
template <class Prob>
struct BatchLines {
  Prob prob;
  LineTab T;
  int nlanes, gx;
  long long si, so, sl;
  double ih;
};

// Workgroup local (of ngroup in its grid) -> line group g, row-major over
// (group along the lines, outer index).  A grid's rows start anywhere in a
// 128-byte cache line (its nx is any size), so neighbouring groups along a
// row share a cache line at every point.  Workgroups go to the eight XCDs
// round robin, each XCD with its own L2: within each set of 8K workgroups
// (K = 16, fewer in the grid's last set), XCD x takes the K consecutive
// groups xK .. xK+K-1, in order, so that all but one of their shared lines
// are fetched once into that XCD's L2 and read there by the neighbour
// dispatched next.  A bijection on [0, ngroup): the last local % 8 ids of
// the grid keep their own group.
__device__ __forceinline__ int xcd_run_group(int local, int ngroup) {
  constexpr int K = 16;
  const int base = local & ~(8 * K - 1);
  const int k = min(K, (ngroup - base) >> 3);  // groups per XCD in this set
  const int r = local - base;
  if (r >= 8 * k) return local;
  return base + (r & 7) * k + (r >> 3);
}

// MAXT: the largest workgroup the launch uses.  1024 serves lines of up to
// 512 points at any shape; 256 (lines of up to 256/LG chunks) lets a
// shape keep its registers instead of spilling to fit 1024 items.
template <int S, int C, int LG, class Prob, int MAXT = 1024>
__global__ void __launch_bounds__(MAXT)
k_lines_batched(const BatchLines<Prob>* grids, const int* blk_start, int ngrid) {
  const int b = blockIdx.x;
  int lo = 0, hi = ngrid;  // blk_start[lo] <= b < blk_start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (blk_start[mid] <= b) lo = mid;
    else hi = mid;
  }
  const BatchLines<Prob>& G = grids[lo];
  const int g = xcd_run_group(b - blk_start[lo], blk_start[lo + 1] - blk_start[lo]);
  lines_strided_body<S, C, LG>(G.prob, G.T, G.nlanes, G.si, G.so, G.sl, G.ih, g % G.gx,
                               g / G.gx);
}

// The wavefront kind for many grids in one launch (xi lines, contiguous in
// memory): workgroup b belongs to grid g as above and is workgroup
// b - blk_start[g] of the nblk that grid's lines are dealt to (its four
// waves take every (4*nblk)-th line).  Each grid has its own chunk length
// (its table's C, at most CMAX; 3 as well as the powers of two, so that
// lines of 129-192 points fill three quarters of a wave's 256 slots at
// least): the branch is uniform over a workgroup.
template <class Prob>
struct BatchWaveLines {
  Prob prob;
  LineTab T;
  long long nlines, sj, sk, sl;
  int nyl, nblk;
  double ih;
};

template <int S, int CMAX, class Prob>
__global__ void __launch_bounds__(256)
k_lines_batched_wave(const BatchWaveLines<Prob>* grids, const int* blk_start, int ngrid) {
  const int b = blockIdx.x;
  int lo = 0, hi = ngrid;  // blk_start[lo] <= b < blk_start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (blk_start[mid] <= b) lo = mid;
    else hi = mid;
  }
  const BatchWaveLines<Prob>& G = grids[lo];
  const Prob& prob = G.prob;
  const int blk = b - blk_start[lo];
  if (CMAX >= 8 && G.T.C == 8)
    lines_wave_body<S, (CMAX >= 8 ? 8 : 1)>(prob, G.T, G.nlines, G.nyl, G.sj, G.sk, G.sl,
                                             G.ih, blk, G.nblk);
  else if (CMAX >= 4 && G.T.C == 4)
    lines_wave_body<S, (CMAX >= 4 ? 4 : 1)>(prob, G.T, G.nlines, G.nyl, G.sj, G.sk, G.sl,
                                             G.ih, blk, G.nblk);
  else if (CMAX >= 4 && G.T.C == 3)
    lines_wave_body<S, (CMAX >= 4 ? 3 : 1)>(prob, G.T, G.nlines, G.nyl, G.sj, G.sk, G.sl,
                                             G.ih, blk, G.nblk);
  else if (CMAX >= 2 && G.T.C == 2)
    lines_wave_body<S, (CMAX >= 2 ? 2 : 1)>(prob, G.T, G.nlines, G.nyl, G.sj, G.sk, G.sl,
                                             G.ih, blk, G.nblk);
  else
    lines_wave_body<S, 1>(prob, G.T, G.nlines, G.nyl, G.sj, G.sk, G.sl, G.ih, blk, G.nblk);
}

// ---------------------------------------------------------------------------
// The strided shape for problems whose point values are expensive to form
// (many input fields, e.g. the flux divergence): each work-item loads and
// forms the NV values of its own C points only, then the fields are solved
// one after another, the halo window of each read from an LDS tile
// [n][LG] of the field.  Same launch geometry as k_lines_strided.
template <int S, int C, int LG, class Prob>
__global__ void __launch_bounds__(1024)
k_lines_tile(Prob prob, LineTab T, int nlanes, long long si, long long so,
             long long sl, double ih) {
  constexpr int NV = Prob::NV;
  constexpr int W = C + 6;
  constexpr int PMAX = 1024 / LG;
  extern __shared__ double lds[];
  const int n = T.n, P = T.P, K = T.K, KB = T.KB;
  const int ntab = 5 * n + 1;
  double* s_ib = lds;
  double* s_m = s_ib + n;
  double* s_gam = s_m + n;
  double* s_H = s_gam + n + 1;
  double* s_G = s_H + n;
  double* s_wf = s_G + n;        // [P*KB]
  double* s_wb = s_wf + P * KB;  // [P*KB]
  double* sh = s_wb + P * KB;    // [PMAX][LG] carries
  double* tile = sh + PMAX * LG; // [n][LG]
  const int tid = threadIdx.x;
  const int li = tid % LG;
  const int c = tid / LG;
  const int i = line_group<LG>(blockIdx.x, gridDim.x) * LG + li;
  const bool act = i < nlanes && c < P;
  const int ic = i < nlanes ? i : nlanes - 1;
  const long long base = (long long)ic * si + (long long)blockIdx.y * so;
  const int s = c * C;

  double v[C][NV];
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q < n ? s + q : n - 1;
    prob.load(base + (long long)l * sl, v[q]);
  }
  {
    const int nstage = ntab + 2 * P * KB;
    const int woff = (int)(T.WfB - T.blob);
    for (int t = tid; t < nstage; t += blockDim.x)
      lds[t] = gref(T.blob + (t < ntab ? t : t - ntab + woff));
  }
#pragma unroll
  for (int a = 0; a < NV; a++) {
    __syncthreads();
    if (c < P) {
#pragma unroll
      for (int q = 0; q < C; q++)
        if (s + q < n) tile[(s + q) * LG + li] = v[q][a];
    }
    __syncthreads();
    double w[W];
#pragma unroll
    for (int q = 0; q < W; q++) w[q] = tile[clampi(s - 3 + q, 0, n - 1) * LG + li];
    double y[C];
    double yp = 0.0;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (act && l < n) {
        const double r = rhs_row<S>(&w[q + 3], l, n, ih);
        yp = r * s_ib[l] - s_m[l] * yp;
        y[q] = yp;
      } else {
        y[q] = 0.0;
      }
    }
    sh[c * LG + li] = act ? yp : 0.0;
    __syncthreads();
    if (act) {
      const int kk = c < K ? c : K;
      double Y = 0.0;
      for (int k = 0; k < kk; k++) Y += s_wf[c * KB + k] * sh[(c - 1 - k) * LG + li];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) y[q] += s_H[l] * Y;
      }
    }
    double xs = 0.0;
#pragma unroll
    for (int q = C - 1; q >= 0; q--) {
      const int l = s + q;
      if (act && l < n) {
        xs = y[q] - s_gam[l + 1] * xs;
        y[q] = xs;
      }
    }
    __syncthreads();
    sh[c * LG + li] = act ? xs : 0.0;
    __syncthreads();
    if (act) {
      const int kb = P - 1 - c < K ? P - 1 - c : K;
      double X = 0.0;
      for (int k = 0; k < kb; k++) X += s_wb[c * KB + k] * sh[(c + 1 + k) * LG + li];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) v[q][a] = y[q] + s_G[l] * X;
      }
    }
  }
  if (!act) return;  // after the last barrier
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (l < n) prob.store(base + (long long)l * sl, v[q]);
  }
}

// ---------------------------------------------------------------------------
// Transposed compact derivative along lines: out = (ih A^{-1} B)^T g
// = ih B^T A^{-T} g.  T is a TAB_TRANSPOSED table: the Thomas factors and
// chunk-carry weights of A^T, and B^T as seven diagonals.  Each work-item
// solves its chunk of A^T y = g (carries as in k_lines_tile), the solution
// goes to an LDS tile [n][LG] of the field, and B^T is applied from the
// tile.  The NV fields of a point are solved one after another.  Launch
// geometry of k_lines_tile.
// CHUNK (xi lines, contiguous in memory; problems with CHUNK_STORE): the
// scatter of a work-item's chunk goes through store_w<C> (k_lines_adj_xi).
template <int C, int LG, class Prob, bool CHUNK>
__device__ __forceinline__ void lines_adj_body(const Prob& prob, const LineTab& T, int nlanes,
                                               long long si, long long so, long long sl,
                                               double ih) {
  constexpr int NV = Prob::NV;
  constexpr int PMAX = 1024 / LG;
  extern __shared__ double lds[];
  const int n = T.n, P = T.P, K = T.K, KB = T.KB;
  const int ntab = 5 * n + 1;
  double* s_ib = lds;
  double* s_m = s_ib + n;
  double* s_gam = s_m + n;
  double* s_H = s_gam + n + 1;
  double* s_G = s_H + n;
  double* s_wf = s_G + n;         // [P*KB]
  double* s_wb = s_wf + P * KB;   // [P*KB]
  double* s_bt = s_wb + P * KB;   // [7n]
  double* sh = s_bt + 7 * n;      // [PMAX][LG] carries
  double* tile = sh + PMAX * LG;  // [n][LG]
  const int tid = threadIdx.x;
  const int li = tid % LG;
  const int c = tid / LG;
  const int i = blockIdx.x * LG + li;
  const bool act = i < nlanes && c < P;
  const int ic = i < nlanes ? i : nlanes - 1;
  const long long base = (long long)ic * si + (long long)blockIdx.y * so;
  const int s = c * C;

  double v[C][NV];
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q < n ? s + q : n - 1;
    prob.load(base + (long long)l * sl, v[q]);
  }
  {
    const int nstage = ntab + 2 * P * KB;
    const int woff = (int)(T.WfB - T.blob);
    for (int t = tid; t < nstage; t += blockDim.x)
      lds[t] = gref(T.blob + (t < ntab ? t : t - ntab + woff));
    for (int t = tid; t < 7 * n; t += blockDim.x) s_bt[t] = T.BT[t];
  }
#pragma unroll
  for (int a = 0; a < NV; a++) {
    __syncthreads();
    double y[C];
    double yp = 0.0;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int l = s + q;
      if (act && l < n) {
        yp = v[q][a] * s_ib[l] - s_m[l] * yp;
        y[q] = yp;
      } else {
        y[q] = 0.0;
      }
    }
    sh[c * LG + li] = act ? yp : 0.0;
    __syncthreads();
    if (act) {
      const int kk = c < K ? c : K;
      double Y = 0.0;
      for (int k = 0; k < kk; k++) Y += s_wf[c * KB + k] * sh[(c - 1 - k) * LG + li];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) y[q] += s_H[l] * Y;
      }
    }
    double xs = 0.0;
#pragma unroll
    for (int q = C - 1; q >= 0; q--) {
      const int l = s + q;
      if (act && l < n) {
        xs = y[q] - s_gam[l + 1] * xs;
        y[q] = xs;
      }
    }
    __syncthreads();
    sh[c * LG + li] = act ? xs : 0.0;
    __syncthreads();
    if (act) {
      const int kb = P - 1 - c < K ? P - 1 - c : K;
      double X = 0.0;
      for (int k = 0; k < kb; k++) X += s_wb[c * KB + k] * sh[(c + 1 + k) * LG + li];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) tile[l * LG + li] = y[q] + s_G[l] * X;
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int l = s + q;
        if (l < n) {
          double r = 0.0;
#pragma unroll
          for (int k = 0; k < 7; k++) {
            const int t = l + k - 3;
            if (t >= 0 && t < n) r += s_bt[7 * l + k] * tile[t * LG + li];
          }
          v[q][a] = ih * r;
        }
      }
    }
  }
  if (!act) return;  // after the last barrier
  if constexpr (CHUNK) {
    if (s + C <= n) {  // the chunk's C points are adjacent (unit stride)
      prob.template store_w<C>(base + s, v);
      return;
    }
  }
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int l = s + q;
    if (l < n) prob.store(base + (long long)l * sl, v[q]);
  }
}

template <int C, int LG, class Prob>
__global__ void __launch_bounds__(1024)
k_lines_adj(Prob prob, LineTab T, int nlanes, long long si, long long so, long long sl,
            double ih) {
  lines_adj_body<C, LG, Prob, false>(prob, T, nlanes, si, so, sl, ih);
}

// The xi form: at most 256 work-items per workgroup and three waves per
// SIMD, which leaves the chunk scatter the registers of its C points.
template <int C, int LG, class Prob>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_lines_adj_xi(Prob prob, LineTab T, int nlanes, long long si, long long so, long long sl,
               double ih) {
  lines_adj_body<C, LG, Prob, true>(prob, T, nlanes, si, so, sl, ih);
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

// Workgroups of kernel kern (256 work-items, lds bytes) the device holds at
// once, over all its compute units.
template <class K>
int resident_blocks(K kern, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 1 << 30;  // unknown: one workgroup per four lines
  }
  return (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
}

template <int S, int C, class Prob>
hipError_t launch_wave_c(const Prob& prob, const LineTab& T, const DirGeom& g,
                         double ih, hipStream_t st) {
  // one resident set of workgroups; each wave takes every (4*blocks)-th line
  const size_t lds = (size_t)(5 * 64 * C + 768 + 4 * 64 * (C + 1)) * sizeof(double);
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_lines_wave<S, C, Prob>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  static const int resident = resident_blocks(k_lines_wave<S, C, Prob>, lds);
  long long blocks = (g.nlines + 3) / 4;
  if (blocks > resident) blocks = resident;
  hipLaunchKernelGGL((k_lines_wave<S, C, Prob>), dim3((unsigned)blocks),
                     dim3(256), lds, st, prob, T, g.nlines, g.nyl, g.sj, g.sk,
                     g.sl, ih);
  return hipGetLastError();
}

template <int S, int C, int LG, class Prob>
hipError_t launch_strided_c(const Prob& prob, const LineTab& T,
                            const DirGeom& g, double ih, hipStream_t st) {
  dim3 grid((g.nlanes + LG - 1) / LG, g.nouter);
  const int threads = ((T.P * LG + 63) / 64) * 64;
  if constexpr (uses_tile<Prob>::value) {
    const size_t lds = (size_t)(5 * T.n + 1 + 2 * T.P * T.KB + (1024 / LG) * LG +
                                T.n * LG) * sizeof(double);
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_lines_tile<S, C, LG, Prob>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_lines_tile<S, C, LG, Prob>), grid, dim3(threads), lds,
                       st, prob, T, g.nlanes, g.si, g.so, g.sl, ih);
    return hipGetLastError();
  } else {
  const size_t lds = (size_t)(5 * T.n + 1 + 2 * T.P * T.KB +
                              Prob::NV * T.P * LG) * sizeof(double);
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_lines_strided<S, C, LG, Prob>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_lines_strided<S, C, LG, Prob>), grid, dim3(threads), lds,
                     st, prob, T, g.nlanes, g.si, g.so, g.sl, ih);
  return hipGetLastError();
  }
}

template <int S, class Prob>
hipError_t launch_lines_s(const Prob& prob, const LineTab& T, const DirGeom& g,
                          double ih, hipStream_t st) {
  if constexpr (uses_tile<Prob>::value) {
    // many-input problems: the wavefront kind, or the tile form of the
    // 32/16-lines-per-wave kinds with short chunks (the formed values of a
    // work-item's chunk must fit its registers)
    switch (T.kind) {
      case KIND_WAVE:
        switch (T.C) {
          case 1: return launch_wave_c<S, 1>(prob, T, g, ih, st);
          case 2: return launch_wave_c<S, 2>(prob, T, g, ih, st);
          case 4: return launch_wave_c<S, 4>(prob, T, g, ih, st);
          case 8: return launch_wave_c<S, 8>(prob, T, g, ih, st);
          case 16: return launch_wave_c<S, 16>(prob, T, g, ih, st);
        }
        break;
      case KIND_STRIDED32:
        if (T.C == 4) return launch_strided_c<S, 4, 32>(prob, T, g, ih, st);
        if (T.C == 8) return launch_strided_c<S, 8, 32>(prob, T, g, ih, st);
        break;
      case KIND_STRIDED16:
        if (T.C == 4) return launch_strided_c<S, 4, 16>(prob, T, g, ih, st);
        if (T.C == 8) return launch_strided_c<S, 8, 16>(prob, T, g, ih, st);
        break;
      case KIND_STRIDED8:
        if (T.C == 2) return launch_strided_c<S, 2, 8>(prob, T, g, ih, st);
        if (T.C == 4) return launch_strided_c<S, 4, 8>(prob, T, g, ih, st);
        break;
    }
    return hipErrorInvalidValue;
  } else if constexpr (Prob::NV > 4) {
    // more than four solves per work-item only fit the wavefront kind
    if (T.kind != KIND_WAVE) return hipErrorInvalidValue;
    switch (T.C) {
      case 1: return launch_wave_c<S, 1>(prob, T, g, ih, st);
      case 2: return launch_wave_c<S, 2>(prob, T, g, ih, st);
      case 4: return launch_wave_c<S, 4>(prob, T, g, ih, st);
      case 8: return launch_wave_c<S, 8>(prob, T, g, ih, st);
    }
    return hipErrorInvalidValue;
  } else {
    switch (T.kind) {
      case KIND_WAVE:
        switch (T.C) {
          case 1: return launch_wave_c<S, 1>(prob, T, g, ih, st);
          case 2: return launch_wave_c<S, 2>(prob, T, g, ih, st);
          case 4: return launch_wave_c<S, 4>(prob, T, g, ih, st);
          case 8: return launch_wave_c<S, 8>(prob, T, g, ih, st);
          case 16: return launch_wave_c<S, 16>(prob, T, g, ih, st);
        }
        break;
      case KIND_STRIDED:
        switch (T.C) {
          case 8: return launch_strided_c<S, 8, 64>(prob, T, g, ih, st);
          case 16: return launch_strided_c<S, 16, 64>(prob, T, g, ih, st);
          case 32: return launch_strided_c<S, 32, 64>(prob, T, g, ih, st);
        }
        break;
      case KIND_STRIDED32:
        switch (T.C) {
          case 4: return launch_strided_c<S, 4, 32>(prob, T, g, ih, st);
          case 8: return launch_strided_c<S, 8, 32>(prob, T, g, ih, st);
          case 16: return launch_strided_c<S, 16, 32>(prob, T, g, ih, st);
        }
        break;
      case KIND_STRIDED16:
        switch (T.C) {
          case 4: return launch_strided_c<S, 4, 16>(prob, T, g, ih, st);
          case 8: return launch_strided_c<S, 8, 16>(prob, T, g, ih, st);
        }
        break;
      case KIND_STRIDED8:
        switch (T.C) {
          case 4: return launch_strided_c<S, 4, 8>(prob, T, g, ih, st);
          case 8: return launch_strided_c<S, 8, 8>(prob, T, g, ih, st);
        }
        break;
    }
    return hipErrorInvalidValue;
  }
}

// Solve all lines of direction dir with the context plan.
template <class Prob>
int launch_lines(emx_ctx* ctx, int dir, const Prob& prob, double ih) {
  const DirGeom g = dir_geom(ctx->d, dir);
  const int op = uses_tile<Prob>::value ? PLAN_FLUX : PLAN_DERIV;
  const LineTab& T = ctx->plan[op][dir].tab;
  LaunchTimer timer(ctx, Prob::label(), dir, prob.bytes_pt() * (double)ctx->d.N());
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

// Longest line of the transposed solve (its LDS tile holds a whole line of
// 16 lanes).
constexpr int ADJ_MAX_N = 512;

// Chunk length of the transposed solve (16 lines per wavefront).
inline int adj_chunk(int n) { return n <= 256 ? 4 : 8; }

// Solve all lines of direction dir with the transposed operator; T is the
// context's TAB_TRANSPOSED table of that direction.
template <class Prob>
int launch_lines_adj(emx_ctx* ctx, int dir, const Prob& prob, double ih) {
  const DirGeom g = dir_geom(ctx->d, dir);
  const LineTab& T = ctx->adj[dir];
  LaunchTimer timer(ctx, Prob::label(), dir, prob.bytes_pt() * (double)ctx->d.N());
  auto go = [&](auto kern, int LG) {
    dim3 grid((g.nlanes + LG - 1) / LG, g.nouter);
    const int threads = ((T.P * LG + 63) / 64) * 64;
    const size_t lds = (size_t)(5 * T.n + 1 + 2 * T.P * T.KB + 7 * T.n + (1024 / LG) * LG +
                                T.n * LG) * sizeof(double);
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)kern,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, grid, dim3(threads), lds, ctx->stream, prob, T, g.nlanes,
                       g.si, g.so, g.sl, ih);
  };
  // xi lines of a chunk-scattering problem with chunks of four (lines of up
  // to 256): as many lines per workgroup (4, 8 or 16) as keep it within
  // 256 items, so the chunk scatter has the registers of such a workgroup
  // (with chunks of eight it does not fit them); otherwise 16 lines per
  // workgroup and the point-by-point scatter
  bool chunked = false;
  if constexpr (chunk_store<Prob>::value) {
    if (g.sl == 1 && T.C == 4 && T.P <= 64) {
      const int lg = T.P <= 16 ? 16 : (T.P <= 32 ? 8 : 4);
      if (lg == 16) go(k_lines_adj_xi<4, 16, Prob>, 16);
      else if (lg == 8) go(k_lines_adj_xi<4, 8, Prob>, 8);
      else go(k_lines_adj_xi<4, 4, Prob>, 4);
      chunked = true;
    }
  }
  if (!chunked) {
    if (T.C == 4) go(k_lines_adj<4, 16, Prob>, 16);
    else go(k_lines_adj<8, 16, Prob>, 16);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("transposed line-solver launch (dir %d, C %d) failed: %s", dir, T.C,
              hipGetErrorString(e));
    return EMX_ERR_HIP;
  }
  return EMX_OK;
}

// Valid chunk sizes per kind.
inline bool valid_chunk(int kind, int C) {
  switch (kind) {
    case KIND_WAVE: return C == 1 || C == 2 || C == 4 || C == 8 || C == 16;
    case KIND_STRIDED: return C == 8 || C == 16 || C == 32;
    case KIND_STRIDED32: return C == 4 || C == 8 || C == 16;
    case KIND_STRIDED16: return C == 4 || C == 8;
    case KIND_STRIDED8: return C == 2 || C == 4 || C == 8;
  }
  return false;
}

// Most chunks a line may have under each kind (64 lanes of one wave, or
// 1024 work-items of a workgroup divided by the lines per wave).
inline int max_chunks(int kind) {
  switch (kind) {
    case KIND_WAVE: return 64;
    case KIND_STRIDED: return 16;
    case KIND_STRIDED32: return 32;
    case KIND_STRIDED16: return 64;
    case KIND_STRIDED8: return 128;
  }
  return 0;
}

// Default plans.  Derivatives: xi lines one wave per line with the smallest
// C that fits 64 lanes; eta/zeta lines 16 per wave, C=8 (lines up to 512
// points), else 64 per wave with the smallest C giving <= 16 waves.  Flux
// divergence: the same xi shape; eta/zeta the materialized form (projected
// fluxes into scratch, then single-field solves on the derivative plan).
inline void default_plan(int op, int dir, int n, int* kind, int* C) {
  if (op == PLAN_FLUX) {
    // the materialized form (one projection pass for all three directions,
    // then single-field solves) measures fastest on MI355X
    *kind = KIND_MATERIALIZED;
    *C = 0;
  } else if (dir == 0) {
    *kind = KIND_WAVE;
    int c = 1;
    while (c * 64 < n) c *= 2;
    *C = c;
  } else if (n >= 128 && n <= 512) {
    // 32 lines per wavefront, 16-point chunks: the fastest single-field
    // solve at 256 points per line with a working set beyond the L3
    *kind = KIND_STRIDED32;
    *C = 16;
  } else if (n <= 512) {
    *kind = KIND_STRIDED16;
    *C = 8;
  } else {
    *kind = KIND_STRIDED;
    int c = 8;
    while (c * 16 < n && c < 32) c *= 2;
    *C = c;
  }
}

// Shapes each plan may take.
inline bool valid_plan(int op, int kind, int C) {
  if (kind == KIND_MATERIALIZED) return op == PLAN_FLUX && C == 0;
  if (!valid_chunk(kind, C)) return false;
  if (op == PLAN_FLUX) {
    if (kind == KIND_WAVE) return C <= 16;
    if (kind == KIND_STRIDED8) return C <= 4;
    return kind != KIND_STRIDED && C <= 8;
  }
  return kind != KIND_STRIDED8 || C >= 4;
}

}  // namespace emx
