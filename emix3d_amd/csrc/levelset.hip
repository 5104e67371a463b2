// Level-set kernels: pressure extrapolation into the solid (the path the
// reference's advectLevelSet really runs) and WENO5 / SSP-RK3 advection of
// phi (cuAdvectLsJB).  Both work on the 5-cell mirror-padded domain that
// advect_CUDA builds (src/LevelSet/lsTransportEc.cu:39-166), with the
// reference's per-point boundary substitutions on the outermost padded
// layers restated exactly.  Unlike the reference (floor-divided grids,
// lsTransportEc.cu:58-79), every point is computed, for any size.
#include <cmath>

#include "ops.hpp"

namespace emx {

namespace {

constexpr int GC = 5;  // gcells, lsTransportEc.cu:39

struct PDims {
  int x, y, z;
  __host__ __device__ long long N() const { return (long long)x * y * z; }
};

__device__ __forceinline__ int mirror(int i, int n) {
  // padded index -> interior index of the even mirror used by
  // cuGhostCellsMirror3d{X,Y,Z} (lsTools.cu:482-585)
  int s = i < GC ? 2 * GC - 1 - i : (i >= n + GC ? 2 * (n + GC) - 1 - i : i);
  return s - GC;
}

// cuSwapToGhost + the three mirror passes, as one gather.
__global__ void k_pad(const double* in, int nx, int ny, int nz, double* out,
                      PDims g) {
  const long long NG = g.N();
  for (long long p = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; p < NG;
       p += (long long)gridDim.x * blockDim.x) {
    const unsigned up = (unsigned)p, row = up / (unsigned)g.x;  // NG < 2^32
    const int i = (int)(up - row * (unsigned)g.x), j = (int)(row % (unsigned)g.y),
              k = (int)(row / (unsigned)g.y);
    const int is = mirror(i, nx), js = mirror(j, ny), ks = mirror(k, nz);
    out[p] = in[(long long)is + (long long)nx * ((long long)js + (long long)ny * ks)];
  }
}

// cuSwapFromGhost (lsTools.cu:798-822).
__global__ void k_unpad(const double* in, PDims g, int nx, int ny, int nz,
                        double* out) {
  const long long N = (long long)nx * ny * nz;
  for (long long p = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const unsigned up = (unsigned)p, row = up / (unsigned)nx;
    const int i = (int)(up - row * (unsigned)nx), j = (int)(row % (unsigned)ny),
              k = (int)(row / (unsigned)ny);
    out[p] = in[(long long)(i + GC) + (long long)g.x * ((long long)(j + GC) + (long long)g.y * (k + GC))];
  }
}

// DevFirstOrder_LS (src/LevelSet/extrapol.cu:105-153) at padded point id:
// the metric-scaled central difference of S along each axis, one-sided with
// factor 1 for all three when the point lies on any face of the padded grid.
__device__ __forceinline__ void ls_grad1_point(const double* S, const double m[3], PDims g,
                                               double iX, double iY, double iZ, long long id,
                                               int x, int y, int z, double dphi[3]) {
  const long long sxy = (long long)g.x * g.y;
  long long ip = id + 1, im = id - 1, jp = id + g.x, jm = id - g.x, kp = id + sxy,
            km = id - sxy;
  double factor = 0.5;
  if (x == 0) { im = id; factor = 1.0; }
  if (y == 0) { jm = id; factor = 1.0; }
  if (z == 0) { km = id; factor = 1.0; }
  if (x == g.x - 1) { ip = id; factor = 1.0; }
  if (y == g.y - 1) { jp = id; factor = 1.0; }
  if (z == g.z - 1) { kp = id; factor = 1.0; }
  dphi[0] = factor * iX * m[0] * (S[ip] - S[im]);
  dphi[1] = factor * iY * m[1] * (S[jp] - S[jm]);
  dphi[2] = factor * iZ * m[2] * (S[kp] - S[km]);
}

// extrapolKernel (extrapol.cu:5-103, extFlag>0) at one solid point (phiS >
// 0) of the padded grid: the upwind metric-scaled one-sided differences of
// ext along the sign of the solid's gradient.  Elsewhere the reference's rs
// is exactly zero (so = 0 and finite operands).  geo: the point's metrics
// 1, 5, 9 and gradient of the solid (k_ls_solid_geom).
__device__ __forceinline__ double ls_extrap_rs(long long id, const double* ext,
                                               const double* geo, PDims g, double iX,
                                               double iY, double iZ) {
  const long long sxy = (long long)g.x * g.y;
  const unsigned uid = (unsigned)id, row = uid / (unsigned)g.x;  // NG < 2^32
  const int x = (int)(uid - row * (unsigned)g.x), y = (int)(row % (unsigned)g.y),
            z = (int)(row / (unsigned)g.y);
  long long ip = id + 1, im = id - 1, jp = id + g.x, jm = id - g.x, kp = id + sxy,
            km = id - sxy;
  long long id2 = id;
  if (x == 0) { id2 = ip; im = id; }
  if (y == 0) { id2 = jp; jm = id; }
  if (z == 0) { id2 = kp; km = id; }
  if (x == g.x - 1) { id2 = im; ip = id; }
  if (y == g.y - 1) { id2 = jm; jp = id; }
  if (z == g.z - 1) { id2 = km; kp = id; }
  const double so = 1.0;
  const double m1 = geo[0], m5 = geo[1], m9 = geo[2];
  const double pdx = so * geo[3], pdy = so * geo[4], pdz = so * geo[5];
  const double e2 = ext[id2];
  const double xd = iX * m1 * (e2 - ext[im]);
  const double yd = iY * m5 * (e2 - ext[jm]);
  const double zd = iZ * m9 * (e2 - ext[km]);
  const double xu = iX * m1 * (ext[ip] - e2);
  const double yu = iY * m5 * (ext[jp] - e2);
  const double zu = iZ * m9 * (ext[kp] - e2);
  double s = (double)(pdx > 0.0) - (double)(pdx < 0.0);
  double r = (0.5 * (s + 1.0) * xd + 0.5 * fabs(s - 1.0) * xu) * pdx;
  s = (double)(pdy > 0.0) - (double)(pdy < 0.0);
  r += (0.5 * (s + 1.0) * yd + 0.5 * fabs(s - 1.0) * yu) * pdy;
  s = (double)(pdz > 0.0) - (double)(pdz < 0.0);
  r += (0.5 * (s + 1.0) * zd + 0.5 * fabs(s - 1.0) * zu) * pdz;
  return r;
}

// Ordered list of the solid points (S > 0) of the padded grid: per-chunk
// counts, one exclusive scan, then each chunk writes its indices in order
// (wavefront ballots, no atomics).
constexpr int LS_CHUNK = 4096;

__global__ void k_ls_solid_count(const double* S, long long NG, int* bcount) {
  __shared__ int wsum[4];
  const long long base = (long long)blockIdx.x * LS_CHUNK;
  int c = 0;
  for (int t = threadIdx.x; t < LS_CHUNK; t += 256) {
    const long long id = base + t;
    c += (id < NG && S[id] > 0.0) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Exclusive scan of nb chunk counts in place (one workgroup of 1024); the
// total into *total.
__global__ void __launch_bounds__(1024) k_ls_scan(int* bcount, int nb, int* total) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per < nb ? t * per : nb;
  const int b1 = b0 + per < nb ? b0 + per : nb;
  int s = 0;
  for (int b = b0; b < b1; b++) s += bcount[b];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int b = b0; b < b1; b++) {
    const int c = bcount[b];
    bcount[b] = run;
    run += c;
  }
  if (t == 1023) *total = part[1023];
}

__global__ void k_ls_solid_write(const double* S, long long NG, const int* boff, int* list) {
  __shared__ int wsum[4];
  const long long base = (long long)blockIdx.x * LS_CHUNK;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int run = boff[blockIdx.x];
  for (int r = 0; r < LS_CHUNK / 256; r++) {
    const long long id = base + r * 256 + threadIdx.x;
    const bool f = id < NG && S[id] > 0.0;
    const unsigned long long m = __ballot(f);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(m);
    __syncthreads();
    int off = run;
    for (int w = 0; w < wv; w++) off += wsum[w];
    if (f) list[off + pre] = (int)id;
    run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

__global__ void k_copy2(const double* a, double* b, double* c, long long n) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const double v = a[p];
    b[p] = v;
    c[p] = v;
  }
}

// One SSP-RK3 stage of the extrapolation (RunGK_{First,Second,Third}S,
// extrapol.cu:155-212) at the listed solid points:
//   stage 0: out = ext - dt*rs
//   stage 1: out = 0.75*P + 0.25*(ext - dt*rs)
//   stage 2: out = (P + 2*(ext - dt*rs))/3      (out may alias P)
// The fluid points keep P in all three buffers: their rs is zero, and the
// stages only re-round them (0.75*P + 0.25*P, (P + 2*P)/3), which the
// reference does and this build does not.
// The metrics 1, 5, 9 (read from the interior grid through the mirror map
// of the padding) and DevFirstOrder_LS's gradient of S at each listed solid
// point, six doubles per entry: the only points where the pseudo-stages use
// them.
__global__ void k_ls_solid_geom(const int* list, const int* count, const double* S,
                                const double* jbn, int nx, int ny, int nz, PDims g, double iX,
                                double iY, double iZ, double* geo) {
  const int n = *count;
  const long long N = (long long)nx * ny * nz;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const long long id = list[t];
    const unsigned uid = (unsigned)id, row = uid / (unsigned)g.x;  // NG < 2^32
    const int x = (int)(uid - row * (unsigned)g.x), y = (int)(row % (unsigned)g.y),
              z = (int)(row / (unsigned)g.y);
    const long long q = mirror(x, nx) + (long long)nx * (mirror(y, ny) + (long long)ny * mirror(z, nz));
    const double m[3] = {jbn[q], jbn[4 * N + q], jbn[8 * N + q]};
    double dphi[3];
    ls_grad1_point(S, m, g, iX, iY, iZ, id, x, y, z, dphi);
    double* e = geo + 6 * (long long)t;
    e[0] = m[0];
    e[1] = m[1];
    e[2] = m[2];
    e[3] = dphi[0];
    e[4] = dphi[1];
    e[5] = dphi[2];
  }
}

__global__ void k_ls_extrap_list(int stage, const int* list, const int* count, const double* ext,
                                 const double* P, const double* geo, PDims g, double iX,
                                 double iY, double iZ, double dt, double* out) {
  const int n = *count;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const long long id = list[t];
    const double rs = ls_extrap_rs(id, ext, geo + 6 * (long long)t, g, iX, iY, iZ);
    const double ev = ext[id];
    if (stage == 0) out[id] = ev - dt * rs;
    else if (stage == 1) out[id] = 0.75 * P[id] + 0.25 * (ev - dt * rs);
    else out[id] = (P[id] + 2.0 * (ev - dt * rs)) / 3.0;
  }
}

// 1/x for x in the normal range: the hardware reciprocal refined by two
// Newton steps (each doubles the correct bits), within an ulp of the
// correctly rounded quotient.
__device__ __forceinline__ double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// Phi_x_WENO (src/LevelSet/lsTools.cu:9-50).  The level-set kernels let
// the compiler fuse multiplies and adds (the library is built without
// contraction elsewhere): they are issue-bound, and the fused forms change
// results in the last bits only.
__device__ __forceinline__ double weno5(double b1, double b2, double b3,
                                        double b4, double b5) {
#pragma clang fp contract(fast)
  const double s1 = (13.0 / 12.0) * (b1 - 2.0 * b2 + b3) * (b1 - 2.0 * b2 + b3) +
                    (0.25) * (b1 - 4.0 * b2 + 3.0 * b3) * (b1 - 4.0 * b2 + 3.0 * b3);
  const double s2 = (13.0 / 12.0) * (b2 - 2.0 * b3 + b4) * (b2 - 2.0 * b3 + b4) +
                    (0.25) * (b2 - b4) * (b2 - b4);
  const double s3 = (13.0 / 12.0) * (b3 - 2.0 * b4 + b5) * (b3 - 2.0 * b4 + b5) +
                    (0.25) * (3.0 * b3 - 4.0 * b4 + b5) * (3.0 * b3 - 4.0 * b4 + b5);
  // weights a_k = c_k/d_k, d_k = (s_k + 1e-6)^2, normalised: o_k = c_k p_k /
  // sum_j c_j p_j with p_k the product of the other two d (one division
  // instead of four; d_k lies in [1e-12, ~1e16], so the products stay far
  // from the double range's ends)
  const double d1 = (s1 + 1.0e-6) * (s1 + 1.0e-6);
  const double d2 = (s2 + 1.0e-6) * (s2 + 1.0e-6);
  const double d3 = (s3 + 1.0e-6) * (s3 + 1.0e-6);
  const double w1 = 0.1 * (d2 * d3), w2 = 0.6 * (d1 * d3), w3 = 0.3 * (d1 * d2);
  const double isum = recip_nr(w1 + w2 + w3);
  const double o1 = w1 * isum;
  const double o2 = w2 * isum;
  const double o3 = w3 * isum;
  return ((o1 * (2.0 * b1 - 7.0 * b2 + 11.0 * b3) +
           o2 * (-1.0 * b2 + 5.0 * b3 + 2.0 * b4) +
           o3 * (2.0 * b3 + 5.0 * b4 - b5)) *
          (1.0 / 6.0));
}

// Backward difference of phi along axis `ax` at point (x,y,z), with the
// index-0 rule of Dev1thO_Downwind (lsTools.cu:53-95).
__device__ __forceinline__ double bdiff(const double* phi, PDims g, int ax,
                                        int x, int y, int z, double ih) {
  const long long sxy = (long long)g.x * g.y;
  const long long id = (long long)x + (long long)g.x * y + sxy * z;
  if (ax == 0) {
    if (x == 0) return ih * (gref(phi + id + 1) - gref(phi + id));
    return ih * (gref(phi + id) - gref(phi + id - 1));
  }
  if (ax == 1) {
    if (y == 0) return ih * (gref(phi + id + g.x) - gref(phi + id));
    return ih * (gref(phi + id) - gref(phi + id - g.x));
  }
  if (z == 0) return ih * (gref(phi + id + sxy) - gref(phi + id));
  return ih * (gref(phi + id) - gref(phi + id - sxy));
}

// The backward differences (Dev1thO_Downwind, lsTools.cu:53-95) times the
// metric m along axis c at offsets -2..+3 from padded point uid (axis
// position q of n, index stride `stride`).  phi is read once at each of the
// seven positions q-3..q+3, clamped into the grid (a clamped position's
// difference is never used by the stencils below); at position 0 the
// difference is the forward one, phi(1) - phi(0), which is the backward
// difference at position 1.  Indices are 32-bit: the padded grid has fewer
// than 2^32 points, so the wrapped unsigned sum is the exact index.
__device__ __forceinline__ void axis_bd(const double* src, unsigned stride, int q, int n,
                                        unsigned uid, double ihc, double m, double bd[6]) {
#pragma clang fp contract(fast)
  double s[7];
#pragma unroll
  for (int o = 0; o < 7; o++) {
    int t = q - 3 + o;
    t = t < 0 ? 0 : (t > n - 1 ? n - 1 : t);
    s[o] = gref(src + (uid + (unsigned)(t - q) * stride));
  }
#pragma unroll
  for (int o = 0; o < 6; o++) bd[o] = ihc * (s[o + 1] - s[o]) * m;
  if (q == 0) bd[2] = bd[3];
  if (q == 1) bd[1] = bd[2];
  if (q == 2) bd[0] = bd[1];
}

// axis_bd for an axis position q with 3 <= q <= n - 4: every stencil
// position lies inside the grid, so no clamps and no edge rules.
__device__ __forceinline__ void axis_bd_inner(const double* src, unsigned stride,
                                              unsigned uid, double ihc, double m,
                                              double bd[6]) {
#pragma clang fp contract(fast)
  // signed offsets from the point itself: on the x axis (stride 1) they
  // become the loads' immediate offsets
  const double* c0 = src + uid;
  const int st = (int)stride;
  double s[7];
#pragma unroll
  for (int o = 0; o < 7; o++) s[o] = gref(c0 + (o - 3) * st);
#pragma unroll
  for (int o = 0; o < 6; o++) bd[o] = ihc * (s[o + 1] - s[o]) * m;
}

// The five WENO5 operands of the left-biased side along axis c from bd
// (PhiDevPlusParameterJB, lsTools.cu:587-673: positions q-2..q+2), with the
// reference's substitutions at the edges.  On the z axis the last plane's
// fifth operand is the z difference one row down (id_kp2 = id_jm1,
// lsTools.cu:636).
__device__ __forceinline__ void stencil_left(const double* src, PDims g, int c, int q, int n,
                                             int x, int y, int z, double ihc, double m,
                                             const double bd[6], double b[5]) {
#pragma unroll
  for (int t = 0; t < 5; t++) b[t] = bd[t];
  if (q == 0) { b[1] = bd[2]; b[0] = bd[3]; }
  if (q == 1) b[0] = bd[1];
  if (q == n - 2) b[4] = bd[3];
  if (q == n - 1) {
    b[3] = bd[2];
    if (c < 2) b[4] = bd[1];
    else b[4] = bdiff(src, g, 2, x, y == 0 ? y : y - 1, z, ihc) * m;
  }
}

// The right-biased side (PhiDevMinusParameterJB, lsTools.cu:676-770:
// positions q+3, q+2, q+1, q, q-1) with its edge substitutions.
__device__ __forceinline__ void stencil_right(int q, int n, const double bd[6], double b[5]) {
#pragma unroll
  for (int t = 0; t < 5; t++) b[t] = bd[5 - t];
  if (q == 0) b[4] = bd[2];
  if (q == n - 1) { b[2] = bd[2]; b[1] = bd[1]; b[0] = bd[0]; }
  if (q == n - 2) { b[1] = bd[3]; b[0] = bd[2]; }
  if (q == n - 3) b[0] = bd[4];
}

// One fused SSP-RK3 stage of cuAdvectLsJB (lsTransportEc.cu:243-399):
// backward differences, the upwind JB WENO derivative, the right-hand side
// advect_RS_WENO (:361-417) and the stage update.  src is the stage input,
// P the step's initial phi; the stage writes out at every padded point, or,
// given phi_int (the last stage), only the interior points, straight into
// the unpadded phi (the values the reference's cuSwapFromGhost copies
// back).  vel is the interior velocity (3 components of
// (nx, ny, nz)).  cuAdvectLsJB takes an already padded velocity from a
// caller the reference does not have, so its ghost values are this
// library's choice: the even mirror image of the interior for every
// component (as phi is padded), read at the mirrored coordinates of the
// interior point instead of from a padded copy.  Near the faces the
// advected phi therefore follows that choice (an odd reflection of the
// wall-normal component would be as plausible for a wall); the reference
// does not pin it.  The metrics (jbn, 11 components of the interior grid;
// 1, 5, 9 used) are read the same way.
__device__ __forceinline__ void weno_stage_point(int stage, const double* src,
                                                 const double* P, const double* vel,
                                                 const double* jbn, PDims g, double iX,
                                                 double iY, double iZ, double dt,
                                                 double* out, double* phi_int, long long id) {
#pragma clang fp contract(fast)
  const unsigned uid = (unsigned)id, row = uid / (unsigned)g.x;  // NG < 2^32
  const int x = (int)(uid - row * (unsigned)g.x), y = (int)(row % (unsigned)g.y),
            z = (int)(row / (unsigned)g.y);
  if (phi_int && (x < GC || x >= g.x - GC || y < GC || y >= g.y - GC || z < GC ||
                  z >= g.z - GC))
    return;  // a ghost point of the last stage: not part of the result
  // The last stage reads the second stage's output only within 3 points of
  // the interior (its stencils' reach), so the second stage skips the two
  // outermost ghost layers (whose values the reference computes and never
  // uses for the interior).  The first stage computes every padded point:
  // the second stage's points reach its outermost layer.
  constexpr int SKIP = GC - 3;
  if (stage == 1 && (x < SKIP || x >= g.x - SKIP || y < SKIP || y >= g.y - SKIP ||
                     z < SKIP || z >= g.z - SKIP))
    return;
  const double ih[3] = {iX, iY, iZ};
  const unsigned nx = (unsigned)(g.x - 2 * GC), ny = (unsigned)(g.y - 2 * GC),
                 nz = (unsigned)(g.z - 2 * GC);
  const unsigned vid = (unsigned)mirror(x, (int)nx) +
                       nx * ((unsigned)mirror(y, (int)ny) + ny * (unsigned)mirror(z, (int)nz));
  const unsigned vN = nx * ny * nz;
  const double met[3] = {gref(jbn + vid), gref(jbn + (size_t)4 * vN + vid),
                         gref(jbn + (size_t)8 * vN + vid)};
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    // upwinding (advect_RS_WENO, lsTools.cu:361-417) uses the left-biased
    // derivative where the velocity is positive and the right-biased one
    // where it is negative, so only that one is formed:
    //   0.5(s+1) v L + 0.5|s-1| v R  =  v L (v > 0),  v R (v < 0),  0
    const double v = gref(vel + (size_t)c * vN + vid);
    if (v == 0.0) {
      r[c] = 0.0;
      continue;
    }
    const int pos[3] = {x, y, z}, ext[3] = {g.x, g.y, g.z};
    const unsigned stride[3] = {1u, (unsigned)g.x, (unsigned)g.x * (unsigned)g.y};
    double bd[6], b[5];
    if (pos[c] >= 3 && pos[c] <= ext[c] - 4) {
      // away from the padded grid's faces: the interior operands, left
      // (q-2..q+2) or right (q+3..q-1) biased, with no edge rules
      axis_bd_inner(src, stride[c], uid, ih[c], met[c], bd);
#pragma unroll
      for (int t = 0; t < 5; t++) b[t] = v > 0.0 ? bd[t] : bd[5 - t];
    } else {
      axis_bd(src, stride[c], pos[c], ext[c], uid, ih[c], met[c], bd);
      if (v > 0.0) stencil_left(src, g, c, pos[c], ext[c], x, y, z, ih[c], met[c], bd, b);
      else stencil_right(pos[c], ext[c], bd, b);
    }
    r[c] = v * weno5(b[0], b[1], b[2], b[3], b[4]);
  }
  const double rs = r[0] + r[1] + r[2];
  const double sv = gref(src + id);
  double o;
  if (stage == 0) o = sv - dt * rs;
  else if (stage == 1) o = 0.75 * gref(P + id) + 0.25 * (sv - dt * rs);
  else o = (gref(P + id) + 2.0 * (sv - dt * rs)) * (1.0 / 3.0);
  if (phi_int)
    gref(phi_int + ((unsigned)(x - GC) + nx * ((unsigned)(y - GC) + ny * (unsigned)(z - GC)))) = o;
  else
    gref(out + id) = o;
}

__global__ void k_weno_stage(int stage, const double* src, const double* P,
                             const double* vel, const double* jbn, PDims g,
                             double iX, double iY, double iZ, double dt,
                             double* out, double* phi_int) {
  const long long NG = g.N();
  for (long long id = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; id < NG;
       id += (long long)gridDim.x * blockDim.x)
    weno_stage_point(stage, src, P, vel, jbn, g, iX, iY, iZ, dt, out, phi_int, id);
}

// ---------------------------------------------------------------------------
// Level-set reinitialisation (cuReinitLsJB, lsTransportEc.cu:422-676).

// The padded coordinates whose mirror image (cuGhostCellsMirror3d{Z,Y,X},
// lsTools.cu:482-585, applied over the full padded ranges) is the interior
// padded coordinate X of a line of n interior points: X itself, its
// reflection into the low ghost layers and into the high ones.
__device__ __forceinline__ int mirror_images(int X, int n, int c[3]) {
  int k = 0;
  c[k++] = X;
  if (X < 2 * GC) c[k++] = 2 * GC - 1 - X;
  if (X >= n) c[k++] = 2 * (n + GC) - 1 - X;
  return k;
}

// One SSP-RK3 stage of the reinitialisation at every padded point: the two
// JB WENO derivatives per axis, the Godunov |grad phi| by the sign of phi0
// and rs = S(phi0)(|grad phi| - 1), S = phi0/sqrt(phi0^2 + dmin^2)
// (reini_RS_WENOJB, lsTools.cu:824-900), then RunGK_{First,Second,Third}S
// (extrapol.cu:155-212).  src is the stage input, P the iteration's phi,
// which is also its phi0 (the reference copies phi to phi0 at the start of
// every iteration).  The last stage computes the interior points and writes
// each into out, a different array, together with its mirror images in the
// ghost layers: out is then the next iteration's phi with the ghost
// mirroring that iteration would start with already applied.
__global__ void k_ls_reinit_stage(int stage, const double* src, const double* P,
                                  const double* dmin, const double* j1, const double* j5,
                                  const double* j9, PDims g, double iX, double iY, double iZ,
                                  double dt, double* out) {
  const long long NG = g.N();
  const double* P0 = P;
  for (long long id = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
       id < NG; id += (long long)gridDim.x * blockDim.x) {
    const unsigned uid = (unsigned)id, row = uid / (unsigned)g.x;  // NG < 2^32
    const int x = (int)(uid - row * (unsigned)g.x), y = (int)(row % (unsigned)g.y),
              z = (int)(row / (unsigned)g.y);
    if (stage == 2 && (x < GC || x >= g.x - GC || y < GC || y >= g.y - GC || z < GC ||
                       z >= g.z - GC))
      continue;  // a ghost point: written as an interior point's image
    const double ih[3] = {iX, iY, iZ};
    const double met[3] = {j1[id], j5[id], j9[id]};
    const double p0 = P0[id];
    const double ta = (double)(p0 > 0.0) - (double)(p0 < 0.0);
    double r[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int pos[3] = {x, y, z}, ext[3] = {g.x, g.y, g.z};
      const unsigned stride[3] = {1u, (unsigned)g.x, (unsigned)g.x * (unsigned)g.y};
      double bd[6], L, R;
      if (pos[c] >= 3 && pos[c] <= ext[c] - 4) {
        // away from the padded grid's faces: the interior operands, left
        // (q-2..q+2) and right (q+3..q-1) biased, with no edge rules
        axis_bd_inner(src, stride[c], uid, ih[c], met[c], bd);
        L = weno5(bd[0], bd[1], bd[2], bd[3], bd[4]);
        R = weno5(bd[5], bd[4], bd[3], bd[2], bd[1]);
      } else {
        double b[5];
        axis_bd(src, stride[c], pos[c], ext[c], uid, ih[c], met[c], bd);
        stencil_left(src, g, c, pos[c], ext[c], x, y, z, ih[c], met[c], bd, b);
        L = weno5(b[0], b[1], b[2], b[3], b[4]);
        stencil_right(pos[c], ext[c], bd, b);
        R = weno5(b[0], b[1], b[2], b[3], b[4]);
      }
      const double lp = L > 0.0 ? L : 0.0, rm = R < 0.0 ? R : 0.0;
      double hi = lp * lp, lo = rm * rm;
      r[c] = 0.5 * (ta + 1.0) * (hi > lo ? hi : lo);
      const double rp = R > 0.0 ? R : 0.0, lm = L < 0.0 ? L : 0.0;
      hi = rp * rp;
      lo = lm * lm;
      r[c] += 0.5 * fabs(ta - 1.0) * (hi > lo ? hi : lo);
    }
    const double grad_mod = sqrt(r[0] + r[1] + r[2]);
    const double d = dmin[id];
    const double so = p0 / sqrt(p0 * p0 + d * d);
    const double rs = 1.0 * so * (grad_mod - 1.0);
    const double sv = src[id];
    if (stage == 0) {
      out[id] = sv - dt * rs;
    } else if (stage == 1) {
      out[id] = 0.75 * P[id] + 0.25 * (sv - dt * rs);
    } else {
      const double o = (P[id] + 2.0 * (sv - dt * rs)) / 3.0;
      int cx[3], cy[3], cz[3];
      const int kx = mirror_images(x, g.x - 2 * GC, cx), ky = mirror_images(y, g.y - 2 * GC, cy),
                kz = mirror_images(z, g.z - 2 * GC, cz);
      for (int c = 0; c < kz; c++)
        for (int b = 0; b < ky; b++)
          for (int a = 0; a < kx; a++)
            out[(unsigned)cx[a] + (unsigned)g.x * ((unsigned)cy[b] + (unsigned)g.y * (unsigned)cz[c])] = o;
    }
  }
}

// ---------------------------------------------------------------------------
// The same step for many grids in one launch per phase (the batched stage):
// workgroup b serves grid gi with blk_start[gi] <= b < blk_start[gi+1],
// points (b - blk_start[gi])*256 + threadIdx.x of the padded grid (the pad
// and the three stages; the last writes the interior of phi).
__device__ __forceinline__ int batch_grid(const int* blk_start, int ngrid) {
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  int lo = 0, hi = ngrid;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (blk_start[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256)
k_batch_pad(const LsBatchGrid* grids, const int* blk_start, int ngrid) {
  const int gi = batch_grid(blk_start, ngrid);
  const LsBatchGrid& G = grids[gi];
  const PDims g{G.nx + 2 * GC, G.ny + 2 * GC, G.nz + 2 * GC};
  const long long p = (long long)((int)xcd_block(blockIdx.x, gridDim.x) - blk_start[gi]) * 256 + threadIdx.x;
  if (p >= g.N()) return;
  const unsigned up = (unsigned)p, row = up / (unsigned)g.x;
  const int i = (int)(up - row * (unsigned)g.x), j = (int)(row % (unsigned)g.y),
            k = (int)(row / (unsigned)g.y);
  const long long src = mirror(i, G.nx) + (long long)G.nx * (mirror(j, G.ny) +
                                                             (long long)G.ny * mirror(k, G.nz));
  gref(G.P + p) = gref(G.phi + src);
}

__global__ void __launch_bounds__(256)
k_batch_weno_stage(const LsBatchGrid* grids, const int* blk_start, int ngrid, int stage) {
  const int gi = batch_grid(blk_start, ngrid);
  const LsBatchGrid& G = grids[gi];
  const PDims g{G.nx + 2 * GC, G.ny + 2 * GC, G.nz + 2 * GC};
  const long long NG = g.N();
  const long long id = (long long)((int)xcd_block(blockIdx.x, gridDim.x) - blk_start[gi]) * 256 + threadIdx.x;
  if (id >= NG) return;
  const double* src = stage == 0 ? G.P : (stage == 1 ? G.T1 : G.T2);
  double* out = stage == 0 ? G.T1 : G.T2;
  weno_stage_point(stage, src, G.P, G.vel, G.jbn, g, G.iX, G.iY, G.iZ,
                   G.dt, out, stage == 2 ? G.phi : nullptr, id);
}

unsigned blocks_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 131072) b = 131072;
  return (unsigned)(b < 1 ? 1 : b);
}

PDims padded(const Dims& d) { return PDims{d.nx + 2 * GC, d.ny + 2 * GC, d.nz + 2 * GC}; }

}  // namespace

// ten padded fields, then the solid-point list of the extrapolation: chunk
// counts, the count and up to one index per padded point (ints)
size_t ls_scratch_doubles(const Dims& d) {
  const size_t NG = (size_t)padded(d).N();
  const size_t ints = (NG + LS_CHUNK - 1) / LS_CHUNK + 1 + NG;
  return 10 * NG + (ints + 1) / 2 + 1;
}

long long ls_padded_points(int nx, int ny, int nz) {
  return (long long)(nx + 2 * GC) * (ny + 2 * GC) * (nz + 2 * GC);
}

int op_batch_ls_step(hipStream_t st, const LsBatchGrid* grids, const int* blk_pad,
                     int nblk_pad, int ngrid) {
  hipLaunchKernelGGL(k_batch_pad, dim3(nblk_pad), dim3(256), 0, st, grids, blk_pad, ngrid);
  for (int stage = 0; stage < 3; stage++)
    hipLaunchKernelGGL(k_batch_weno_stage, dim3(nblk_pad), dim3(256), 0, st, grids, blk_pad,
                       ngrid, stage);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_ls_extrapolate(emx_ctx* ctx, double* pres, const double* jbn,
                      const double* sobject, double dx, double dy, double dz,
                      double deltamin) {
  const Dims d = ctx->d;
  const PDims g = padded(d);
  const long long N = d.N(), NG = g.N();
  double* P = ctx->ls_buf;
  double* S = P + NG;
  double* geo = S + NG;  // 6 per solid point, at most 6*NG
  double* D1 = geo + 6 * NG;
  double* D2 = D1 + NG;
  hipStream_t st = ctx->stream;
  const unsigned bg = blocks_for(NG);
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
          hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, pres, d.nx, d.ny, d.nz, P, g));
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
          hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, sobject, d.nx, d.ny, d.nz, S, g));
  // cuExtrapolation(d_press, d_sobject, d_jbn, deltaX.., NxG.., deltamin*0.5, 1)
  // hands 1.0/deltaX to every kernel (interface.cu:373-421)
  const double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz, dt = deltamin * 0.5;
  // the solid points, in order; D1 = D2 = P at every point (the fluid
  // points are never written again)
  const int nb = (int)((NG + LS_CHUNK - 1) / LS_CHUNK);
  int* bcount = reinterpret_cast<int*>(D2 + NG);
  int* count = bcount + nb;
  int* list = count + 1;
  EMX_TIMED(ctx, "ls_solid_list", 8.0 * NG,
            hipLaunchKernelGGL(k_ls_solid_count, dim3(nb), dim3(256), 0, st, S, NG, bcount));
  EMX_TIMED(ctx, "ls_solid_list", 0.0,
            hipLaunchKernelGGL(k_ls_scan, dim3(1), dim3(1024), 0, st, bcount, nb, count));
  EMX_TIMED(ctx, "ls_solid_list", 8.0 * NG,
            hipLaunchKernelGGL(k_ls_solid_write, dim3(nb), dim3(256), 0, st, S, NG, bcount, list));
  // algorithmic bytes per solid point (the count is only known on the
  // device; read back while the context records timings): geometry = list
  // entry, S, three metrics in, six values out; a stage = list entry, ext,
  // the six values, P (stages 1, 2), out
  double nsolid = 0.0;
  if (ctx->timing && (ctx->timing_label.empty() || ctx->timing_label == "ls_extrap_stage" ||
                      ctx->timing_label == "ls_solid_geom")) {
    int h = 0;
    EMX_TRY_HIP(hipMemcpyAsync(&h, count, sizeof(int), hipMemcpyDeviceToHost, st));
    EMX_TRY_HIP(hipStreamSynchronize(st));
    nsolid = h;
  }
  const unsigned bl = bg < 4096u ? bg : 4096u;
  EMX_TIMED(ctx, "ls_solid_geom", (4.0 + 8.0 * 10) * nsolid,
            hipLaunchKernelGGL(k_ls_solid_geom, dim3(bl), dim3(256), 0, st, list, count, S, jbn,
                               d.nx, d.ny, d.nz, g, iX, iY, iZ, geo));
  EMX_TIMED(ctx, "ls_copy", 24.0 * NG,
            hipLaunchKernelGGL(k_copy2, dim3(bg), dim3(256), 0, st, P, D1, D2, NG));
  EMX_TRY_HIP(hipGetLastError());
  for (int it = 1; it <= 10; it++) {
    EMX_TIMED(ctx, "ls_extrap_stage", (4.0 + 8.0 * 8) * nsolid,
              hipLaunchKernelGGL(k_ls_extrap_list, dim3(bl), dim3(256), 0, st, 0, list, count, P,
                                 P, geo, g, iX, iY, iZ, dt, D1));
    EMX_TIMED(ctx, "ls_extrap_stage", (4.0 + 8.0 * 9) * nsolid,
              hipLaunchKernelGGL(k_ls_extrap_list, dim3(bl), dim3(256), 0, st, 1, list, count, D1,
                                 P, geo, g, iX, iY, iZ, dt, D2));
    EMX_TIMED(ctx, "ls_extrap_stage", (4.0 + 8.0 * 9) * nsolid,
              hipLaunchKernelGGL(k_ls_extrap_list, dim3(bl), dim3(256), 0, st, 2, list, count, D2,
                                 P, geo, g, iX, iY, iZ, dt, P));
  }
  EMX_TIMED(ctx, "ls_unpad", 16.0 * N,
          hipLaunchKernelGGL(k_unpad, dim3(blocks_for(N)), dim3(256), 0, st, P, g,
                     d.nx, d.ny, d.nz, pres));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_ls_reinit(emx_ctx* ctx, double* phi, const double* jbn, const double* dmin,
                 double dx, double dy, double dz, double deltamin) {
  const Dims d = ctx->d;
  const PDims g = padded(d);
  const long long N = d.N(), NG = g.N();
  double* P = ctx->ls_buf;  // the iteration's phi, ping-pong with Q
  double* Q = P + NG;
  double* J1 = Q + NG;
  double* J5 = J1 + NG;
  double* J9 = J5 + NG;
  double* Dm = J9 + NG;
  double* T1 = Dm + NG;
  double* T2 = T1 + NG;
  hipStream_t st = ctx->stream;
  const unsigned bg = blocks_for(NG);
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
            hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, phi, d.nx, d.ny, d.nz, P, g));
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
            hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, dmin, d.nx, d.ny, d.nz, Dm, g));
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
            hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, jbn, d.nx, d.ny, d.nz, J1, g));
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
            hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, jbn + 4 * N, d.nx, d.ny, d.nz,
                               J5, g));
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
            hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, jbn + 8 * N, d.nx, d.ny, d.nz,
                               J9, g));
  // Dev1thO_Downwind receives 1.0/deltaX (lsTransportEc.cu:487-495)
  const double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz, dt = 0.5 * deltamin;
  // the padding already mirrors the ghost layers of the first iteration
  for (int it = 1; it <= 10; it++) {
    // per point: src, three metrics, phi (= phi0), dmin, out
    EMX_TIMED(ctx, "ls_reinit_stage", 56.0 * NG,
              hipLaunchKernelGGL(k_ls_reinit_stage, dim3(bg), dim3(256), 0, st, 0, P, P, Dm,
                                 J1, J5, J9, g, iX, iY, iZ, dt, T1));
    EMX_TIMED(ctx, "ls_reinit_stage", 56.0 * NG,
              hipLaunchKernelGGL(k_ls_reinit_stage, dim3(bg), dim3(256), 0, st, 1, T1, P, Dm,
                                 J1, J5, J9, g, iX, iY, iZ, dt, T2));
    EMX_TIMED(ctx, "ls_reinit_stage", 56.0 * NG,
              hipLaunchKernelGGL(k_ls_reinit_stage, dim3(bg), dim3(256), 0, st, 2, T2, P, Dm,
                                 J1, J5, J9, g, iX, iY, iZ, dt, Q));
    double* t = P;
    P = Q;
    Q = t;
  }
  EMX_TIMED(ctx, "ls_unpad", 16.0 * N,
            hipLaunchKernelGGL(k_unpad, dim3(blocks_for(N)), dim3(256), 0, st, P, g, d.nx, d.ny,
                               d.nz, phi));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_ls_advect_weno5(emx_ctx* ctx, double* phi, const double* vel,
                       const double* jbn, double dx, double dy, double dz,
                       double dt) {
  const Dims d = ctx->d;
  const PDims g = padded(d);
  const long long N = d.N(), NG = g.N();
  double* P = ctx->ls_buf;
  double* T1 = P + NG;
  double* T2 = T1 + NG;
  hipStream_t st = ctx->stream;
  const unsigned bg = blocks_for(NG);
  EMX_TIMED(ctx, "ls_pad", 16.0 * NG,
          hipLaunchKernelGGL(k_pad, dim3(bg), dim3(256), 0, st, phi, d.nx, d.ny, d.nz, P, g));
  const double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz;
  // per stage: phi in (and the step's phi) and phi out on the padded grid;
  // the velocity and the metrics once on the interior grid
  EMX_TIMED(ctx, "ls_weno_stage", 16.0 * NG + 48.0 * N,
          hipLaunchKernelGGL(k_weno_stage, dim3(bg), dim3(256), 0, st, 0, P, P, vel, jbn, g,
                     iX, iY, iZ, dt, T1, nullptr));
  EMX_TIMED(ctx, "ls_weno_stage", 24.0 * NG + 48.0 * N,
          hipLaunchKernelGGL(k_weno_stage, dim3(bg), dim3(256), 0, st, 1, T1, P, vel, jbn, g,
                     iX, iY, iZ, dt, T2, nullptr));
  EMX_TIMED(ctx, "ls_weno_stage", 16.0 * NG + 56.0 * N,
          hipLaunchKernelGGL(k_weno_stage, dim3(bg), dim3(256), 0, st, 2, T2, P, vel, jbn, g,
                     iX, iY, iZ, dt, nullptr, phi));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

namespace {

// heviside (solver/advectLS.f90:29-59): the smoothed Heaviside of phi over
// the band |phi| <= eps, in the reference's order of operations.  A point
// that none of the three cases takes (NaN phi) keeps its hevi, as there.
__global__ void k_heviside(const double* phi, long long N, double eps, double* hevi) {
  const double pi = 3.14159265358979323846;  // 2 acos(0) in double
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
    const double f = phi[p];
    if (f < -eps) hevi[p] = 0.0;
    if (f >= -eps && f <= eps) hevi[p] = 0.5 + f / (2.0 * eps) + sin(pi * f / eps) / (2.0 * pi);
    if (f > eps) hevi[p] = 1.0;
  }
}

}  // namespace

int op_ls_heviside(emx_ctx* ctx, const double* phi, double deltamax, double* hevi) {
  const long long N = ctx->d.N();
  EMX_TIMED(ctx, "ls_heviside", 16.0 * N,
            hipLaunchKernelGGL(k_heviside, dim3(blocks_for(N)), dim3(256), 0, ctx->stream, phi, N,
                               1.5 * deltamax, hevi));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx
