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

int op_sgs_geometry(emx_ctx* ctx, const double* x, double* sgd,
                    double* dmins, double* blockmax, int nblocks) {
  hipLaunchKernelGGL(k_sgs_geometry, dim3(nblocks), dim3(256), 0, ctx->stream,
                     x, ctx->d, sgd, dmins, blockmax);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

// viscosidad_t + viscosidad in one pass: f_structura
// (solver/sgdm_incomp.f90:180-298), amut = amu0*sqrt(F2)*dxyzsgd (:129-137),
// VORT (:330-341), SELECTIVO (:360-392) and viscosidad
// (solver/viscosidad_incomp.f90:12-20), in the reference's operation order.  A workgroup owns a 32 x 8 tile of (i, j) and marches k
// through a slab; the velocity and the vorticity of planes k-1, k, k+1 are
// held in LDS with a one-point halo in i and j (rolling over three slots),
// so every input is read from memory once per workgroup instead of once
// per neighbour, and the vorticity never goes to memory.  Tiles are
// numbered i fastest, then j, then slab, and dealt to the XCDs in
// contiguous ranges (xcd_block): the tiles that share a halo run on the
// same XCD at the same time, so the halo's cache lines come from its L2.
constexpr int SG_TX = 32, SG_TY = 8, SG_HX = SG_TX + 2, SG_HY = SG_TY + 2,
              SG_HP = SG_HX * SG_HY;

__global__ void __launch_bounds__(256)
k_sgs_fused(const double* u, const double* dvel, const double* jbn, const double* sgd,
            Dims d, int kz, double c4, double* amut, double* vis, double* vis5) {
  __shared__ double su[3][3][SG_HP];  // [slot][component][halo point]
  __shared__ double sw[3][3][SG_HP];
  const long long N = d.N();
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  const int tx = threadIdx.x % SG_TX, ty = threadIdx.x / SG_TX;
  const unsigned nbx = (d.nx + SG_TX - 1) / SG_TX, nby = (d.ny + SG_TY - 1) / SG_TY;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (int)(b % nbx) * SG_TX, j0 = (int)((b / nbx) % nby) * SG_TY;
  const int k0 = (int)(b / (nbx * nby)) * kz, k1 = min(k0 + kz, d.nz);
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i < d.nx && j < d.ny;
  const double amu0 = 0.1781;
  const double cutoff = cos(20.0 * atan(1.0) / 45.0);
  // plane kk (clamped into the grid) into slot kk mod 3; halo points
  // outside the grid hold clamped copies and are never used
  auto load_plane = [&](int kk) {
    const int slot = ((kk % 3) + 3) % 3;
    const int gk = kk < 0 ? 0 : (kk >= d.nz ? d.nz - 1 : kk);
    for (int h = threadIdx.x; h < SG_HP; h += blockDim.x) {
      int gi = i0 - 1 + h % SG_HX, gj = j0 - 1 + h / SG_HX;
      gi = gi < 0 ? 0 : (gi >= d.nx ? d.nx - 1 : gi);
      gj = gj < 0 ? 0 : (gj >= d.ny ? d.ny - 1 : gj);
      const long long q = gi + gj * sy + gk * sz;
#pragma unroll
      for (int c = 0; c < 3; c++) su[slot][c][h] = u[c * N + q];
      const double j1 = jbn[q], j5 = jbn[4 * N + q], j9 = jbn[8 * N + q];
      sw[slot][0][h] = j5 * dvel[7 * N + q] - j9 * dvel[5 * N + q];
      sw[slot][1][h] = j9 * dvel[2 * N + q] - j1 * dvel[6 * N + q];
      sw[slot][2][h] = j1 * dvel[3 * N + q] - j5 * dvel[N + q];
    }
  };
  load_plane(k0 - 1);
  load_plane(k0);
  const int hc = (ty + 1) * SG_HX + (tx + 1);  // own halo index
  for (int k = k0; k < k1; k++) {
    load_plane(k + 1);
    __syncthreads();
    if (own) {
      const long long p = i + j * sy + k * sz;
      const int sm = ((k - 1) % 3 + 3) % 3, s0 = k % 3, sp = (k + 1) % 3;
      // structure function, six one-sided differences
      int a_s, a_h, b_s, b_h;  // (slot, halo index) of the two points
      double s = 0.0, du, dv, dw, fac;
#define TERM(FAC)                                                           \
  du = (su[a_s][0][a_h] - su[b_s][0][b_h]) * (FAC);                         \
  dv = (su[a_s][1][a_h] - su[b_s][1][b_h]) * (FAC);                         \
  dw = (su[a_s][2][a_h] - su[b_s][2][b_h]) * (FAC);
      // x+
      if (i < d.nx - 1) { a_s = s0; a_h = hc + 1; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc; b_s = s0; b_h = hc - 1; }
      fac = sgd[N + p];
      TERM(fac);
      s = du * du + dv * dv + dw * dw;
      // x-
      if (i > 0) { a_s = s0; a_h = hc - 1; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc + 1; b_s = s0; b_h = hc; }
      fac = sgd[2 * N + p];
      TERM(fac);
      s = s + du * du + dv * dv + dw * dw;
      // y+
      if (j < d.ny - 1) { a_s = s0; a_h = hc + SG_HX; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc; b_s = s0; b_h = hc - SG_HX; }
      fac = sgd[3 * N + p];
      TERM(fac);
      s = s + du * du + dv * dv + dw * dw;
      // y-
      if (j > 0) { a_s = s0; a_h = hc - SG_HX; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc; b_s = s0; b_h = hc + SG_HX; }
      fac = sgd[4 * N + p];
      TERM(fac);
      s = s + du * du + dv * dv + dw * dw;
      // z+
      if (k < d.nz - 1) { a_s = sp; a_h = hc; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc; b_s = sm; b_h = hc; }
      fac = sgd[5 * N + p];
      TERM(fac);
      s = s + du * du + dv * dv + dw * dw;
      // z-
      if (k > 0) { a_s = sm; a_h = hc; b_s = s0; b_h = hc; }
      else { a_s = s0; a_h = hc; b_s = sp; b_h = hc; }
      fac = sgd[6 * N + p];
      TERM(fac);
      s = s + du * du + dv * dv + dw * dw;
#undef TERM
      const double f2 = s / 6.e0;
      double m = amu0 * sqrt(f2) * sgd[p];
      // SELECTIVO: neighbours n1..n6 = x+, x-, z+, z-, y+, y- (self at walls)
      const int n1s = s0, n1h = i == d.nx - 1 ? hc : hc + 1;
      const int n2s = s0, n2h = i == 0 ? hc : hc - 1;
      const int n3s = k == d.nz - 1 ? s0 : sp, n3h = hc;
      const int n4s = k == 0 ? s0 : sm, n4h = hc;
      const int n5s = s0, n5h = j == d.ny - 1 ? hc : hc + SG_HX;
      const int n6s = s0, n6h = j == 0 ? hc : hc - SG_HX;
      double wn[3];
#pragma unroll
      for (int c = 0; c < 3; c++)
        wn[c] = sw[n1s][c][n1h] + sw[n2s][c][n2h] + sw[n3s][c][n3h] + sw[n4s][c][n4h] +
                sw[n5s][c][n5h] + sw[n6s][c][n6h];
      const double a = sw[s0][0][hc], b = sw[s0][1][hc], c3 = sw[s0][2][hc];
      double wnn = a * a + b * b + c3 * c3;
      wnn = sqrt(wnn * (wn[0] * wn[0] + wn[1] * wn[1] + wn[2] * wn[2]));
      const double s1 = (a * wn[0] + b * wn[1] + c3 * wn[2]) / wnn;
      if (s1 > cutoff) m = 0.0;
      amut[p] = m;
      if (vis5) vis5[p] = m / 0.7;
      vis[p] = c4 * 1.0 + m;
    }
    __syncthreads();  // the slot of plane k-1 is reloaded next
  }
}

int op_sgdm(emx_ctx* ctx, const double* u, const double* dvel,
            const double* jbn, const double* sgd, double c4, double* amut,
            double* vis, double* vis5) {
  const Dims d = ctx->d;
  const int kz = 16;
  const unsigned tiles = (unsigned)((d.nx + SG_TX - 1) / SG_TX) *
                         (unsigned)((d.ny + SG_TY - 1) / SG_TY) * (unsigned)((d.nz + kz - 1) / kz);
  const dim3 grid(tiles);
  // reads u (3), dvel (6), jbn (3), sgd (7); writes amut, vis (, vis5)
  EMX_TIMED(ctx, "sgdm", (vis5 ? 176.0 : 168.0) * d.N(),
            hipLaunchKernelGGL(k_sgs_fused, grid, dim3(256), 0, ctx->stream, u, dvel, jbn,
                               sgd, d, kz, c4, amut, vis, vis5));
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
// indices at the walls, applied along x, then y, then z.
//
// Two kernels, both on the plane march of k_sgs_fused: a workgroup owns a
// 32 x 8 (i, j) tile and a slab of k planes.  For each plane (and one
// clamped plane either side of the slab) it forms the filtered quantities
// at the tile and its one-point halo into LDS, filters them along x and y
// from LDS, and filters along z in registers as the planes pass:
//   Z(k) = (1/4 Y(k-1) + 1/2 Y(k)) + 1/4 Y(k+1),
// the term of plane k+1 added when that plane arrives, so every quantity
// of a point is formed once per plane and read once from memory.
//   k_dynsmag_lm:  the 21 quantities S (6), |S| S (6), u_a u_b (6), u (3)
//                  -> L:M and M:M (two fields of scratch);
//   k_dynsmag_nut: the filtered L:M and M:M -> C, nu_t, vis.
constexpr int DS_TX = 32, DS_TY = 8, DS_HX = DS_TX + 2, DS_HY = DS_TY + 2,
              DS_HP = DS_HX * DS_HY, DS_KS = 16;

// Strain-rate pairs (0,0) (1,1) (2,2) (0,1) (0,2) (1,2); off-diagonal pairs
// count twice in contractions.
__device__ __forceinline__ void strain(const double* dvel, const double* jbn, long long N,
                                       long long q, double S[6], double* aS) {
  const double m[3] = {jbn[q], jbn[4 * N + q], jbn[8 * N + q]};
  double g[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) g[3 * a + b] = m[b] * dvel[(3 * b + a) * N + q];
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  double s2 = 0.0;
#pragma unroll
  for (int t = 0; t < 6; t++) {
    S[t] = 0.5 * (g[3 * ia[t] + ib[t]] + g[3 * ib[t] + ia[t]]);
    s2 += (t < 3 ? 1.0 : 2.0) * S[t] * S[t];
  }
  *aS = sqrt(2.0 * s2);
}

// The plane march shared by both kernels.  form(q, v) writes the NQ
// quantities of grid point q; done(p, Z) receives the filtered quantities
// of an owned point p of the slab.  Planes k0-1 .. k1 are formed (clamped
// into the grid); when plane kk arrives, Z(kk-1) is completed and handed
// to done(), Z(kk) gets its middle term and Z(kk+1) its first.
template <int NQ, class Form, class Done>
__device__ __forceinline__ void dyn_march(const Dims& d, double (*Q)[DS_HP], Form form,
                                          Done done) {
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  const int tx = threadIdx.x % DS_TX, ty = threadIdx.x / DS_TX;
  const unsigned nbx = (d.nx + DS_TX - 1) / DS_TX, nby = (d.ny + DS_TY - 1) / DS_TY;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (int)(b % nbx) * DS_TX, j0 = (int)((b / nbx) % nby) * DS_TY;
  const int k0 = (int)(b / (nbx * nby)) * DS_KS, k1 = min(k0 + DS_KS, d.nz);
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i < d.nx && j < d.ny;
  const int hc = (ty + 1) * DS_HX + (tx + 1);
  double Zn[NQ] = {}, Zc[NQ] = {};  // Z(kk+1) after its first term, Z(kk) after two
  for (int kk = k0 - 1; kk <= k1; kk++) {
    const int gk = kk < 0 ? 0 : (kk >= d.nz ? d.nz - 1 : kk);
    for (int h = threadIdx.x; h < DS_HP; h += blockDim.x) {
      int gi = i0 - 1 + h % DS_HX, gj = j0 - 1 + h / DS_HX;
      gi = gi < 0 ? 0 : (gi >= d.nx ? d.nx - 1 : gi);
      gj = gj < 0 ? 0 : (gj >= d.ny ? d.ny - 1 : gj);
      double v[NQ];
      form(gi + gj * sy + gk * sz, v);
#pragma unroll
      for (int t = 0; t < NQ; t++) Q[t][h] = v[t];
    }
    __syncthreads();
    if (own) {
      double Zd[NQ];
#pragma unroll
      for (int t = 0; t < NQ; t++) {
        double X[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const int c = hc + (r - 1) * DS_HX;
          X[r] = 0.25 * Q[t][c - 1] + 0.5 * Q[t][c] + 0.25 * Q[t][c + 1];
        }
        const double Y = 0.25 * X[0] + 0.5 * X[1] + 0.25 * X[2];
        Zd[t] = Zc[t] + 0.25 * Y;     // Z(kk-1), valid for kk > k0
        Zc[t] = Zn[t] + 0.5 * Y;      // Z(kk), valid for kk >= k0
        Zn[t] = 0.25 * Y;             // Z(kk+1)
      }
      if (kk > k0) done(i + j * sy + (long long)(kk - 1) * sz, Zd);
    }
    __syncthreads();  // Q is rewritten for the next plane
  }
}

// L:M and M:M at each point: with hat() the test filter,
//   L_ij = hat(u_i u_j) - hat(u_i) hat(u_j),
//   M_ij = 2 delta^2 (hat(|S| S_ij) - 4 |hat S| hat S_ij).
__global__ void __launch_bounds__(256)
k_dynsmag_lm(const double* u, const double* dvel, const double* jbn, const double* delta,
             Dims d, double* LM, double* MM) {
  __shared__ double Q[21][DS_HP];
  const long long N = d.N();
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  dyn_march<21>(
      d, Q,
      [&](long long q, double* v) {
        double S[6], aS;
        strain(dvel, jbn, N, q, S, &aS);
        const double uq[3] = {u[q], u[N + q], u[2 * N + q]};
#pragma unroll
        for (int t = 0; t < 6; t++) {
          v[t] = S[t];
          v[6 + t] = aS * S[t];
          v[12 + t] = uq[ia[t]] * uq[ib[t]];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) v[18 + c] = uq[c];
      },
      [&](long long p, const double* Z) {
        double s2 = 0.0;
#pragma unroll
        for (int t = 0; t < 6; t++) s2 += (t < 3 ? 1.0 : 2.0) * Z[t] * Z[t];
        const double ahS = sqrt(2.0 * s2);
        const double d2 = delta[p] * delta[p];
        double lm = 0.0, mm = 0.0;
#pragma unroll
        for (int t = 0; t < 6; t++) {
          const double L = Z[12 + t] - Z[18 + ia[t]] * Z[18 + ib[t]];
          const double M = 2.0 * d2 * (Z[6 + t] - 4.0 * ahS * Z[t]);
          const double wq = t < 3 ? 1.0 : 2.0;
          lm += wq * L * M;
          mm += wq * M * M;
        }
        LM[p] = lm;
        MM[p] = mm;
      });
}

// nu_t = max(0, C delta^2 |S|) with C = hat(L:M) / hat(M:M) (0 where
// hat(M:M) <= 0); vis = c4 + nu_t when vis is not NULL.
__global__ void __launch_bounds__(256)
k_dynsmag_nut(const double* LM, const double* MM, const double* dvel, const double* jbn,
              const double* delta, Dims d, double* nut, double c4, double* vis) {
  __shared__ double Q[2][DS_HP];
  const long long N = d.N();
  dyn_march<2>(
      d, Q,
      [&](long long q, double* v) {
        v[0] = LM[q];
        v[1] = MM[q];
      },
      [&](long long p, const double* Z) {
        double S[6], aS;
        strain(dvel, jbn, N, p, S, &aS);
        const double C = Z[1] > 0.0 ? Z[0] / Z[1] : 0.0;
        const double v = C * delta[p] * delta[p] * aS;
        const double nu = v > 0.0 ? v : 0.0;
        nut[p] = nu;
        if (vis) vis[p] = c4 * 1.0 + nu;  // viscosidad (solver/viscosidad_incomp.f90:15-17)
      });
}

// Launches the two kernels; LM and MM are 2*N doubles of scratch.
int launch_dynsmag(emx_ctx* ctx, hipStream_t stream, const Dims& d, const double* u,
                   const double* dvel, const double* jbn, const double* delta, double* nut,
                   double c4, double* vis, double* LM, double* MM) {
  const long long N = d.N();
  const unsigned tiles = (unsigned)((d.nx + DS_TX - 1) / DS_TX) *
                         (unsigned)((d.ny + DS_TY - 1) / DS_TY) *
                         (unsigned)((d.nz + DS_KS - 1) / DS_KS);
  // reads u(3), dvel(9), jbn(3), delta; writes LM, MM / reads LM, MM,
  // dvel(9), jbn(3), delta; writes nut (and vis)
  auto lm = [&] {
    hipLaunchKernelGGL(k_dynsmag_lm, dim3(tiles), dim3(256), 0, stream, u, dvel, jbn, delta, d,
                       LM, MM);
  };
  auto nu = [&] {
    hipLaunchKernelGGL(k_dynsmag_nut, dim3(tiles), dim3(256), 0, stream, LM, MM, dvel, jbn,
                       delta, d, nut, c4, vis);
  };
  if (ctx) {
    EMX_TIMED(ctx, "dynsmag_lm", 18.0 * 8.0 * N, lm());
    EMX_TIMED(ctx, "dynsmag_nut", (vis ? 17.0 : 16.0) * 8.0 * N, nu());
  } else {
    lm();
    nu();
  }
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_dynsmag(emx_ctx* ctx, const double* u, const double* dvel, const double* jbn,
               const double* delta, double* nut, double c4, double* vis) {
  const long long N = ctx->d.N();
  return launch_dynsmag(ctx, ctx->stream, ctx->d, u, dvel, jbn, delta, nut, c4, vis,
                        ctx->sgs_tmp, ctx->sgs_tmp + N);
}

}  // namespace emx
