// Operator kernels of the RK-stage spatial operator: gradient, fluxes,
// divergence, the fused flux-divergence, grid metrics and the
// structure-function eddy viscosity.
#include <cmath>

#include "compact.hpp"
#include "ops.hpp"

namespace emx {

// ---------------------------------------------------------------------------
// Problems for the line solver.

// NV independent fields in, NV derivatives out.
template <int N_>
struct ProbDeriv {
  static constexpr int NV = N_;
  const double* in[N_];
  double* out[N_];
  __device__ __forceinline__ void load(long long p, double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) v[a] = in[a][p];
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) out[a][p] = v[a];
  }
};

// Inputs of flujos (solver/flujos_compact.f90) addressed by linear index.
struct FluxIn {
  const double* um;    // 5 comps
  const double* u;     // 3 comps
  const double* pres;
  const double* vis;
  const double* jbn;   // 11 comps
  const double* dvel;  // 9 comps
  long long N;
  __device__ __forceinline__ double UM(int m, long long p) const { return um[(m - 1) * N + p]; }
  __device__ __forceinline__ double U(int m, long long p) const { return u[(m - 1) * N + p]; }
  __device__ __forceinline__ double J(int c, long long p) const { return jbn[(c - 1) * N + p]; }
  __device__ __forceinline__ double D(int c, long long p) const { return dvel[(c - 1) * N + p]; }
};

// Convective flux of equation neq along dir (0:e, 1:f, 2:g),
// solver/flujos_compact.f90:21-185.
__device__ __forceinline__ double conv_flux(const FluxIn& F, int neq, int dir,
                                            long long p) {
  switch (neq) {
    case 1:
      return -F.UM(2 + dir, p);
    case 2:
    case 3:
    case 4: {
      const double v = -F.UM(2 + dir, p) * F.U(neq - 1, p);
      return (dir == neq - 2) ? v - F.pres[p] : v;
    }
    default:
      return -F.U(1 + dir, p) * F.UM(5, p);
  }
}

// Viscous flux of equation neq (2..4) along dir (0:ev, 1:fv, 2:gv),
// solver/flujos_compact.f90:54-173.  The 3x3 table of (metric, dvel)
// pairs is symmetric: entry (a,b) couples velocity a and direction b.
__device__ __forceinline__ double visc_flux(const FluxIn& F, int neq, int dir,
                                            long long p) {
  const double mu = F.vis[p];
  const int a = neq - 2;  // velocity component 0..2
  if (a == dir) {
    // 2*mu*metric_a * d(u_a)/d(xi_a): dvel index 1, 5, 9
    const int mc = 1 + 4 * a;
    return mu * (2.0 * (F.J(mc, p) * F.D(1 + 4 * a, p)));
  }
  // off-diagonal pairs, with the reference's operand order:
  //  (0,1),(1,0): J1*D2 + J5*D4    (0,2),(2,0): J1*D3 + J9*D7
  //  (1,2),(2,1): J9*D8 + J5*D6
  const int lo = a < dir ? a : dir, hi = a < dir ? dir : a;
  if (lo == 0 && hi == 1) return mu * (F.J(1, p) * F.D(2, p) + F.J(5, p) * F.D(4, p));
  if (lo == 0 && hi == 2) return mu * (F.J(1, p) * F.D(3, p) + F.J(9, p) * F.D(7, p));
  return mu * (F.J(9, p) * F.D(8, p) + F.J(5, p) * F.D(6, p));
}

// Flux fields already formed (cuDivergence / divergencia input).
struct FluxArr {
  const double* flux;  // 6 comps e,f,g,ev,fv,gv
  long long N;
  __device__ __forceinline__ double conv(int dir, long long p) const { return flux[dir * N + p]; }
  __device__ __forceinline__ double visc(int dir, long long p) const { return flux[(3 + dir) * N + p]; }
};

// Fluxes formed on the fly from the primitive inputs.
struct FluxLive {
  FluxIn in;
  int neq;
  __device__ __forceinline__ double conv(int dir, long long p) const { return conv_flux(in, neq, dir, p); }
  __device__ __forceinline__ double visc(int dir, long long p) const { return visc_flux(in, neq, dir, p); }
};

// Projection E = J*(xi_x*e) (solver/divergencia.f90:228-240) fused with
// the three-pass accumulation rs = (1/J)*((E_xi + F_eta) + G_zeta)
// (:349-351).  Pass 0 writes the accumulators, pass 1 adds, pass 2
// finishes with 1/J.  VISC selects the viscous pair.
template <class Src, bool VISC>
struct ProbDiv {
  static constexpr int NV = VISC ? 2 : 1;
  Src src;
  const double* jbn;
  long long N;
  int dir, pass;
  double* acc;
  double* accv;
  double* rs;
  double* rsv;
  __device__ __forceinline__ void load(long long p, double* v) const {
    const double det = jbn[9 * N + p];
    const double met = jbn[(4 * dir) * N + p];  // comps 1, 5, 9
    v[0] = det * (met * src.conv(dir, p));
    if constexpr (VISC) v[1] = det * (met * src.visc(dir, p));
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
    if (pass == 0) {
      acc[p] = v[0];
      if constexpr (VISC) accv[p] = v[1];
    } else if (pass == 1) {
      acc[p] = acc[p] + v[0];
      if constexpr (VISC) accv[p] = accv[p] + v[1];
    } else {
      const double ij = jbn[10 * N + p];
      rs[p] = ij * (acc[p] + v[0]);
      if constexpr (VISC) rsv[p] = ij * (accv[p] + v[1]);
    }
  }
};

// ---------------------------------------------------------------------------
// Pointwise kernels.

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

__global__ void k_fill(double* a, double v, long long n) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x)
    a[p] = v;
}

__global__ void k_fill_int(int* a, int v) {
  if (threadIdx.x == 0) *a = v;
}

// det / inverse of the metric tensor in place, solver/jacob_malla.f90:127-169.
__global__ void k_metric_invert(double* jbn, long long N, int* bad) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long long)gridDim.x * blockDim.x) {
#define J(c) jbn[((c) - 1) * N + p]
    const double j1 = J(1), j2 = J(2), j3 = J(3), j4 = J(4), j5 = J(5),
                 j6 = J(6), j7 = J(7), j8 = J(8), j9 = J(9);
    const double det = j1 * j5 * j9 - j1 * j6 * j8 + j2 * j6 * j7 -
                       j2 * j4 * j9 + j3 * j4 * j8 - j3 * j5 * j7;
    if (fabs(det) <= 1.e-20) *bad = 1;
    const double r = 1. / det;
    const double t1 = j5 * j9 - j6 * j8;
    const double t4 = j3 * j8 - j2 * j9;
    const double t7 = j2 * j6 - j3 * j5;
    const double t2 = j6 * j7 - j4 * j9;
    const double t5 = j1 * j9 - j3 * j7;
    const double t8 = j3 * j4 - j1 * j6;
    const double t3 = j4 * j8 - j5 * j7;
    const double t6 = j2 * j7 - j1 * j8;
    const double t9 = j1 * j5 - j2 * j4;
    J(10) = det;
    J(11) = r;
    J(1) = t1 * r;
    J(2) = -t2 * r;
    J(3) = t3 * r;
    J(4) = -t4 * r;
    J(5) = t5 * r;
    J(6) = -t6 * r;
    J(7) = t7 * r;
    J(8) = -t8 * r;
    J(9) = t9 * r;
#undef J
  }
}

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

// ---------------------------------------------------------------------------
// Host entry points (device pointers, context stream).

static unsigned pointwise_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

int op_deriv(emx_ctx* ctx, int dir, const double* f, double* df, double ih) {
  ProbDeriv<1> p;
  p.in[0] = f;
  p.out[0] = df;
  return launch_lines(ctx, dir, p, ih);
}

// Derivatives of m independent fields along dir, two fields per launch (a
// line solve holds 2*(2C+6) doubles per field in registers; pairs keep the
// strided kernel within the 128 VGPRs its 1024-thread workgroups allow).
static int deriv_fields(emx_ctx* ctx, int dir, int m, const double* const* in,
                        double* const* out, double ih) {
  int f = 0;
  for (; f + 1 < m; f += 2) {
    ProbDeriv<2> p;
    p.in[0] = in[f];
    p.in[1] = in[f + 1];
    p.out[0] = out[f];
    p.out[1] = out[f + 1];
    EMX_TRY(launch_lines(ctx, dir, p, ih));
  }
  if (f < m) {
    ProbDeriv<1> p;
    p.in[0] = in[f];
    p.out[0] = out[f];
    EMX_TRY(launch_lines(ctx, dir, p, ih));
  }
  return EMX_OK;
}

int op_grad(emx_ctx* ctx, const double* u, const double* temp, double* dvel,
            double* dtemp, const double ih[3]) {
  const long long N = ctx->d.N();
  for (int dir = 0; dir < 3; dir++) {
    const double* in[4] = {u, u + N, u + 2 * N, temp};
    double* out[4] = {dvel + (3 * dir) * N, dvel + (3 * dir + 1) * N,
                      dvel + (3 * dir + 2) * N, temp ? dtemp + dir * N : nullptr};
    EMX_TRY(deriv_fields(ctx, dir, temp ? 4 : 3, in, out, ih[dir]));
  }
  return EMX_OK;
}

int op_flux(emx_ctx* ctx, int neq, const FluxPtrs& fp, double* flux) {
  FluxIn F{fp.um, fp.u, fp.pres, fp.vis, fp.jbn, fp.dvel, ctx->d.N()};
  hipLaunchKernelGGL(k_flujos, dim3(pointwise_blocks(F.N)), dim3(256), 0,
                     ctx->stream, F, neq, flux);
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
  hipLaunchKernelGGL(k_fill, dim3(pointwise_blocks(N)), dim3(256), 0,
                     ctx->stream, rsv, 0.0, N);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

int op_metrics(emx_ctx* ctx, const double* x, double* jbn, int* bad_dev) {
  const long long N = ctx->d.N();
  const double ih[3] = {(double)(ctx->d.nx - 1), (double)(ctx->d.ny - 1),
                        (double)(ctx->d.nz - 1)};
  // jacob_malla.f90:14-122: comp(x_c along dir d) = 1 + 3*c + d
  for (int dir = 0; dir < 3; dir++) {
    const double* in[3] = {x, x + N, x + 2 * N};
    double* out[3] = {jbn + dir * N, jbn + (3 + dir) * N, jbn + (6 + dir) * N};
    EMX_TRY(deriv_fields(ctx, dir, 3, in, out, ih[dir]));
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, bad_dev, 0);
  hipLaunchKernelGGL(k_metric_invert, dim3(pointwise_blocks(N)), dim3(256), 0,
                     ctx->stream, jbn, N, bad_dev);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
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
  hipLaunchKernelGGL(k_sgs_structure, grid, block, 0, ctx->stream, u, dvel,
                     jbn, sgd, d, kz, amut, w);
  EMX_TRY_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sgs_select, grid, block, 0, ctx->stream, w, d, kz, c4,
                     amut, vis, vis5);
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx
