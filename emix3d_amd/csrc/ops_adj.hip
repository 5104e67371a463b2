// Adjoints (transposes) of the spatial operator: the compact derivative,
// the velocity gradient, the divergence and the linearised flux
// divergence of all five equations.  The reference has no adjoint; each
// function here is the exact transpose of the forward operator it names
// (same scheme, same metric projection), up to rounding.
#include "ops_common.hpp"

namespace emx {

static __global__ void k_zero(double* a, long long n) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x)
    a[p] = 0.0;
}

// NV fields, each through D^T along the pass direction; with accumulate
// the result is added to out (the passes of a gradient adjoint).
template <int N_>
struct ProbDerivAdj {
  static constexpr int NV = N_;
  static const char* label() { return "deriv_adj"; }
  double bytes_pt() const { return 8.0 * N_ * (accumulate ? 3 : 2); }
  const double* in[N_];
  double* out[N_];
  bool accumulate;
  __device__ __forceinline__ void load(long long p, double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) v[a] = in[a][p];
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) out[a][p] = accumulate ? out[a][p] + v[a] : v[a];
  }
};

// Transpose of the divergence (solver/divergencia.f90:225-354 forward):
// rs = (1/J) sum_d D_d(J m_d e_d) gives e_d_bar = J m_d D_d^T(rs_bar / J),
// and the same for the viscous pair.  flux_bar holds e,f,g,ev,fv,gv.
struct ProbDivAdj {
  static constexpr int NV = 2;
  static const char* label() { return "divergence_adj"; }
  double bytes_pt() const { return 8.0 * 7; }
  const double* rs_bar;
  const double* rsv_bar;
  const double* jbn;
  double* flux_bar;
  long long N;
  int dir;
  __device__ __forceinline__ void load(long long p, double* v) const {
    const double ij = jbn[10 * N + p];
    v[0] = ij * rs_bar[p];
    v[1] = ij * rsv_bar[p];
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
    const double jm = jbn[9 * N + p] * jbn[(4 * dir) * N + p];
    flux_bar[dir * N + p] = jm * v[0];
    flux_bar[(3 + dir) * N + p] = jm * v[1];
  }
};

// Transpose of the linearisation of the five-equation flux divergence
// (flujos_compact.f90:20-217 fluxes, divergencia.f90 projection) about
// (um, u, vis, dvel); pres enters linearly, jbn is fixed.  Pass dir forms
// the eight flux adjoints along dir (convective of 1..5, viscous of 2..4)
// and scatters them to the input adjoints:
//   conv 1   e = -um(2+d)                 -> um(2+d)
//   conv 2-4 e = -um(2+d) u(a+1) [-pres]  -> um(2+d), u(a+1), pres (a = d)
//   conv 5   e = -u(1+d) um5              -> u(1+d), um5
//   visc a=d  mu*2*(J m_a D_a)            -> vis, dvel(1+4a)
//   visc a!=d mu*(J_x D_y + J_z D_w)      -> vis, two dvel
// Pass 0 writes each adjoint, later passes add, except the dvel
// components a pass touches first (first[] below).
struct ProbDivAllAdj {
  static constexpr int NV = 8;
  static const char* label() { return "flux_div_all_adj"; }
  double bytes_pt() const {
    // loads: 8 adjoint fields + 1/J; primal: J, um(2+d), u(1..3), um5,
    // vis, jbn 1/5/9, five dvel; writes: um(2+d), u(1..3), um5, pres, vis,
    // five dvel; read back in later passes: u(1..3), um5, pres, vis and
    // the dvel components touched before
    const int rmw = pass == 0 ? 0 : 6 + (pass == 1 ? 2 : 4);
    return 8.0 * (9 + 10 + 5 + 12 + rmw);
  }
  FluxIn in;
  const double* rs_bar;   // 5 fields
  const double* rsv_bar;  // 5 fields (2..4 used)
  double* um_bar;         // 5 fields (1 is zero)
  double* u_bar;          // 3 fields
  double* pres_bar;
  double* vis_bar;
  double* dvel_bar;       // 9 fields
  int dir, pass;
  __device__ __forceinline__ void load(long long p, double* v) const {
    const long long N = in.N;
    const double ij = in.jbn[10 * N + p];
#pragma unroll
    for (int a = 0; a < 5; a++) v[a] = ij * rs_bar[a * N + p];
#pragma unroll
    for (int a = 0; a < 3; a++) v[5 + a] = ij * rsv_bar[(1 + a) * N + p];
  }
  // The scatter of a work-item's W consecutive points p .. p+W-1 of one
  // line, used where the line is contiguous in memory (the xi pass): in
  // two phases (convective, viscous), each of which reads all its input
  // fields for the W points before it computes and writes its outputs, so
  // every cache line a wavefront touches is used whole by a few consecutive
  // instructions and the loads of a phase overlap.  (Point by point, each
  // line of every field was visited once per point, two dozen fields
  // apart, and was evicted from L2 between the visits.)  Per point the
  // same arithmetic in the same order as store().
  static constexpr bool CHUNK_STORE = true;
  template <int W>
  __device__ __forceinline__ void putw(double* f, const double* v, bool first) const {
    if (first) {
#pragma unroll
      for (int i = 0; i < W; i++) f[i] = v[i];
    } else {
      double o[W];
#pragma unroll
      for (int i = 0; i < W; i++) o[i] = f[i];
#pragma unroll
      for (int i = 0; i < W; i++) f[i] = o[i] + v[i];
    }
  }
  template <int W>
  __device__ __forceinline__ void store_w(long long p, const double (*x)[NV]) const {
    const long long N = in.N;
    const int d = dir;
    const bool first = pass == 0;
    const double* jb = in.jbn + p;
    double jm[W];
#pragma unroll
    for (int i = 0; i < W; i++) jm[i] = jb[9 * N + i] * jb[(4 * d) * N + i];
    // projected flux adjoint k of point i (convective k < 5, viscous 5..7)
    auto e = [&](int i, int k) { return jm[i] * x[i][k]; };
    // --- convective: inputs um(2+d), u(1..3), um5 ---
    {
      double umd[W], uu[3][W], um5[W];
#pragma unroll
      for (int i = 0; i < W; i++) umd[i] = in.um[(1 + d) * N + p + i];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < W; i++) uu[a][i] = in.u[a * N + p + i];
#pragma unroll
      for (int i = 0; i < W; i++) um5[i] = in.um[4 * N + p + i];
      double r[W];
      // um_bar(2+d) = -e1 - sum_a e(2+a) u(1+a): only this pass writes it
#pragma unroll
      for (int i = 0; i < W; i++) {
        double t = -e(i, 0);
#pragma unroll
        for (int a = 0; a < 3; a++) t -= e(i, 1 + a) * uu[a][i];
        r[i] = t;
      }
#pragma unroll
      for (int i = 0; i < W; i++) um_bar[(1 + d) * N + p + i] = r[i];
#pragma unroll
      for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int i = 0; i < W; i++) {
          double g = -e(i, 1 + a) * umd[i];
          if (a == d) g -= e(i, 4) * um5[i];
          r[i] = g;
        }
        putw<W>(u_bar + a * N + p, r, first);
      }
      // u(1+d) by selects (a run-time index would put the arrays in scratch)
#pragma unroll
      for (int i = 0; i < W; i++)
        r[i] = -e(i, 4) * (d == 0 ? uu[0][i] : (d == 1 ? uu[1][i] : uu[2][i]));
      putw<W>(um_bar + 4 * N + p, r, first);
#pragma unroll
      for (int i = 0; i < W; i++) r[i] = -(d == 0 ? e(i, 1) : (d == 1 ? e(i, 2) : e(i, 3)));
      putw<W>(pres_bar + p, r, first);
    }
    // --- viscous: inputs vis, jbn 1/5/9 and the five dvel components of
    // this direction; pass that first touches dvel component m (1-based):
    // 1,2,3,4,7 -> 0; 5,6,8 -> 1; 9 -> 2 ---
    {
      // the two (metric, gradient) pairs of viscous flux a != d, and the
      // diagonal one of a == d (metric jbn 1+4a, gradient 1+4a)
      auto pair = [&](int a, int& jx, int& mx, int& jz, int& mz) {
        const int lo = a < d ? a : d, hi = a < d ? d : a;
        if (lo == 0 && hi == 1) { jx = 1; mx = 2; jz = 5; mz = 4; }       // J1*D2 + J5*D4
        else if (lo == 0) { jx = 1; mx = 3; jz = 9; mz = 7; }             // J1*D3 + J9*D7
        else { jx = 9; mx = 8; jz = 5; mz = 6; }                          // J9*D8 + J5*D6
      };
      double mu[W], J1[W], J5[W], J9[W];
#pragma unroll
      for (int i = 0; i < W; i++) mu[i] = in.vis[p + i];
#pragma unroll
      for (int i = 0; i < W; i++) J1[i] = jb[i];
#pragma unroll
      for (int i = 0; i < W; i++) J5[i] = jb[4 * N + i];
#pragma unroll
      for (int i = 0; i < W; i++) J9[i] = jb[8 * N + i];
      // gradient components: Dp[a][0] the diagonal one of a == d, else
      // Dp[a][0], Dp[a][1] the pair of flux a (indices fixed at compile
      // time; d only chooses which loads run)
      double Dp[3][2][W];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (a == d) {
#pragma unroll
          for (int i = 0; i < W; i++) Dp[a][0][i] = in.dvel[(4 * a) * N + p + i];
        } else {
          int jx, mx, jz, mz;
          pair(a, jx, mx, jz, mz);
#pragma unroll
          for (int i = 0; i < W; i++) Dp[a][0][i] = in.dvel[(mx - 1) * N + p + i];
#pragma unroll
          for (int i = 0; i < W; i++) Dp[a][1][i] = in.dvel[(mz - 1) * N + p + i];
        }
      }
      auto Jv = [&](int jc, int i) { return jc == 1 ? J1[i] : (jc == 5 ? J5[i] : J9[i]); };
      double vb[W], r[W];
#pragma unroll
      for (int i = 0; i < W; i++) vb[i] = 0.0;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (a == d) {
          const int m = 1 + 4 * a, jc = 1 + 4 * a;
          const int f0 = (m == 5 || m == 6 || m == 8) ? 1 : (m == 9 ? 2 : 0);
#pragma unroll
          for (int i = 0; i < W; i++) {
            const double w = e(i, 5 + a);
            vb[i] += w * (2.0 * (Jv(jc, i) * Dp[a][0][i]));
            r[i] = w * mu[i] * 2.0 * Jv(jc, i);
          }
          putw<W>(dvel_bar + (m - 1) * N + p, r, f0 == d);
        } else {
          int jx, mx, jz, mz;
          pair(a, jx, mx, jz, mz);
#pragma unroll
          for (int i = 0; i < W; i++)
            vb[i] += e(i, 5 + a) * (Jv(jx, i) * Dp[a][0][i] + Jv(jz, i) * Dp[a][1][i]);
          const int fx = (mx == 5 || mx == 6 || mx == 8) ? 1 : (mx == 9 ? 2 : 0);
          const int fz = (mz == 5 || mz == 6 || mz == 8) ? 1 : (mz == 9 ? 2 : 0);
#pragma unroll
          for (int i = 0; i < W; i++) r[i] = e(i, 5 + a) * mu[i] * Jv(jx, i);
          putw<W>(dvel_bar + (mx - 1) * N + p, r, fx == d);
#pragma unroll
          for (int i = 0; i < W; i++) r[i] = e(i, 5 + a) * mu[i] * Jv(jz, i);
          putw<W>(dvel_bar + (mz - 1) * N + p, r, fz == d);
        }
      }
      putw<W>(vis_bar + p, vb, first);
    }
  }
  __device__ __forceinline__ void put(double* f, long long p, double v, bool first) const {
    f[p] = first ? v : f[p] + v;
  }
  __device__ __forceinline__ void store(long long p, const double* x) const {
    const long long N = in.N;
    const int d = dir;
    const bool first = pass == 0;
    const double jm = in.jbn[9 * N + p] * in.jbn[(4 * d) * N + p];
    double e[5], ev[3];
#pragma unroll
    for (int k = 0; k < 5; k++) e[k] = jm * x[k];
#pragma unroll
    for (int k = 0; k < 3; k++) ev[k] = jm * x[5 + k];
    // convective
    const double umd = in.UM(2 + d, p);
    double t = -e[0];
#pragma unroll
    for (int a = 0; a < 3; a++) t -= e[1 + a] * in.U(1 + a, p);
    um_bar[(1 + d) * N + p] = t;  // component 2+d: only this pass
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double g = -e[1 + a] * umd;
      if (a == d) g -= e[4] * in.UM(5, p);
      put(u_bar + a * N, p, g, first);
    }
    put(um_bar + 4 * N, p, -e[4] * in.U(1 + d, p), first);
    put(pres_bar, p, -e[1 + d], first);
    // viscous; pass that first touches dvel component m (1-based):
    // 1,2,3,4,7 -> 0; 5,6,8 -> 1; 9 -> 2
    const double mu = in.vis[p];
    const double J1 = in.J(1, p), J5 = in.J(5, p), J9 = in.J(9, p);
    auto dput = [&](int m, double v) {
      const int f0 = (m == 5 || m == 6 || m == 8) ? 1 : (m == 9 ? 2 : 0);
      put(dvel_bar + (m - 1) * N, p, v, f0 == d);
    };
    double vb = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double w = ev[a];
      if (a == d) {
        const int m = 1 + 4 * a;
        const double Jd = a == 0 ? J1 : (a == 1 ? J5 : J9);
        vb += w * (2.0 * (Jd * in.D(m, p)));
        dput(m, w * mu * 2.0 * Jd);
      } else {
        const int lo = a < d ? a : d, hi = a < d ? d : a;
        if (lo == 0 && hi == 1) {  // J1*D2 + J5*D4
          vb += w * (J1 * in.D(2, p) + J5 * in.D(4, p));
          dput(2, w * mu * J1);
          dput(4, w * mu * J5);
        } else if (lo == 0) {      // J1*D3 + J9*D7
          vb += w * (J1 * in.D(3, p) + J9 * in.D(7, p));
          dput(3, w * mu * J1);
          dput(7, w * mu * J9);
        } else {                   // J9*D8 + J5*D6
          vb += w * (J9 * in.D(8, p) + J5 * in.D(6, p));
          dput(8, w * mu * J9);
          dput(6, w * mu * J5);
        }
      }
    }
    put(vis_bar, p, vb, first);
  }
};

int adj_ready(emx_ctx* ctx) {
  for (int d = 0; d < 3; d++)
    if (!ctx->adj[d].blob) {
      set_error("adjoint operators need lines of at most %d points (grid %d x %d x %d)",
                ADJ_MAX_N, ctx->d.nx, ctx->d.ny, ctx->d.nz);
      return EMX_ERR_SIZE;
    }
  return EMX_OK;
}

int op_deriv_adj(emx_ctx* ctx, int dir, const double* g, double* out, double ih) {
  EMX_TRY(adj_ready(ctx));
  ProbDerivAdj<1> p{{g}, {out}, false};
  return launch_lines_adj(ctx, dir, p, ih);
}

int op_grad_adj(emx_ctx* ctx, const double* dvel_bar, const double* dtemp_bar,
                double* u_bar, double* temp_bar, const double ih[3]) {
  EMX_TRY(adj_ready(ctx));
  const long long N = ctx->d.N();
  for (int d = 0; d < 3; d++) {
    if (dtemp_bar) {
      ProbDerivAdj<4> p{{dvel_bar + (3 * d) * N, dvel_bar + (3 * d + 1) * N,
                         dvel_bar + (3 * d + 2) * N, dtemp_bar + d * N},
                        {u_bar, u_bar + N, u_bar + 2 * N, temp_bar}, d > 0};
      EMX_TRY(launch_lines_adj(ctx, d, p, ih[d]));
    } else {
      ProbDerivAdj<3> p{{dvel_bar + (3 * d) * N, dvel_bar + (3 * d + 1) * N,
                         dvel_bar + (3 * d + 2) * N},
                        {u_bar, u_bar + N, u_bar + 2 * N}, d > 0};
      EMX_TRY(launch_lines_adj(ctx, d, p, ih[d]));
    }
  }
  return EMX_OK;
}

int op_divergence_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                      const double* jbn, const double ih[3], double* flux_bar) {
  EMX_TRY(adj_ready(ctx));
  for (int d = 0; d < 3; d++) {
    ProbDivAdj p{rs_bar, rsv_bar, jbn, flux_bar, ctx->d.N(), d};
    EMX_TRY(launch_lines_adj(ctx, d, p, ih[d]));
  }
  return EMX_OK;
}

int op_flux_div_all_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                        const FluxPtrs& fp, const double ih[3], double* um_bar,
                        double* u_bar, double* pres_bar, double* vis_bar,
                        double* dvel_bar) {
  EMX_TRY(adj_ready(ctx));
  const long long N = ctx->d.N();
  const FluxIn in{fp.um, fp.u, nullptr, fp.vis, fp.jbn, fp.dvel, N};
  for (int d = 0; d < 3; d++) {
    ProbDivAllAdj p{in, rs_bar, rsv_bar, um_bar, u_bar, pres_bar, vis_bar, dvel_bar, d, d};
    EMX_TRY(launch_lines_adj(ctx, d, p, ih[d]));
  }
  // um(1) does not enter any flux
  EMX_TIMED(ctx, "fill", 8.0 * N,
            hipLaunchKernelGGL(k_zero, dim3(pointwise_blocks(N)), dim3(256), 0,
                               ctx->stream, um_bar, N));
  EMX_TRY_HIP(hipGetLastError());
  return EMX_OK;
}

}  // namespace emx
