// Shared device definitions of the operator kernels: the line-solver
// problems (what a solve loads and stores), the flux formulas of flujos and
// small pointwise helpers.
#pragma once
#include <cmath>

#include "compact.hpp"
#include "ops.hpp"

namespace emx {

// ---------------------------------------------------------------------------
// Problems for the line solver.  Besides load/store each problem names
// itself (label) and states its algorithmic bytes per grid point
// (bytes_pt: every field the launch must read or write once, 8 B each),
// which the per-launch timing reports.

// NV independent fields in, NV derivatives out.
template <int N_>
struct ProbDeriv {
  static constexpr int NV = N_;
  static constexpr bool PREFETCH = true;
  static const char* label() { return "deriv"; }
  double bytes_pt() const { return 16.0 * N_; }
  const double* in[N_];
  double* out[N_];
  __device__ __forceinline__ void load(long long p, double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) v[a] = in[a][p];
  }
  // Outputs are streamed with nontemporal stores: no kernel of the pass
  // reads them back, and the hint measured 10% faster on the strided kinds.
  __device__ __forceinline__ void store(long long p, const double* v) const {
#pragma unroll
    for (int a = 0; a < N_; a++) __builtin_nontemporal_store(v[a], &out[a][p]);
  }
};

// Velocity gradient (DERIV_VEL, solver/deriv_vel.f90:17-161) fused with the
// velocity divergence in the projection of divergencia
// (solver/divergencia.f90:225-354 with the velocity as the flux):
// pass dir solves u, v, w and J*(m_dir*u_dir); dvel(3*dir+c) is written, the
// fourth field accumulates and the last pass writes
// div = (1/J)*((E_xi + F_eta) + G_zeta).
struct ProbGradDiv {
  static constexpr int NV = 4;
  static const char* label() { return "grad_div"; }
  double bytes_pt() const {
    // u (3), J, metric; dvel (3); accumulator / 1/J and div
    return 8.0 * (5 + 3 + (dir == 0 ? 1 : (dir == 1 ? 2 : 3)));
  }
  const double* u;    // 3 fields
  const double* jbn;  // 11 fields
  double* dvel;       // 9 fields
  double* acc;
  double* div;
  long long N;
  int dir;
  // Every access goes through gref: the batched launches read this problem
  // from a descriptor array, and its pointers would otherwise be generic.
  __device__ __forceinline__ void load(long long p, double* v) const {
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = gref(u + c * N + p);
    v[3] = gref(jbn + 9 * N + p) * (gref(jbn + (4 * dir) * N + p) * v[dir]);
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
#pragma unroll
    for (int c = 0; c < 3; c++) gref(dvel + (3 * dir + c) * N + p) = v[c];
    if (dir == 0) gref(acc + p) = v[3];
    else if (dir == 1) gref(acc + p) = gref(acc + p) + v[3];
    else gref(div + p) = gref(jbn + 10 * N + p) * (gref(acc + p) + v[3]);
  }
  // Q points p + i*sl at once: the accumulator (and 1/J) reads first, as
  // ProbAccum::store_batch.  Same arithmetic as store().
  static constexpr bool BATCH_STORE = true;
  template <int Q>
  __device__ __forceinline__ void store_batch(long long p, long long sl,
                                              const double (*v)[NV]) const {
    double a[Q], q[Q];
    if (dir > 0) {
#pragma unroll
      for (int i = 0; i < Q; i++) {
        a[i] = gref(acc + p + i * sl);
        q[i] = dir == 2 ? gref(jbn + 10 * N + p + i * sl) : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const long long pi = p + i * sl;
#pragma unroll
      for (int c = 0; c < 3; c++) gref(dvel + (3 * dir + c) * N + pi) = v[i][c];
      if (dir == 0) gref(acc + pi) = v[i][3];
      else if (dir == 1) gref(acc + pi) = a[i] + v[i][3];
      else gref(div + pi) = q[i] * (a[i] + v[i][3]);
    }
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

// The inputs of every flux at one point, held in registers (accessors as
// FluxIn, the point argument ignored).  Loads that no flux of the caller
// uses are dropped by the compiler.
struct FluxPoint {
  double um[5], u[3], pres, vis, j[10], d[9];
  __device__ __forceinline__ FluxPoint() {}
  __device__ __forceinline__ FluxPoint(const FluxIn& F, long long p) {
#pragma unroll
    for (int m = 1; m < 5; m++) um[m] = F.UM(m + 1, p);
    um[0] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) u[m] = F.U(m + 1, p);
    pres = F.pres[p];
    vis = F.vis[p];
#pragma unroll
    for (int c = 0; c < 10; c++) j[c] = F.J(c + 1, p);
#pragma unroll
    for (int c = 0; c < 9; c++) d[c] = F.D(c + 1, p);
  }
  __device__ __forceinline__ double UM(int m, long long) const { return um[m - 1]; }
  __device__ __forceinline__ double U(int m, long long) const { return u[m - 1]; }
  __device__ __forceinline__ double J(int c, long long) const { return j[c - 1]; }
  __device__ __forceinline__ double D(int c, long long) const { return d[c - 1]; }
};

__device__ __forceinline__ double pres_of(const FluxIn& F, long long p) { return F.pres[p]; }
__device__ __forceinline__ double pres_of(const FluxPoint& F, long long) { return F.pres; }
__device__ __forceinline__ double vis_of(const FluxIn& F, long long p) { return F.vis[p]; }
__device__ __forceinline__ double vis_of(const FluxPoint& F, long long) { return F.vis; }

// Convective flux of equation neq along dir (0:e, 1:f, 2:g),
// solver/flujos_compact.f90:21-185.
template <class F_>
__device__ __forceinline__ double conv_flux(const F_& F, int neq, int dir, long long p) {
  switch (neq) {
    case 1:
      return -F.UM(2 + dir, p);
    case 2:
    case 3:
    case 4: {
      const double v = -F.UM(2 + dir, p) * F.U(neq - 1, p);
      return (dir == neq - 2) ? v - pres_of(F, p) : v;
    }
    default:
      return -F.U(1 + dir, p) * F.UM(5, p);
  }
}

// Viscous flux of equation neq (2..4) along dir (0:ev, 1:fv, 2:gv),
// solver/flujos_compact.f90:54-173.  The 3x3 table of (metric, dvel)
// pairs is symmetric: entry (a,b) couples velocity a and direction b.
template <class F_>
__device__ __forceinline__ double visc_flux(const F_& F, int neq, int dir, long long p) {
  const double mu = vis_of(F, p);
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
  static const char* label() { return "divergence"; }
  int in_fields(int, bool visc) const { return visc ? 2 : 1; }
};

// Fluxes formed on the fly from the primitive inputs.
struct FluxLive {
  FluxIn in;
  int neq;
  __device__ __forceinline__ double conv(int dir, long long p) const { return conv_flux(in, neq, dir, p); }
  __device__ __forceinline__ double visc(int dir, long long p) const { return visc_flux(in, neq, dir, p); }
  static const char* label() { return "flux_div"; }
  // distinct input fields of the fluxes of neq along dir, besides the
  // metric of dir: um(2+dir) [neq 1]; u(1+dir), um5 [neq 5]; um(2+dir),
  // u(neq-1), pres on the diagonal, and for the viscous flux vis plus one
  // dvel (diagonal) or one more metric and two dvel [neq 2..4]
  int in_fields(int dir, bool visc) const {
    if (neq == 1) return 1;
    if (neq == 5) return 2;
    const bool diag = neq - 2 == dir;
    return 2 + (diag ? 1 : 0) + (visc ? (diag ? 2 : 4) : 0);
  }
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
  static const char* label() { return Src::label(); }
  // fluxes + J + metric (+ 1/J in the last pass), accumulators
  double bytes_pt() const {
    const int in = src.in_fields(dir, VISC) + 2 + (pass == 2 ? 1 : 0);
    return 8.0 * (in + (pass > 0 ? NV : 0) + NV);
  }
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

// All five equations of the RUNGEK loop `DO m=1,nd: flujos(m);
// divergencia(m)` (solver/rungek_ico.f90:83-95) in three directional passes.
// Eight solves per direction: the convective fluxes of equations 1..5 and
// the viscous fluxes of 2..4 (those of 1 and 5 are zero).  G selects the
// group a launch solves: 0 = convective 1..4, 1 = convective 5 and viscous
// 2..4 (four solves each, for the strided kinds), 2 = all eight.
// Accumulator slots: 0..4 convective, 5..7 viscous.
template <int G>
struct ProbDivAll {
  static constexpr int NV = G == 2 ? 8 : 4;
  static constexpr bool TILE = true;  // many inputs per point: form once
  static constexpr int SLOT0 = G == 1 ? 4 : 0;
  FluxIn in;
  int dir, pass;
  double* acc;  // 8 fields
  double* rs;   // 5 fields
  double* rsv;  // 5 fields
  static const char* label() {
    return G == 0 ? "flux_div_all[conv1-4]"
                  : (G == 1 ? "flux_div_all[conv5,visc2-4]" : "flux_div_all[8]");
  }
  // inputs along dir: G=0 um(2+dir), u(1..3), pres, J, metric;
  // G=1 u(1+dir), um5, vis, jbn 1/5/9, five dvel, J; G=2 the union (16);
  // 1/J in the last pass; accumulators
  double bytes_pt() const {
    const int in = (G == 0 ? 7 : (G == 1 ? 12 : 16)) + (pass == 2 ? 1 : 0);
    return 8.0 * (in + (pass > 0 ? NV : 0) + NV);
  }
  __device__ __forceinline__ double val(int slot, long long p) const {
    return slot < 5 ? conv_flux(in, slot + 1, dir, p) : visc_flux(in, slot - 3, dir, p);
  }
  __device__ __forceinline__ void load(long long p, double* v) const {
    const long long N = in.N;
    const double det = in.jbn[9 * N + p];
    const double met = in.jbn[(4 * dir) * N + p];
#pragma unroll
    for (int a = 0; a < NV; a++) v[a] = det * (met * val(SLOT0 + a, p));
  }
  __device__ __forceinline__ void store(long long p, const double* v) const {
    const long long N = in.N;
    if (pass == 0) {
#pragma unroll
      for (int a = 0; a < NV; a++) acc[(SLOT0 + a) * N + p] = v[a];
    } else if (pass == 1) {
#pragma unroll
      for (int a = 0; a < NV; a++) {
        const long long q = (SLOT0 + a) * N + p;
        acc[q] = acc[q] + v[a];
      }
    } else {
      const double ij = in.jbn[10 * N + p];
#pragma unroll
      for (int a = 0; a < NV; a++) {
        const int slot = SLOT0 + a;
        const double r = ij * (acc[slot * N + p] + v[a]);
        if (slot < 5) rs[slot * N + p] = r;
        else rsv[(slot - 4) * N + p] = r;
      }
    }
  }
};

// Two projected fluxes (already J*(metric*flux)) through the derivative
// and the three-pass accumulation of ProbDivAll (same arithmetic): the
// materialized form of the flux divergence.  Two slots per launch keep
// twice the bytes in flight per work-item and read 1/J once for both in the
// last pass (measured 5% faster on the 256^3 stage than one slot).
struct ProbAccum {
  static constexpr int NV = 2;
  static constexpr bool PREFETCH = true;
  static const char* label() { return "flux_div_all[accumulate]"; }
  double bytes_pt() const { return 8.0 * (4 + (pass > 0 ? 2 : 0) + (pass == 2 ? 1 : 0)); }
  const double* f;   // slots s, s+1 (stride N)
  double* acc;       // their accumulators (stride N)
  double* out0;      // written in the last pass
  double* out1;
  const double* ij;  // 1/J
  long long N;
  int pass;
  __device__ __forceinline__ void load(long long p, double* v) const {
    v[0] = f[p];
    v[1] = f[N + p];
  }
  // Q points p + i*sl at once (a work-item's points along its line): the
  // accumulator (and 1/J) reads of all Q are issued before the first
  // write, so their latencies overlap instead of one read-modify-write
  // waiting for the last.  Same arithmetic as store().
  static constexpr bool BATCH_STORE = true;
  template <int Q>
  __device__ __forceinline__ void store_batch(long long p, long long sl,
                                              const double (*v)[NV]) const {
    if (pass == 0) {
#pragma unroll
      for (int i = 0; i < Q; i++) {
        __builtin_nontemporal_store(v[i][0], &acc[p + i * sl]);
        __builtin_nontemporal_store(v[i][1], &acc[N + p + i * sl]);
      }
      return;
    }
    double a0[Q], a1[Q], q[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
      a0[i] = acc[p + i * sl];
      a1[i] = acc[N + p + i * sl];
      q[i] = pass == 2 ? ij[p + i * sl] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
      if (pass == 1) {
        __builtin_nontemporal_store(a0[i] + v[i][0], &acc[p + i * sl]);
        __builtin_nontemporal_store(a1[i] + v[i][1], &acc[N + p + i * sl]);
      } else {
        __builtin_nontemporal_store(q[i] * (a0[i] + v[i][0]), &out0[p + i * sl]);
        __builtin_nontemporal_store(q[i] * (a1[i] + v[i][1]), &out1[p + i * sl]);
      }
    }
  }
  // nontemporal stores, as ProbDeriv::store
  __device__ __forceinline__ void store(long long p, const double* v) const {
    if (pass == 0) {
      __builtin_nontemporal_store(v[0], &acc[p]);
      __builtin_nontemporal_store(v[1], &acc[N + p]);
    } else if (pass == 1) {
      __builtin_nontemporal_store(acc[p] + v[0], &acc[p]);
      __builtin_nontemporal_store(acc[N + p] + v[1], &acc[N + p]);
    } else {
      const double q = ij[p];
      __builtin_nontemporal_store(q * (acc[p] + v[0]), &out0[p]);
      __builtin_nontemporal_store(q * (acc[N + p] + v[1]), &out1[p]);
    }
  }
};

inline unsigned pointwise_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace emx
