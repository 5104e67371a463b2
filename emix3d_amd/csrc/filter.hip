// Compact filter (FILTRADO, src/MSVSingle/herramientas_incomp2.f90:227-410)
// on the device: each conservative field times J is filtered along the
// three directions in the order itf selects, each pass one tridiagonal
// solve per line (FILTRONPER, src/MSVSingle/filtro_incomp.f90:307-362) with
// an 11-point explicit right-hand side, then times 1/J; u, temp and pres are
// recomputed from the result.  The solves run on the line-solver kernels of
// the derivatives (chunked Thomas with precomputed carries), with the
// filter's matrix in the line table and a 5-point halo.
#include "ops_common.hpp"

namespace emx {

// One field filtered in place along a line: J*f formed at load, the
// filtered value times 1/J stored (a nontemporal store: the pass does not
// read it back, and the hint measured 5% faster over FILTRADO).
struct ProbFilter {
  static constexpr int NV = 1;
  static const char* label() { return "filter"; }
  double bytes_pt() const { return 8.0 * 4; }  // f, J, 1/J in; f out
  double* f;
  const double* jac;
  const double* ijac;
  __device__ __forceinline__ void load(long long p, double* v) const { v[0] = f[p] * jac[p]; }
  __device__ __forceinline__ void store(long long p, const double* v) const {
    __builtin_nontemporal_store(v[0] * ijac[p], &f[p]);
  }
  // Q points p + i*sl at once, the 1/J reads first (each would otherwise
  // wait for the store before it)
  static constexpr bool BATCH_STORE = true;
  template <int Q>
  __device__ __forceinline__ void store_batch(long long p, long long sl,
                                              const double (*v)[NV]) const {
    double q[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) q[i] = ijac[p + i * sl];
#pragma unroll
    for (int i = 0; i < Q; i++) __builtin_nontemporal_store(v[i][0] * q[i], &f[p + i * sl]);
  }
};

namespace {

// xi: one wavefront per line, the shortest chunk that covers the line;
// eta, zeta: 32 lines per wavefront, chunks of 16 points.
int filter_shape(int dir, int n, int* kind, int* C) {
  if (dir == 0) {
    int c = 1;
    while (c * 64 < n) c *= 2;
    *kind = KIND_WAVE;
    *C = c;
  } else {
    *kind = KIND_STRIDED32;
    *C = 16;
  }
  return EMX_OK;
}

template <int S>
hipError_t launch_filter(const ProbFilter& p, const LineTab& T, const DirGeom& g, double cfilt,
                         hipStream_t st) {
  if (T.kind == KIND_WAVE) {
    switch (T.C) {
      case 1: return launch_wave_c<S, 1>(p, T, g, cfilt, st);
      case 2: return launch_wave_c<S, 2>(p, T, g, cfilt, st);
      case 4: return launch_wave_c<S, 4>(p, T, g, cfilt, st);
      case 8: return launch_wave_c<S, 8>(p, T, g, cfilt, st);
      default: return launch_wave_c<S, 16>(p, T, g, cfilt, st);
    }
  }
  return launch_strided_c<S, 16, 32>(p, T, g, cfilt, st);
}

}  // namespace

int op_filtrado(emx_ctx* ctx, int itf, double* um1, const double* jbn, double cfilt,
                double mach, double* u, double* pres, double* temp) {
  // direction order of each itf (herramientas_incomp2.f90:245-392)
  static const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const long long N = ctx->d.N();
  LineTab tab[3];
  for (int dir = 0; dir < 3; dir++) {
    const int n = dir == 0 ? ctx->d.nx : (dir == 1 ? ctx->d.ny : ctx->d.nz);
    int kind, C;
    EMX_TRY(filter_shape(dir, n, &kind, &C));
    EMX_TRY(get_line_tab(dir == 2 ? FILTER_P : FILTER_NP, n, C, kind, &tab[dir], cfilt));
  }
  for (int k = 0; k < 3; k++) {
    const int dir = order[itf - 1][k];
    const DirGeom g = dir_geom(ctx->d, dir);
    for (int m = 0; m < 5; m++) {
      ProbFilter p{um1 + m * N, jbn + 9 * N, jbn + 10 * N};
      LaunchTimer timer(ctx, ProbFilter::label(), dir, p.bytes_pt() * (double)N);
      const hipError_t e = dir == 2 ? launch_filter<FILTER_P>(p, tab[dir], g, cfilt, ctx->stream)
                                    : launch_filter<FILTER_NP>(p, tab[dir], g, cfilt, ctx->stream);
      if (e != hipSuccess) {
        set_error("filter launch (dir %d) failed: %s", dir, hipGetErrorString(e));
        return EMX_ERR_HIP;
      }
    }
  }
  // u, temp, pres from the filtered variables (herramientas_incomp2.f90:395-405)
  return op_varc_to_varnc(ctx, um1, mach, u, pres, temp, nullptr);
}

}  // namespace emx
