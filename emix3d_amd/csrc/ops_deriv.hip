// Compact derivatives of fields: single derivative, velocity/temperature
// gradient (DERIV_VEL) and grid metrics (JACOBEANO).
#include "ops_common.hpp"

namespace emx {

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

int op_deriv(emx_ctx* ctx, int dir, const double* f, double* df, double ih) {
  ProbDeriv<1> p;
  p.in[0] = f;
  p.out[0] = df;
  return launch_lines(ctx, dir, p, ih);
}

// Derivatives of m independent fields along dir, two fields per launch (a
// line solve holds 2*(2C+6) doubles per field in registers; pairs keep the
// strided kernel within the 128 VGPRs its 1024-thread workgroups allow).
// Pairs keep twice the bytes in flight per work-item: measured faster than
// one field per launch on every kind with a working set beyond the L3.
static int deriv_fields(emx_ctx* ctx, int dir, int m, const double* const* in,
                        double* const* out, double ih) {
  int f = 0;
  if (m == 4 && ctx->plan[PLAN_DERIV][dir].kind == KIND_WAVE) {
    // the wavefront kind (one line per wave, 256-item workgroups) holds
    // four fields per line: one launch for the gradient's four
    ProbDeriv<4> p;
    for (int a = 0; a < 4; a++) {
      p.in[a] = in[a];
      p.out[a] = out[a];
    }
    return launch_lines(ctx, dir, p, ih);
  }
  for (; f + 1 < m; f += 2) {
    ProbDeriv<2> p;
    p.in[0] = in[f];
    p.in[1] = in[f + 1];
    p.out[0] = out[f];
    p.out[1] = out[f + 1];
    EMX_TRY(launch_lines(ctx, dir, p, ih));
  }
  for (; f < m; f++) {
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

int op_grad_div(emx_ctx* ctx, const double* u, const double* jbn, double* dvel, double* div,
                const double ih[3]) {
  const long long N = ctx->d.N();
  for (int d = 0; d < 3; d++) {
    ProbGradDiv p{u, jbn, dvel, ctx->acc, div, N, d};
    EMX_TRY(launch_lines(ctx, d, p, ih[d]));
  }
  return EMX_OK;
}

}  // namespace emx
