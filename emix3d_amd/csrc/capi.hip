// extern "C" boundary of libemix3d_hip.so (declared in include/emix3d.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "compact.hpp"
#include "ops.hpp"

using namespace emx;

namespace {

int check_ctx(const emx_ctx* ctx) {
  if (!ctx) {
    set_error("null context");
    return EMX_ERR_ARG;
  }
  return EMX_OK;
}

#define CHECK_PTR(p)                            \
  do {                                          \
    if (!(p)) {                                 \
      set_error("%s: null pointer " #p, __func__); \
      return EMX_ERR_ARG;                       \
    }                                           \
  } while (0)

int make_plan(emx_ctx* ctx, int op, int dir, int kind, int C) {
  const int n = dir == 0 ? ctx->d.nx : (dir == 1 ? ctx->d.ny : ctx->d.nz);
  if (!valid_plan(op, kind, C)) {
    set_error("line plan kind %d C %d is not available for operation class %d",
              kind, C, op);
    return EMX_ERR_ARG;
  }
  LineTab t{};
  if (kind != KIND_MATERIALIZED)  // that form solves with the derivative plan
    EMX_TRY(get_line_tab(ctx->scheme, n, C, kind, &t));
  ctx->plan[op][dir].tab = t;
  ctx->plan[op][dir].kind = kind;
  ctx->plan[op][dir].C = C;
  return EMX_OK;
}

// Grow the projected-flux scratch to 8 fields per direction on the
// materialized kind of the flux-divergence plan.
int ensure_fluxbuf(emx_ctx* c) {
  int need = 0;
  for (int dir = 0; dir < 3; dir++)
    if (c->plan[PLAN_FLUX][dir].kind == KIND_MATERIALIZED) need++;
  if (need <= c->fluxbuf_dirs) return EMX_OK;
  const size_t N = (size_t)c->d.N();
  if (c->fluxbuf) EMX_TRY_HIP(hipFree(c->fluxbuf));
  c->fluxbuf = nullptr;
  c->fluxbuf_dirs = 0;
  if (hipMalloc(&c->fluxbuf, (size_t)need * 8 * N * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    c->fluxbuf = nullptr;
    set_error("device allocation of %d x 8 projected-flux fields of %zu points failed", need, N);
    return EMX_ERR_NOMEM;
  }
  c->fluxbuf_dirs = need;
  return EMX_OK;
}

void free_ctx(emx_ctx* c) {
  if (!c) return;
  (void)hipFree(c->acc);
  (void)hipFree(c->sgs_tmp);
  (void)hipFree(c->fluxbuf);
  (void)hipFree(c->mlp_pack);
  (void)hipFree(c->ls_buf);
  (void)hipFree(c->stage);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  delete c;
}

}  // namespace

extern "C" {

const char* emx_last_error(void) { return emx::last_error(); }

int emx_version(void) { return EMX_VERSION; }

int emx_create(emx_ctx** out, int nx, int ny, int nz, int scheme,
               void* stream) {
  if (!out) {
    set_error("emx_create: null output pointer");
    return EMX_ERR_ARG;
  }
  *out = nullptr;
  if (scheme < EMX_SCHEME_C4 || scheme > EMX_SCHEME_C6C) {
    set_error("emx_create: unknown scheme %d", scheme);
    return EMX_ERR_ARG;
  }
  const int nmin = scheme_min_n(scheme);
  if (nx < nmin || ny < nmin || nz < nmin) {
    set_error("emx_create: grid %dx%dx%d below the %d points per line the "
              "scheme's boundary stencils need", nx, ny, nz, nmin);
    return EMX_ERR_SIZE;
  }
  if (nx > 1024 || ny > 512 || nz > 512) {
    set_error("emx_create: grid %dx%dx%d above the supported line lengths "
              "(xi <= 1024, eta/zeta <= 512)", nx, ny, nz);
    return EMX_ERR_SIZE;
  }
  emx_ctx* c = new emx_ctx();
  c->scheme = scheme;
  c->d = Dims{nx, ny, nz};
  c->stream = (hipStream_t)stream;
  if (hipGetDevice(&c->device) != hipSuccess) {
    set_error("emx_create: no HIP device");
    delete c;
    return EMX_ERR_HIP;
  }
  for (int op = 0; op < PLAN_OPS; op++)
    for (int dir = 0; dir < 3; dir++) {
      const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
      int kind, C;
      default_plan(op, dir, n, &kind, &C);
      int rc = make_plan(c, op, dir, kind, C);
      if (rc != EMX_OK) {
        free_ctx(c);
        return rc;
      }
    }
  // tables of the transposed operator (adjoint entry points), for lines
  // the transposed solve holds
  for (int dir = 0; dir < 3; dir++) {
    const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    if (n > ADJ_MAX_N) continue;
    int rc = get_line_tab(scheme | TAB_TRANSPOSED, n, adj_chunk(n), KIND_STRIDED16,
                          &c->adj[dir]);
    if (rc != EMX_OK) {
      free_ctx(c);
      return rc;
    }
  }
  const size_t N = (size_t)c->d.N();
  c->ls_bytes = ls_scratch_doubles(c->d) * sizeof(double);
  if (hipMalloc(&c->acc, 8 * N * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->sgs_tmp, 2 * N * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->mlp_pack, mlp_pack_elems() * 2) != hipSuccess ||
      hipMalloc(&c->ls_buf, c->ls_bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("emx_create: device allocation of scratch for %zu points failed", N);
    free_ctx(c);
    return EMX_ERR_NOMEM;
  }
  if (int rc = ensure_fluxbuf(c); rc != EMX_OK) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return EMX_OK;
}

int emx_destroy(emx_ctx* ctx) {
  EMX_TRY(check_ctx(ctx));
  free_ctx(ctx);
  return EMX_OK;
}

int emx_set_stream(emx_ctx* ctx, void* stream) {
  EMX_TRY(check_ctx(ctx));
  ctx->stream = (hipStream_t)stream;
  return EMX_OK;
}

int emx_set_plan(emx_ctx* ctx, int op, int dir, int kind, int C) {
  EMX_TRY(check_ctx(ctx));
  if (dir < 0 || dir > 2 || op < 0 || op >= PLAN_OPS) {
    set_error("emx_set_plan: bad operation class %d or direction %d", op, dir);
    return EMX_ERR_ARG;
  }
  const int old_kind = ctx->plan[op][dir].kind, old_C = ctx->plan[op][dir].C;
  EMX_TRY(make_plan(ctx, op, dir, kind, C));
  if (op == PLAN_FLUX) {
    if (int rc = ensure_fluxbuf(ctx); rc != EMX_OK) {
      (void)make_plan(ctx, op, dir, old_kind, old_C);
      return rc;
    }
  }
  return EMX_OK;
}

int emx_get_plan(const emx_ctx* ctx, int op, int dir, int* kind, int* C) {
  EMX_TRY(check_ctx(ctx));
  if (dir < 0 || dir > 2 || op < 0 || op >= PLAN_OPS || !kind || !C) {
    set_error("emx_get_plan: bad arguments");
    return EMX_ERR_ARG;
  }
  *kind = ctx->plan[op][dir].kind;
  *C = ctx->plan[op][dir].C;
  return EMX_OK;
}

int emx_timing_begin(emx_ctx* ctx) {
  EMX_TRY(check_ctx(ctx));
  ctx->timed.clear();
  ctx->timing_summary.clear();
  ctx->timing = true;
  return EMX_OK;
}

int emx_timing_select(emx_ctx* ctx, const char* label) {
  EMX_TRY(check_ctx(ctx));
  ctx->timing_label = label ? label : "";
  return EMX_OK;
}

int emx_timing_end(emx_ctx* ctx, int* n_entries) {
  EMX_TRY(check_ctx(ctx));
  const bool was = ctx->timing;
  ctx->timing = false;
  ctx->timing_summary.clear();
  if (!ctx->timed.empty()) {
    EMX_TRY_HIP(hipEventSynchronize(ctx->ev_pool[2 * ctx->timed.size() - 1]));
    for (size_t i = 0; i < ctx->timed.size(); i++) {
      float ms = 0.f;
      EMX_TRY_HIP(hipEventElapsedTime(&ms, ctx->ev_pool[2 * i], ctx->ev_pool[2 * i + 1]));
      const emx::TimedLaunch& t = ctx->timed[i];
      emx::TimingEntry* e = nullptr;
      for (auto& x : ctx->timing_summary)
        if (x.dir == t.dir && std::strcmp(x.label, t.label) == 0) e = &x;
      if (!e) {
        ctx->timing_summary.push_back(emx::TimingEntry{t.label, t.dir, 0, 0.0, 0.0});
        e = &ctx->timing_summary.back();
      }
      e->launches++;
      e->ms += ms;
      e->bytes += t.bytes;
    }
    std::sort(ctx->timing_summary.begin(), ctx->timing_summary.end(),
              [](const emx::TimingEntry& a, const emx::TimingEntry& b) { return a.ms > b.ms; });
  }
  ctx->timed.clear();
  if (n_entries) *n_entries = (int)ctx->timing_summary.size();
  if (!was) {
    set_error("emx_timing_end: timing was not running (or an event failed)");
    return EMX_ERR_ARG;
  }
  return EMX_OK;
}

int emx_timing_entry(const emx_ctx* ctx, int i, const char** label, int* dir,
                     int* launches, double* total_ms, double* bytes) {
  EMX_TRY(check_ctx(ctx));
  if (i < 0 || i >= (int)ctx->timing_summary.size()) {
    set_error("emx_timing_entry: index %d out of range", i);
    return EMX_ERR_ARG;
  }
  const emx::TimingEntry& e = ctx->timing_summary[i];
  if (label) *label = e.label;
  if (dir) *dir = e.dir;
  if (launches) *launches = e.launches;
  if (total_ms) *total_ms = e.ms;
  if (bytes) *bytes = e.bytes;
  return EMX_OK;
}

int emx_sync(emx_ctx* ctx) {
  EMX_TRY(check_ctx(ctx));
  EMX_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return EMX_OK;
}

int emx_deriv(emx_ctx* ctx, int dir, const double* f, double* df,
              double inv_h) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(f);
  CHECK_PTR(df);
  if (dir < 0 || dir > 2) {
    set_error("emx_deriv: bad direction %d", dir);
    return EMX_ERR_ARG;
  }
  if (f == df) {
    set_error("emx_deriv: output aliases input");
    return EMX_ERR_ARG;
  }
  return op_deriv(ctx, dir, f, df, inv_h);
}

int emx_grad(emx_ctx* ctx, const double* u, const double* temp, double* dvel,
             double* dtemp, double ihx, double ihy, double ihz) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(u);
  CHECK_PTR(dvel);
  if ((temp == nullptr) != (dtemp == nullptr)) {
    set_error("emx_grad: temp and dtemp must both be given or both be NULL");
    return EMX_ERR_ARG;
  }
  const double ih[3] = {ihx, ihy, ihz};
  return op_grad(ctx, u, temp, dvel, dtemp, ih);
}

int emx_grad_div(emx_ctx* ctx, const double* u, const double* jbn, double* dvel, double* div,
                 double ihx, double ihy, double ihz) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(u);
  CHECK_PTR(jbn);
  CHECK_PTR(dvel);
  CHECK_PTR(div);
  const double ih[3] = {ihx, ihy, ihz};
  return op_grad_div(ctx, u, jbn, dvel, div, ih);
}

int emx_metrics(emx_ctx* ctx, const double* x, double* jbn) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(x);
  CHECK_PTR(jbn);
  int* bad = reinterpret_cast<int*>(ctx->sgs_tmp);
  EMX_TRY(op_metrics(ctx, x, jbn, bad));
  int hbad = 0;
  EMX_TRY_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost,
                             ctx->stream));
  EMX_TRY_HIP(hipStreamSynchronize(ctx->stream));
  if (hbad) {
    set_error("emx_metrics: zero Jacobian (|J| <= 1e-20) in the grid");
    return EMX_ERR_METRIC;
  }
  return EMX_OK;
}

int emx_sgs_geometry(emx_ctx* ctx, const double* x, double* sgd,
                     double* dmins, double* deltamax) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(x);
  CHECK_PTR(sgd);
  if (ctx->d.nx < 2 || ctx->d.ny < 2 || ctx->d.nz < 2) {
    set_error("emx_sgs_geometry: need at least 2 points per direction");
    return EMX_ERR_SIZE;
  }
  // one partial maximum per block, in the 4N-double scratch: nb <= N/256+1
  // blocks always fit
  const long long N = ctx->d.N();
  const int nb = (int)std::min<long long>(1024, (N + 255) / 256);
  double* bm = ctx->sgs_tmp;
  EMX_TRY(op_sgs_geometry(ctx, x, sgd, dmins, bm, nb));
  if (deltamax) {
    double h[1024];
    EMX_TRY_HIP(hipMemcpyAsync(h, bm, nb * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    EMX_TRY_HIP(hipStreamSynchronize(ctx->stream));
    double m = 0.0;
    for (int b = 0; b < nb; b++) m = h[b] > m ? h[b] : m;
    *deltamax = m;
  }
  return EMX_OK;
}

int emx_sgdm(emx_ctx* ctx, const double* u, const double* dvel,
             const double* jbn, const double* sgd, double c4, double* amut,
             double* vis, double* vis5) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(u);
  CHECK_PTR(dvel);
  CHECK_PTR(jbn);
  CHECK_PTR(sgd);
  CHECK_PTR(amut);
  CHECK_PTR(vis);
  if (ctx->d.nx < 2 || ctx->d.ny < 2 || ctx->d.nz < 2) {
    set_error("emx_sgdm: need at least 2 points per direction");
    return EMX_ERR_SIZE;
  }
  return op_sgdm(ctx, u, dvel, jbn, sgd, c4, amut, vis, vis5);
}

int emx_flux(emx_ctx* ctx, int neq, const double* um, const double* u,
             const double* pres, const double* vis, const double* jbn,
             const double* dvel, double* flux) {
  EMX_TRY(check_ctx(ctx));
  if (neq < 1 || neq > 5) {
    set_error("emx_flux: equation index %d outside 1..5", neq);
    return EMX_ERR_ARG;
  }
  CHECK_PTR(um);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(vis);
  CHECK_PTR(jbn);
  CHECK_PTR(dvel);
  CHECK_PTR(flux);
  return op_flux(ctx, neq, FluxPtrs{um, u, pres, vis, jbn, dvel}, flux);
}

int emx_divergence(emx_ctx* ctx, const double* flux, const double* jbn,
                   double ihx, double ihy, double ihz, double* rs,
                   double* rsv) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(flux);
  CHECK_PTR(jbn);
  CHECK_PTR(rs);
  CHECK_PTR(rsv);
  const double ih[3] = {ihx, ihy, ihz};
  return op_divergence(ctx, flux, jbn, ih, rs, rsv);
}

int emx_flux_div(emx_ctx* ctx, int neq, const double* um, const double* u,
                 const double* pres, const double* vis, const double* jbn,
                 const double* dvel, double ihx, double ihy, double ihz,
                 double* rs, double* rsv) {
  EMX_TRY(check_ctx(ctx));
  if (neq < 1 || neq > 5) {
    set_error("emx_flux_div: equation index %d outside 1..5", neq);
    return EMX_ERR_ARG;
  }
  CHECK_PTR(um);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(jbn);
  CHECK_PTR(rs);
  CHECK_PTR(rsv);
  if (neq >= 2 && neq <= 4) {
    CHECK_PTR(vis);
    CHECK_PTR(dvel);
  }
  const double ih[3] = {ihx, ihy, ihz};
  return op_flux_div(ctx, neq, FluxPtrs{um, u, pres, vis, jbn, dvel}, ih, rs, rsv);
}

int emx_flux_div_all(emx_ctx* ctx, const double* um, const double* u,
                     const double* pres, const double* vis, const double* jbn,
                     const double* dvel, double ihx, double ihy, double ihz,
                     double* rs, double* rsv) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(um);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(vis);
  CHECK_PTR(jbn);
  CHECK_PTR(dvel);
  CHECK_PTR(rs);
  CHECK_PTR(rsv);
  const double ih[3] = {ihx, ihy, ihz};
  return op_flux_div_all(ctx, FluxPtrs{um, u, pres, vis, jbn, dvel}, ih, rs, rsv);
}

int emx_rk_update(emx_ctx* ctx, int sub, double* um1, const double* um0, const double* um,
                  const double* rs, const double* rsv, double dt) {
  EMX_TRY(check_ctx(ctx));
  if (sub < 1 || sub > 3) {
    set_error("emx_rk_update: sub-step %d outside 1..3", sub);
    return EMX_ERR_ARG;
  }
  CHECK_PTR(um1);
  CHECK_PTR(um0);
  CHECK_PTR(rs);
  CHECK_PTR(rsv);
  if (sub > 1 && !um) {
    set_error("emx_rk_update: sub-step %d needs um", sub);
    return EMX_ERR_ARG;
  }
  return op_rk_update(ctx, sub, um1, um0, um, rs, rsv, dt);
}

int emx_heviside(emx_ctx* ctx, const double* phi, double deltamax, double* hevi) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(phi);
  CHECK_PTR(hevi);
  if (!(deltamax > 0.0 && std::isfinite(deltamax))) {
    set_error("emx_heviside: deltamax must be positive and finite (got %g)", deltamax);
    return EMX_ERR_ARG;
  }
  return op_ls_heviside(ctx, phi, deltamax, hevi);
}

int emx_ls_reinit(emx_ctx* ctx, double* phi, const double* jbn, const double* dmin, double dx,
                  double dy, double dz, double deltamin) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(phi);
  CHECK_PTR(jbn);
  CHECK_PTR(dmin);
  if (!(dx > 0.0 && dy > 0.0 && dz > 0.0 && deltamin > 0.0)) {
    set_error("emx_ls_reinit: spacings and deltamin must be positive");
    return EMX_ERR_ARG;
  }
  return op_ls_reinit(ctx, phi, jbn, dmin, dx, dy, dz, deltamin);
}

int emx_frontera_val(emx_ctx* ctx, double* um1) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(um1);
  return op_frontera_val(ctx, um1);
}

static int check_mach(double mach, const char* who) {
  if (!(mach > 0.0)) {
    set_error("%s: Mach number %g must be positive", who, mach);
    return EMX_ERR_ARG;
  }
  return EMX_OK;
}

int emx_varc_to_varnc(emx_ctx* ctx, const double* um1, double mach, double* u, double* pres,
                      double* temp, double* conc) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(um1);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(temp);
  EMX_TRY(check_mach(mach, "emx_varc_to_varnc"));
  return op_varc_to_varnc(ctx, um1, mach, u, pres, temp, conc);
}

int emx_varnc_to_varc(emx_ctx* ctx, const double* u, const double* pres, const double* temp,
                      double mach, double* um0, double* pres0) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(temp);
  CHECK_PTR(um0);
  EMX_TRY(check_mach(mach, "emx_varnc_to_varc"));
  return op_varnc_to_varc(ctx, u, pres, temp, mach, um0, pres0);
}

int emx_rk_clip(emx_ctx* ctx, double* um1, const double* mask, double mach, double* u,
                double* pres, double* temp, double* conc) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(um1);
  CHECK_PTR(mask);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(temp);
  EMX_TRY(check_mach(mach, "emx_rk_clip"));
  return op_rk_clip(ctx, um1, mask, mach, u, pres, temp, conc);
}

int emx_rk_rebuild(emx_ctx* ctx, double* um1, const double* u, const double* pres, double mach) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(um1);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  EMX_TRY(check_mach(mach, "emx_rk_rebuild"));
  return op_rk_rebuild(ctx, um1, u, pres, mach);
}

int emx_filtrado(emx_ctx* ctx, int itf, double* um1, const double* jbn, double cfilt,
                 double mach, double* u, double* pres, double* temp) {
  EMX_TRY(check_ctx(ctx));
  if (itf < 1 || itf > 6) {
    set_error("emx_filtrado: direction order itf %d outside 1..6", itf);
    return EMX_ERR_ARG;
  }
  if (!(cfilt >= 0.0 && cfilt < 0.5)) {
    set_error("emx_filtrado: cfilt %g outside [0, 0.5)", cfilt);
    return EMX_ERR_ARG;
  }
  CHECK_PTR(um1);
  CHECK_PTR(jbn);
  CHECK_PTR(u);
  CHECK_PTR(pres);
  CHECK_PTR(temp);
  EMX_TRY(check_mach(mach, "emx_filtrado"));
  return op_filtrado(ctx, itf, um1, jbn, cfilt, mach, u, pres, temp);
}

int emx_dynsmag(emx_ctx* ctx, const double* u, const double* dvel, const double* jbn,
                const double* delta, double* nut, double c4, double* vis) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(u);
  CHECK_PTR(dvel);
  CHECK_PTR(jbn);
  CHECK_PTR(delta);
  CHECK_PTR(nut);
  return op_dynsmag(ctx, u, dvel, jbn, delta, nut, c4, vis);
}

int emx_mlp_nut(emx_ctx* ctx, const float* params, const float* norm, const double* dvel,
                const double* jbn, const double* delta, double* nut, double c4, double* vis) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(params);
  CHECK_PTR(norm);
  CHECK_PTR(dvel);
  CHECK_PTR(jbn);
  CHECK_PTR(delta);
  CHECK_PTR(nut);
  return op_mlp_nut(ctx, params, norm, dvel, jbn, delta, nut, c4, vis);
}

int emx_mlp_gather(emx_ctx* ctx, const double* dvel, const double* jbn, const double* delta,
                   const long long* idx, int n, float* feat) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(dvel);
  CHECK_PTR(jbn);
  CHECK_PTR(delta);
  CHECK_PTR(idx);
  CHECK_PTR(feat);
  if (n <= 0) {
    set_error("emx_mlp_gather: n = %d", n);
    return EMX_ERR_ARG;
  }
  return op_mlp_gather(ctx, dvel, jbn, delta, idx, n, feat);
}

int emx_deriv_adj(emx_ctx* ctx, int dir, const double* g, double* out,
                  double inv_h) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(g);
  CHECK_PTR(out);
  if (dir < 0 || dir > 2) {
    set_error("emx_deriv_adj: bad direction %d", dir);
    return EMX_ERR_ARG;
  }
  if (g == out) {
    set_error("emx_deriv_adj: output aliases input");
    return EMX_ERR_ARG;
  }
  return op_deriv_adj(ctx, dir, g, out, inv_h);
}

int emx_grad_adj(emx_ctx* ctx, const double* dvel_bar, const double* dtemp_bar,
                 double* u_bar, double* temp_bar, double ihx, double ihy,
                 double ihz) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(dvel_bar);
  CHECK_PTR(u_bar);
  if ((dtemp_bar == nullptr) != (temp_bar == nullptr)) {
    set_error("emx_grad_adj: dtemp_bar and temp_bar must both be given or both be NULL");
    return EMX_ERR_ARG;
  }
  const double ih[3] = {ihx, ihy, ihz};
  return op_grad_adj(ctx, dvel_bar, dtemp_bar, u_bar, temp_bar, ih);
}

int emx_divergence_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                       const double* jbn, double ihx, double ihy, double ihz,
                       double* flux_bar) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(rs_bar);
  CHECK_PTR(rsv_bar);
  CHECK_PTR(jbn);
  CHECK_PTR(flux_bar);
  const double ih[3] = {ihx, ihy, ihz};
  return op_divergence_adj(ctx, rs_bar, rsv_bar, jbn, ih, flux_bar);
}

int emx_flux_div_all_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                         const double* um, const double* u, const double* vis,
                         const double* jbn, const double* dvel, double ihx,
                         double ihy, double ihz, double* um_bar, double* u_bar,
                         double* pres_bar, double* vis_bar, double* dvel_bar) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(rs_bar);
  CHECK_PTR(rsv_bar);
  CHECK_PTR(um);
  CHECK_PTR(u);
  CHECK_PTR(vis);
  CHECK_PTR(jbn);
  CHECK_PTR(dvel);
  CHECK_PTR(um_bar);
  CHECK_PTR(u_bar);
  CHECK_PTR(pres_bar);
  CHECK_PTR(vis_bar);
  CHECK_PTR(dvel_bar);
  const double ih[3] = {ihx, ihy, ihz};
  FluxPtrs fp{um, u, nullptr, vis, jbn, dvel};
  return op_flux_div_all_adj(ctx, rs_bar, rsv_bar, fp, ih, um_bar, u_bar, pres_bar,
                             vis_bar, dvel_bar);
}

int emx_ls_extrapolate(emx_ctx* ctx, double* pres, const double* jbn,
                       const double* sobject, double dx, double dy,
                       double dz, double deltamin) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(pres);
  CHECK_PTR(jbn);
  CHECK_PTR(sobject);
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0)) {
    set_error("emx_ls_extrapolate: spacings must be positive");
    return EMX_ERR_ARG;
  }
  return op_ls_extrapolate(ctx, pres, jbn, sobject, dx, dy, dz, deltamin);
}

int emx_ls_advect_weno5(emx_ctx* ctx, double* phi, const double* vel,
                        const double* jbn, double dx, double dy, double dz,
                        double dt) {
  EMX_TRY(check_ctx(ctx));
  CHECK_PTR(phi);
  CHECK_PTR(vel);
  CHECK_PTR(jbn);
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0)) {
    set_error("emx_ls_advect_weno5: spacings must be positive");
    return EMX_ERR_ARG;
  }
  return op_ls_advect_weno5(ctx, phi, vel, jbn, dx, dy, dz, dt);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host-pointer drop-ins.  A context per grid size (scheme C6C, the scheme of
// the reference's CUDA path) is kept with a staging area, so repeated calls
// allocate nothing; copies are synchronous like the reference's cudaMemcpy.

namespace {

std::mutex g_legacy_mu;

emx_ctx* legacy_ctx(int scheme, int nx, int ny, int nz, size_t stage_doubles) {
  static std::map<std::tuple<int, int, int, int, int>, emx_ctx*> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto key = std::make_tuple(dev, scheme, nx, ny, nz);
  emx_ctx* c = nullptr;
  auto it = cache.find(key);
  if (it != cache.end()) {
    c = it->second;
  } else {
    if (emx_create(&c, nx, ny, nz, scheme, nullptr) != EMX_OK) return nullptr;
    cache[key] = c;
  }
  const size_t need = stage_doubles * sizeof(double);
  if (c->stage_bytes < need) {
    (void)hipFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    if (hipMalloc(&c->stage, need) != hipSuccess) {
      set_error("staging allocation of %zu bytes failed", need);
      return nullptr;
    }
    c->stage_bytes = need;
  }
  return c;
}

[[noreturn]] void legacy_fail(const char* who) {
  fprintf(stderr, "%s error: %s\n", who, emx_last_error());
  exit(-1);
}

#define LEG(call, who)                 \
  do {                                 \
    if ((call) != EMX_OK) legacy_fail(who); \
  } while (0)

int h2d(emx_ctx* c, double* dst, const double* src, size_t n) {
  EMX_TRY_HIP(hipMemcpyAsync(dst, src, n * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  return EMX_OK;
}
int d2h(emx_ctx* c, double* dst, const double* src, size_t n) {
  EMX_TRY_HIP(hipMemcpyAsync(dst, src, n * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  return EMX_OK;
}

}  // namespace

extern "C" {

void cudaDeriv(double* dVel, double* dTemp, double* vel, double* temp,
               double* amut, double* dxyzsgd, double* jbn, double deltaX,
               double deltaY, double deltaZ, int Nx, int Ny, int Nz) {
  (void)amut;
  (void)dxyzsgd;
  (void)jbn;
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  const size_t N = (size_t)Nx * Ny * Nz;
  emx_ctx* c = legacy_ctx(EMX_SCHEME_C6C, Nx, Ny, Nz, 16 * N);
  if (!c) legacy_fail("cudaDeriv");
  double* du = c->stage;       // 3N
  double* dt = du + 3 * N;     // N
  double* ddv = dt + N;        // 9N
  double* ddt = ddv + 9 * N;   // 3N
  LEG(h2d(c, du, vel, 3 * N), "cudaDeriv");
  LEG(h2d(c, dt, temp, N), "cudaDeriv");
  // the reference passes 1/deltax and its kernels multiply by 1.0/delta
  // (src/Schemes/derivCompact.cu:105)
  LEG(emx_grad(c, du, dt, ddv, ddt, 1.0 / deltaX, 1.0 / deltaY, 1.0 / deltaZ),
      "cudaDeriv");
  LEG(d2h(c, dVel, ddv, 9 * N), "cudaDeriv");
  LEG(d2h(c, dTemp, ddt, 3 * N), "cudaDeriv");
  LEG(emx_sync(c), "cudaDeriv");
}

void cuDivergence(double* rs, double* rsv, double* e, double* f, double* g,
                  double* ev, double* fv, double* gv, double* jbn,
                  double deltaX, double deltaY, double deltaZ, int Nx, int Ny,
                  int Nz) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  const size_t N = (size_t)Nx * Ny * Nz;
  emx_ctx* c = legacy_ctx(EMX_SCHEME_C6C, Nx, Ny, Nz, 19 * N);
  if (!c) legacy_fail("cuDivergence");
  double* flux = c->stage;    // 6N
  double* dj = flux + 6 * N;  // 11N (comps 1,5,9,10,11 filled)
  double* drs = dj + 11 * N;  // 2N
  double* src[6] = {e, f, g, ev, fv, gv};
  for (int q = 0; q < 6; q++) LEG(h2d(c, flux + q * N, src[q], N), "cuDivergence");
  const int comps[5] = {0, 4, 8, 9, 10};  // jbn 1,5,9,10,11 (cuDivergence.cu:158-163)
  for (int q = 0; q < 5; q++)
    LEG(h2d(c, dj + comps[q] * N, jbn + comps[q] * N, N), "cuDivergence");
  LEG(emx_divergence(c, flux, dj, 1.0 / deltaX, 1.0 / deltaY, 1.0 / deltaZ,
                     drs, drs + N),
      "cuDivergence");
  LEG(d2h(c, rs, drs, N), "cuDivergence");
  LEG(d2h(c, rsv, drs + N, N), "cuDivergence");
  LEG(emx_sync(c), "cuDivergence");
}

void advect_CUDA(double* press, double* jbn, double* sobject, double deltaX,
                 double deltaY, double deltaZ, int Nx, int Ny, int Nz,
                 double deltamin) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  const size_t N = (size_t)Nx * Ny * Nz;
  // the level set needs no derivative table; any scheme whose minimum line
  // length the grid meets will do for the context
  emx_ctx* c = legacy_ctx(EMX_SCHEME_C6C, Nx, Ny, Nz, 13 * N);
  if (!c) legacy_fail("advect_CUDA");
  double* dp = c->stage;
  double* dj = dp + N;        // 11N (comps 1,5,9 filled)
  double* ds = dj + 11 * N;
  LEG(h2d(c, dp, press, N), "advect_CUDA");
  const int comps[3] = {0, 4, 8};
  for (int q = 0; q < 3; q++)
    LEG(h2d(c, dj + comps[q] * N, jbn + comps[q] * N, N), "advect_CUDA");
  LEG(h2d(c, ds, sobject, N), "advect_CUDA");
  LEG(emx_ls_extrapolate(c, dp, dj, ds, deltaX, deltaY, deltaZ, deltamin),
      "advect_CUDA");
  LEG(d2h(c, press, dp, N), "advect_CUDA");
  LEG(emx_sync(c), "advect_CUDA");
}

int emx_host_flujos(int neq, const double* um, const double* u,
                    const double* pres, const double* vis, const double* jbn,
                    const double* dvel, double* e, double* f, double* g,
                    double* ev, double* fv, double* gv, int Nx, int Ny,
                    int Nz) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  if (!um || !u || !pres || !vis || !jbn || !dvel || !e || !f || !g || !ev ||
      !fv || !gv) {
    set_error("emx_host_flujos: null pointer");
    return EMX_ERR_ARG;
  }
  const size_t N = (size_t)Nx * Ny * Nz;
  emx_ctx* c = legacy_ctx(EMX_SCHEME_C6C, Nx, Ny, Nz, 36 * N);
  if (!c) return EMX_ERR_NOMEM;
  double* dum = c->stage;       // 5N
  double* du = dum + 5 * N;     // 3N
  double* dp = du + 3 * N;      // N
  double* dvs = dp + N;         // N
  double* dj = dvs + N;         // 11N
  double* ddv = dj + 11 * N;    // 9N
  double* dfl = ddv + 9 * N;    // 6N
  EMX_TRY(h2d(c, dum, um, 5 * N));
  EMX_TRY(h2d(c, du, u, 3 * N));
  EMX_TRY(h2d(c, dp, pres, N));
  EMX_TRY(h2d(c, dvs, vis, N));
  EMX_TRY(h2d(c, dj, jbn, 11 * N));
  EMX_TRY(h2d(c, ddv, dvel, 9 * N));
  EMX_TRY(emx_flux(c, neq, dum, du, dp, dvs, dj, ddv, dfl));
  double* dst[6] = {e, f, g, ev, fv, gv};
  for (int q = 0; q < 6; q++) EMX_TRY(d2h(c, dst[q], dfl + q * N, N));
  return emx_sync(c);
}

int emx_host_sgdm(const double* u, const double* dvel, const double* jbn,
                  const double* sgd, double c4, double* amut, double* vis,
                  double* vis5, int Nx, int Ny, int Nz) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  if (!u || !dvel || !jbn || !sgd || !amut || !vis) {
    set_error("emx_host_sgdm: null pointer");
    return EMX_ERR_ARG;
  }
  const size_t N = (size_t)Nx * Ny * Nz;
  emx_ctx* c = legacy_ctx(EMX_SCHEME_C6C, Nx, Ny, Nz, 33 * N);
  if (!c) return EMX_ERR_NOMEM;
  double* du = c->stage;      // 3N
  double* ddv = du + 3 * N;   // 9N
  double* dj = ddv + 9 * N;   // 11N
  double* dsg = dj + 11 * N;  // 7N
  double* dout = dsg + 7 * N; // 3N
  EMX_TRY(h2d(c, du, u, 3 * N));
  EMX_TRY(h2d(c, ddv, dvel, 9 * N));
  EMX_TRY(h2d(c, dj, jbn, 11 * N));
  EMX_TRY(h2d(c, dsg, sgd, 7 * N));
  EMX_TRY(emx_sgdm(c, du, ddv, dj, dsg, c4, dout, dout + N,
                   vis5 ? dout + 2 * N : nullptr));
  EMX_TRY(d2h(c, amut, dout, N));
  EMX_TRY(d2h(c, vis, dout + N, N));
  if (vis5) EMX_TRY(d2h(c, vis5, dout + 2 * N, N));
  return emx_sync(c);
}

}  // extern "C"
