// Host-side operator launchers shared by the C ABI (capi.hip).
#pragma once
#include "emx_internal.hpp"

namespace emx {

struct FluxPtrs {
  const double* um;
  const double* u;
  const double* pres;
  const double* vis;
  const double* jbn;
  const double* dvel;
};

const char* last_error();

int op_deriv(emx_ctx* ctx, int dir, const double* f, double* df, double ih);
int op_grad(emx_ctx* ctx, const double* u, const double* temp, double* dvel,
            double* dtemp, const double ih[3]);
int op_grad_div(emx_ctx* ctx, const double* u, const double* jbn, double* dvel, double* div,
                const double ih[3]);
int op_flux(emx_ctx* ctx, int neq, const FluxPtrs& fp, double* flux);
int op_divergence(emx_ctx* ctx, const double* flux, const double* jbn,
                  const double ih[3], double* rs, double* rsv);
int op_flux_div(emx_ctx* ctx, int neq, const FluxPtrs& fp, const double ih[3],
                double* rs, double* rsv);
int op_flux_div_all(emx_ctx* ctx, const FluxPtrs& fp, const double ih[3],
                    double* rs, double* rsv);
int op_metrics(emx_ctx* ctx, const double* x, double* jbn, int* bad_dev);
int op_rk_update(emx_ctx* ctx, int sub, double* um1, const double* um0, const double* um,
                 const double* rs, const double* rsv, double dt);
int op_frontera_val(emx_ctx* ctx, double* um1);
int op_ls_reinit(emx_ctx* ctx, double* phi, const double* jbn, const double* dmin,
                 double dx, double dy, double dz, double deltamin);
int op_varc_to_varnc(emx_ctx* ctx, const double* um1, double mach, double* u, double* pres,
                     double* temp, double* conc);
int op_varnc_to_varc(emx_ctx* ctx, const double* u, const double* pres, const double* temp,
                     double mach, double* um0, double* pres0);
int op_rk_clip(emx_ctx* ctx, double* um1, const double* mask, double mach, double* u,
               double* pres, double* temp, double* conc);
int op_rk_rebuild(emx_ctx* ctx, double* um1, const double* u, const double* pres, double mach);
int op_filtrado(emx_ctx* ctx, int itf, double* um1, const double* jbn, double cfilt,
                double mach, double* u, double* pres, double* temp);
int op_sgs_geometry(emx_ctx* ctx, const double* x, double* sgd,
                    double* dmins, double* blockmax, int nblocks);
int op_sgdm(emx_ctx* ctx, const double* u, const double* dvel,
            const double* jbn, const double* sgd, double c4, double* amut,
            double* vis, double* vis5);

int op_dynsmag(emx_ctx* ctx, const double* u, const double* dvel, const double* jbn,
               const double* delta, double* nut, double c4, double* vis);
// The dynamic model's two kernels on a stream, with LM, MM (N doubles each)
// as scratch; ctx (may be NULL) only for the per-launch timing.
int launch_dynsmag(emx_ctx* ctx, hipStream_t stream, const Dims& d, const double* u,
                   const double* dvel, const double* jbn, const double* delta, double* nut,
                   double c4, double* vis, double* LM, double* MM);

// learned closure (closure.hip); packed operands in ctx->mlp_pack
size_t mlp_pack_elems();
int op_mlp_nut(emx_ctx* ctx, const float* params, const float* norm, const double* dvel,
               const double* jbn, const double* delta, double* nut, double c4, double* vis);
// packs the network's BF16 operands into ctx->mlp_pack
int op_mlp_pack(emx_ctx* ctx, const float* params, const float* norm);


int op_mlp_gather(emx_ctx* ctx, const double* dvel, const double* jbn, const double* delta,
                  const long long* idx, int n, float* feat);

// adjoints (ops_adj.hip)
int op_deriv_adj(emx_ctx* ctx, int dir, const double* g, double* out, double ih);
int op_grad_adj(emx_ctx* ctx, const double* dvel_bar, const double* dtemp_bar,
                double* u_bar, double* temp_bar, const double ih[3]);
int op_divergence_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                      const double* jbn, const double ih[3], double* flux_bar);
int op_flux_div_all_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                        const FluxPtrs& fp, const double ih[3], double* um_bar,
                        double* u_bar, double* pres_bar, double* vis_bar,
                        double* dvel_bar);

// One grid of the batched level-set step: dimensions, the caller's phi and
// velocity (3 comps), and the batch's padded buffers: P (phi), V (velocity,
// 3), J (jbn 1/5/9, padded once when bound), T1, T2 (RK stages).
struct LsBatchGrid {
  int nx, ny, nz;
  double* phi;
  const double* vel;
  const double* jbn;  // the box's metrics (11 components, interior grid)
  double* P;
  double* T1;
  double* T2;
  double iX, iY, iZ, dt;
};
long long ls_padded_points(int nx, int ny, int nz);
// blk_pad: per-grid first workgroup (ngrid + 1 entries) over the padded
// points, 256 points per workgroup.
int op_batch_ls_step(hipStream_t st, const LsBatchGrid* grids, const int* blk_pad,
                     int nblk_pad, int ngrid);

// level set (levelset.hip); scratch is ctx->ls_buf
size_t ls_scratch_doubles(const Dims& d);
int op_ls_extrapolate(emx_ctx* ctx, double* pres, const double* jbn,
                      const double* sobject, double dx, double dy, double dz,
                      double deltamin);
int op_ls_advect_weno5(emx_ctx* ctx, double* phi, const double* vel,
                       const double* jbn, double dx, double dy, double dz,
                       double dt);
int op_ls_heviside(emx_ctx* ctx, const double* phi, double deltamax, double* hevi);

}  // namespace emx
