/*
 * emx_oracle.h -- CPU oracle for the eMix3d per-RK-stage spatial operator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C restatement of the
 * reference algorithms, used as the checker by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product path (emix3d_amd, libemix3d_hip.so) never links or calls it.
 *
 * Conventions (identical to the reference, solver/mix3d.mk:3 and
 * src/Engine/cuArray.cu:119-124):
 *   - FP64 everywhere;
 *   - arrays are Fortran column-major (nx,ny,nz[,c]); element (i,j,k,c),
 *     0-based, sits at i + nx*(j + ny*(k + nz*c));
 *   - inv_h is the inverse grid spacing in computational space, the
 *     Fortran `delta` = float(N-1) of src/MSVSingle/atl_incomp.f90:30-32.
 *
 * Parity pins (what checks each function; details in tests/):
 *   orc_deriv_line C4/C6F, orc_jacobeano, orc_deriv_vel, orc_flujos,
 *   orc_divergencia, orc_inisgdm, orc_sgdm:
 *       golden vectors from the reference Fortran built by flang
 *       (oracle/_ref), plus closed forms (polynomial exactness, dense solve).
 *   orc_deriv_line C6C: CUDA-only in the reference; pinned by polynomial
 *       exactness (degree <= 4 on every row) and a dense solve of the
 *       matrices written in src/Schemes/derivCompact.cu:83-215.
 *   orc_ls_extrapolate, orc_ls_advect_weno5: CUDA-only in the reference;
 *       pinned by invariants and closed forms (constant preservation,
 *       untouched fluid points, WENO5 exactness on cubics and 5th-order
 *       convergence, 1-D upwind reduction).
 *   orc_dynsmag: no reference implementation exists (the reference's
 *       model is the structure-function sgdm); pinned by closed forms
 *       (zero for uniform flow and rigid rotation; on any linear field
 *       u = A x of a grid with spacings h_a, nu_t = -(A H^2 A^T : S) /
 *       (12 S:S) away from the walls, H = diag(h_a), independent of the
 *       filter width) that fail for a wrong filter ratio, filter weight or
 *       off-diagonal weight, and by the model's definition written with full
 *       3x3 tensors and a 27-point filter on small random fields.
 *   orc_*_adj: the reference has no adjoint.  Each is the transpose of the
 *       oracle function it names, pinned by dot-product tests against that
 *       function (itself pinned as above) and by the transposed dense
 *       operator matrix on short lines.
 */
#ifndef EMX_ORACLE_H
#define EMX_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_C4 = 0, ORC_C6F = 1, ORC_C6C = 2 };

/* Tridiagonal coefficients (a: sub, b: diag, c: super) of a scheme, n rows. */
int orc_coeffs(int scheme, int n, double *a, double *b, double *c);

/* Compact first derivative of one line of n points read with stride `is`
 * from f, written with stride `os` to out (out may alias f).  Returns 0, or
 * -1 when n is below the scheme's minimum. */
int orc_deriv_line(int scheme, int n, const double *f, long is, double inv_h,
                   double *out, long os);

/* Compact derivative of a whole field along dir (0=xi, 1=eta, 2=zeta). */
int orc_deriv3d(int scheme, int dir, int nx, int ny, int nz, const double *f,
                double inv_h, double *df);

/* Grid metrics, solver/jacob_malla.f90:1-171 (x: (nx,ny,nz,3); jbn: 11). */
int orc_jacobeano(int scheme, int nx, int ny, int nz, const double *x,
                  double *jbn);

/* Velocity/temperature gradients, solver/deriv_vel.f90:1-161
 * (u: 3 comps, temp: 1, dvel: 9, dtemp: 3; temp/dtemp may be NULL). */
int orc_deriv_vel(int scheme, int nx, int ny, int nz, const double *u,
                  const double *temp, double *dvel, double *dtemp,
                  double ihx, double ihy, double ihz);

/* Pointwise fluxes, solver/flujos_compact.f90:1-220.
 * flux: 6 comps e,f,g,ev,fv,gv. */
int orc_flujos(int neq, int nx, int ny, int nz, const double *um,
               const double *u, const double *pres, const double *vis,
               const double *jbn, const double *dvel, double *flux);

/* Flux divergence, solver/divergencia.f90:1-154 (flux: 6 comps, not
 * modified). */
int orc_divergencia(int scheme, int nx, int ny, int nz, const double *flux,
                    const double *jbn, double ihx, double ihy, double ihz,
                    double *rs, double *rsv);

/* SGS geometry, solver/sgdm_incomp.f90:2-114.  sgd: 7 comps
 * (dxyzsgd, dxp, dxm, dyp, dym, dzp, dzm); dmins may be NULL. */
int orc_inisgdm(int nx, int ny, int nz, const double *x, double *sgd,
                double *dmins, double *deltamax);

/* Structure-function eddy viscosity + viscosidad,
 * solver/sgdm_incomp.f90:117-395 and solver/viscosidad_incomp.f90:1-23.
 * vis5 may be NULL.  work: 3*N doubles of scratch (vorticity). */
int orc_sgdm(int nx, int ny, int nz, const double *u, const double *dvel,
             const double *jbn, const double *sgd, double c4, double *amut,
             double *vis, double *vis5);

/* Pressure extrapolation into the solid, the advect_CUDA path of
 * src/LevelSet/lsTransportEc.cu:24-197 -> src/LevelSet/interface.cu:341-436.
 * dx,dy,dz are the grid spacings (the reference passes 1/deltax);
 * pres is updated in place. */
int orc_ls_extrapolate(int nx, int ny, int nz, double *pres, const double *jbn,
                       const double *sobject, double dx, double dy, double dz,
                       double deltamin);

/* One SSP-RK3 step of WENO5 level-set advection on ghost-padded arrays,
 * src/LevelSet/lsTransportEc.cu:201-418 (cuAdvectLsJB).  phi, vel (3 comps)
 * and jbn (11 comps, only 1/5/9 read) have padded dims (ngx,ngy,ngz). */
int orc_ls_advect_weno5_padded(int ngx, int ngy, int ngz, double *phi,
                               const double *vel, const double *jbn,
                               double dx, double dy, double dz, double dt);

/* The left- and right-biased WENO5 derivatives (3 comps each) that one
 * stage of orc_ls_advect_weno5_padded forms from phi. */
int orc_weno5_derivs(int ngx, int ngy, int ngz, const double *phi,
                     const double *jbn, double dx, double dy, double dz,
                     double *left, double *right);

/* Same step on unpadded arrays: mirror-pad by 5 cells the way advect_CUDA
 * pads its inputs (lsTransportEc.cu:100-166), step, copy the interior back. */
int orc_ls_advect_weno5(int nx, int ny, int nz, double *phi, const double *vel,
                        const double *jbn, double dx, double dy, double dz,
                        double dt);

/* Even mirror padding by g cells, as cuSwapToGhost followed by
 * cuGhostCellsMirror3d{X,Y,Z} (src/LevelSet/lsTools.cu:482-585,772-796). */
void orc_pad_mirror(int nx, int ny, int nz, int g, const double *in,
                    double *out);

/* Dynamic Smagorinsky (Germano-Lilly) eddy viscosity with a box test filter
 * of ratio 2.  Not in the reference (see header); inputs are the velocity
 * and its physical-space gradient g[3*a+b] = du_a/dx_b (9 comps), and the
 * grid filter width delta (one value per point).  Clipped at nu_t >= 0. */
int orc_dynsmag(int nx, int ny, int nz, const double *u, const double *gphys,
                const double *delta, double *nut);
/* The same model with its constants as arguments: ratio2 = (test filter /
 * grid filter width)^2 (4 for ratio 2), ws = side weight of the separable
 * test filter (ws, 1 - 2 ws, ws) (1/4), offw = weight of an off-diagonal
 * pair in the tensor contractions (2).  orc_dynsmag is (4, 1/4, 2); tests
 * pass other values to show that the closed-form checks fail on them. */
int orc_dynsmag_cfg(int nx, int ny, int nz, const double *u, const double *gphys,
                    const double *delta, double *nut, double ratio2, double ws,
                    double offw);

/* Level-set reinitialisation cuReinitLsJB (src/LevelSet/lsTransportEc.cu:
 * 422-676): ten iterations of ghost mirroring (cuGhostCellsMirror3d{Z,Y,X},
 * lsTools.cu:482-585), phi0 = phi and three SSP-RK3 stages of
 * phi_t = S(phi0)(1 - |grad phi|) (reini_RS_WENOJB, lsTools.cu:824-900, with
 * the WENO5 derivatives of the advection), dt = deltamin/2, on the 5-cell
 * mirror-padded grid; dmin is the smoothing width of S. */
int orc_ls_ghost_mirror(int Nx, int Ny, int Nz, int g, double *a);
int orc_ls_reinit(int nx, int ny, int nz, double *phi, const double *jbn, const double *dmin,
                  double dx, double dy, double dz, double deltamin);

/* heviside (solver/advectLS.f90:29-59) on n values: eps = band * deltamax
 * (band 1.5 in the reference; a parameter so tests can plant a mistake),
 * hevi = 0 below -eps, 1 above eps, 0.5 + phi/(2 eps) + sin(pi phi/eps)/(2 pi)
 * on the band; NaN phi leaves hevi as it was. */
int orc_heviside(long n, const double *phi, double deltamax, double band, double *hevi);

/* SSP-RK3 update of sub-step sub (1..3), solver/rungek_ico.f90:87-91,
 * 152-153, 216-219, on n values (um may be NULL for sub-step 1). */
int orc_rk_update(int sub, long n, double *um1, const double *um0, const double *um,
                  const double *rs, const double *rsv, double dt);

/* Boundary values of um1 (nx,ny,nz,5) in the RUNGEK order z, y, x:
 * frontera_z_val (src/MSVSingle/frontera_z.f90:219-262), frontera_y_val
 * (frontera_y.f90:220-247), frontera_x_val (frontera_x.f90:223-256). */
int orc_frontera_val(int nx, int ny, int nz, double *um1);
/* varc_to_varnc (src/MSVSingle/herramientas_incomp2.f90:2-30): conc and
 * temp = um1(5), pres = um1(1)/(Mach*Mach), u(c) = um1(1+c)/um1(1); conc may
 * be NULL. */
int orc_varc_to_varnc(long n, const double *um1, double mach, double *u, double *pres,
                      double *temp, double *conc);
/* varnc_to_varc (herramientas_incomp2.f90:32-58); pres0 may be NULL. */
int orc_varnc_to_varc(long n, const double *u, const double *pres, const double *temp,
                      double mach, double *um0, double *pres0);
/* The RUNGEK velocity clip (solver/rungek_ico.f90:102-123): varc_to_varnc,
 * u clipped to [-6, 6], um1(2..4) = u*um1(1)*mask, varc_to_varnc again. */
int orc_rk_clip(long n, double *um1, const double *mask, double mach, double *u, double *pres,
                double *temp, double *conc);
/* The rebuild after the level-set step (rungek_ico.f90:126-135):
 * um1(1) = pres*Mach*Mach, um1(1+c) = u(c)*um1(1). */
int orc_rk_rebuild(long n, double *um1, const double *u, const double *pres, double mach);

/* Compact filter.  orc_filter_coeffs: the tridiagonal a, b, c of one line
 * (INIFILTRO, src/MSVSingle/filtro_incomp.f90:2-139): periodic = 0 gives the
 * non-periodic arrays (a = c = cfilt on 1-based rows 6..n-5, zero on the
 * five rows at each end), periodic = 1 the periodic ones (a = c = cfilt on
 * every row).  orc_filtronper: one line in place (filtro_incomp.f90:307-362,
 * RHS constants :140-145).  orc_filtrado: FILTRADO(itf)
 * (herramientas_incomp2.f90:227-410) on um1 (nx,ny,nz,5) in place -- each
 * field times J, filtered along the three directions in the order itf
 * selects (xi and eta with the non-periodic arrays, zeta with the periodic
 * ones), times 1/J -- then u, temp, pres recomputed. */
int orc_filter_coeffs(int periodic, int n, double cfilt, double *a, double *b, double *c);
int orc_filtronper(int n, double *u, const double *a, const double *b, const double *c,
                   double cfilt);
int orc_filtrado(int itf, int nx, int ny, int nz, double *um1, const double *jbn, double cfilt,
                 double mach, double *u, double *pres, double *temp);

/* ---- learned closure network (no reference counterpart) ------------------ */
/* 10 -> 64 -> 64 -> 64 -> 1, SiLU hidden activations, softplus output.
 * params (9089): W1[64][10] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
 * W4[64] b4; norm (21): in_shift[10] in_scale[10] out_scale.
 * x_k = (f_k - in_shift_k) * in_scale_k; nu = out_scale * softplus(z4).
 * Evaluated in FP64 from FP32 parameters.  Pinned by closed forms and
 * central differences (tests/). */
int orc_mlp_forward(const float *params, const float *norm, const double *feat,
                    long n, double *nu);
/* grad (9089, FP64) = sum_p dnu_p * d nu_p / d params. */
int orc_mlp_grad(const float *params, const float *norm, const double *feat,
                 const double *dnu, long n, double *grad);
/* The 10 features of each point: du_a/dx_b = jbn(1+4b) * dvel(3b+a)
 * (index 3a+b), then delta; feat is [N][10]. */
int orc_mlp_features(int nx, int ny, int nz, const double *dvel, const double *jbn,
                     const double *delta, double *feat);

/* ---- adjoints ----------------------------------------------------------- */

/* out = D^T g along dir, D = orc_deriv3d(scheme, dir, ., inv_h). */
int orc_deriv3d_adj(int scheme, int dir, int nx, int ny, int nz, const double *g,
                    double inv_h, double *out);

/* Transpose of orc_deriv_vel: u_bar(c) = sum_d D_d^T dvel_bar(3d+c);
 * temp_bar = sum_d D_d^T dtemp_bar(d) (both may be NULL). */
int orc_deriv_vel_adj(int scheme, int nx, int ny, int nz, const double *dvel_bar,
                      const double *dtemp_bar, const double ih[3], double *u_bar,
                      double *temp_bar);

/* Transpose of orc_divergencia with respect to its 6 flux fields. */
int orc_divergencia_adj(int scheme, int nx, int ny, int nz, const double *rs_bar,
                        const double *rsv_bar, const double *jbn, double ihx,
                        double ihy, double ihz, double *flux_bar);

/* Transpose of the linearisation of orc_flujos(neq) about (um, u, vis,
 * dvel); ADDS to the adjoint fields. */
int orc_flujos_adj(int neq, int nx, int ny, int nz, const double *um,
                   const double *u, const double *vis, const double *jbn,
                   const double *dvel, const double *flux_bar, double *um_bar,
                   double *u_bar, double *pres_bar, double *vis_bar,
                   double *dvel_bar);

/* Transpose of the linearisation of the five-equation flux divergence
 * (orc_flujos + orc_divergencia, m = 1..5); rs_bar, rsv_bar: 5 fields each.
 * Overwrites the adjoint fields. */
int orc_flux_div_all_adj(int scheme, int nx, int ny, int nz, const double *rs_bar,
                         const double *rsv_bar, const double *um, const double *u,
                         const double *vis, const double *jbn, const double *dvel,
                         double ihx, double ihy, double ihz, double *um_bar,
                         double *u_bar, double *pres_bar, double *vis_bar,
                         double *dvel_bar);

#ifdef __cplusplus
}
#endif
#endif
