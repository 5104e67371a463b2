/*
 * emix3d.h -- C ABI of libemix3d_hip.so, the MI355X (gfx950) implementation
 * of eMix3d's per-RK-stage spatial operator.
 *
 * Data: FP64, Fortran column-major.  Element (i,j,k,c) of an (nx,ny,nz,nc)
 * array is at i + nx*(j + ny*(k + nz*c)), 0-based; `N` below is nx*ny*nz.
 * Component numbering follows the reference:
 *   jbn  (11 comps): 1..9 inverse metrics (1=xi_x, 5=eta_y, 9=zeta_z),
 *                    10 = J, 11 = 1/J          (solver/jacob_malla.f90)
 *   dvel (9 comps):  u_xi, v_xi, w_xi, u_eta, v_eta, w_eta, u_zeta, v_zeta,
 *                    w_zeta                    (solver/deriv_vel.f90:17-123)
 *   um   (5 comps):  conservative variables of RUNGEK
 *   flux (6 comps):  e, f, g, ev, fv, gv       (solver/flujos_compact.f90)
 *   sgd  (7 comps):  dxyzsgd, dxpsgd, dxmsgd, dypsgd, dymsgd, dzpsgd, dzmsgd
 *                                              (solver/sgdm_incomp.f90:2-114)
 * `ih*` arguments are inverse computational spacings (the Fortran
 * `deltax` = float(nx-1), src/MSVSingle/atl_incomp.f90:30-32).
 *
 * Two families of entry points:
 *   - emx_* take an emx_ctx and DEVICE pointers; they are asynchronous on
 *     the context's stream, allocate nothing, and can be captured in a HIP
 *     graph (tests/test_gpu_graph.py replays the stage and both closures),
 *     except the set-up calls that report to the host (emx_metrics and
 *     emx_sgs_geometry wait for their result) and the first emx_filtrado
 *     per (line length, cfilt), which builds its line tables.  They return
 *     EMX_OK or a negative status; emx_last_error() describes the last
 *     failure of the calling thread.
 *   - the reference's own bind(C) symbols (cudaDeriv, cuDivergence,
 *     advect_CUDA) with their exact signatures, taking HOST pointers and
 *     returning after the results are back in host memory, plus host
 *     variants of the two routines the reference runs in Fortran.
 */
#ifndef EMIX3D_H
#define EMIX3D_H
#ifdef __cplusplus
extern "C" {
#endif

#define EMX_VERSION 1

/* compact schemes:
 *  C4 : 4th-order Pade alpha=1/4, 4th-order alpha=3 closure
 *       (solver/derivada.f90:2-125)
 *  C6F: 6th-order alpha=1/3, 4th-order rows 2/n-1, 3rd-order alpha=2
 *       closure (src/MSVSingle/derivada_incomp.f90:2-90,169-201)
 *  C6C: 6th-order alpha=1/3, 4th-order rows 1/N-2, 4th-order alpha=3
 *       closure (src/Schemes/derivCompact.cu:83-215) */
#define EMX_SCHEME_C4 0
#define EMX_SCHEME_C6F 1
#define EMX_SCHEME_C6C 2

/* status codes */
#define EMX_OK 0
#define EMX_ERR_ARG (-1)    /* bad argument (null pointer, bad enum)      */
#define EMX_ERR_HIP (-2)    /* HIP runtime error                          */
#define EMX_ERR_SIZE (-3)   /* grid too small / too large for the scheme  */
#define EMX_ERR_METRIC (-4) /* degenerate metric (|J| <= 1e-20)           */
#define EMX_ERR_NOMEM (-5)  /* device allocation failed                   */
#define EMX_ERR_IO (-6)     /* file missing, unreadable or malformed      */

typedef struct emx_ctx emx_ctx;

/* ---- library and context ------------------------------------------------ */

/* Text of the last failure on the calling thread ("" if none). */
const char* emx_last_error(void);
int emx_version(void);

/* Create a context for an nx*ny*nz grid and a scheme, on the current HIP
 * device, launching on `stream` (a hipStream_t; NULL = default stream).
 * Allocates the line-solver tables and 6*N doubles of scratch once. */
int emx_create(emx_ctx** ctx, int nx, int ny, int nz, int scheme,
               void* stream);
int emx_destroy(emx_ctx* ctx);
int emx_set_stream(emx_ctx* ctx, void* stream);
/* Line-solver launch shape of operation class op (0 = derivatives, gradient
 * and metrics; 1 = flux divergence) along direction dir (0=xi,1=eta,2=zeta):
 * kind 0 = one wavefront per line (lanes along the line); kinds 1, 2, 3, 4 =
 * 64, 32, 16 or 8 neighbouring lines per wavefront with the line split into
 * chunks over the lanes and waves of a workgroup; C = points per chunk.
 * Valid for op 0: kind 0 C in {1,2,4,8,16}; kind 1 {8,16,32}; kind 2
 * {4,8,16}; kind 3 {4,8}; kind 4 {4,8}.  Op 1: kind 0 C <= 16, kinds 2 and
 * 3 C in {4,8}, kind 4 C in {2,4}, and kind 5 with C = 0: the fluxes of
 * the direction projected into scratch (one pointwise kernel for all
 * directions on kind 5, scratch of 8 fields per such direction, grown by
 * this call; EMX_ERR_NOMEM and the plan unchanged if that fails), then one
 * single-field solve each on the derivative plan.  Directions on kind 5 run
 * first, so the sum of the three directional terms follows the plan's pass
 * order.  Used by the autotuner. */
int emx_set_plan(emx_ctx* ctx, int op, int dir, int kind, int C);
int emx_get_plan(const emx_ctx* ctx, int op, int dir, int* kind, int* C);

/* Per-launch timing, for diagnostics and the launch-shape autotuner (no
 * counterpart in the reference, which has no timers).  Between
 * emx_timing_begin and emx_timing_end every kernel the context launches is
 * bracketed by two HIP events on the context stream (events are created on
 * first use and reused).  emx_timing_end waits for the last launch and
 * aggregates per (kernel label, direction), sorted by total time; entry i
 * gives the label, direction (-1 for pointwise kernels), launch count,
 * summed duration and summed algorithmic bytes (each field the kernel must
 * read or write once, 8 B per grid point).  emx_timing_select restricts the
 * recording to the launches of one kernel label (NULL: every launch), so a
 * timed loop carries events around that kernel only. */
int emx_timing_begin(emx_ctx* ctx);
int emx_timing_select(emx_ctx* ctx, const char* label);
int emx_timing_end(emx_ctx* ctx, int* n_entries);
int emx_timing_entry(const emx_ctx* ctx, int i, const char** label, int* dir,
                     int* launches, double* total_ms, double* bytes);
int emx_sync(emx_ctx* ctx);

/* ---- device-pointer operator entry points ------------------------------- */

/* One compact first derivative along dir: one call of derivnper per line
 * (solver/derivada.f90:95-125) / cu::cs::dF{X,Y,Z}
 * (src/Schemes/derivCompact.cu:222-409).  df must not alias f. */
int emx_deriv(emx_ctx* ctx, int dir, const double* f, double* df,
              double inv_h);

/* Velocity (3 comps) and temperature gradients in computational space:
 * DERIV_VEL (solver/deriv_vel.f90:1-203) / cudaDeriv
 * (solver/cuDeriv.cu:529-654).  temp/dtemp may both be NULL. */
int emx_grad(emx_ctx* ctx, const double* u, const double* temp, double* dvel,
             double* dtemp, double ihx, double ihy, double ihz);

/* Velocity gradient (as emx_grad, no temperature) fused with the velocity
 * divergence in divergencia's conservative form with the velocity as the
 * flux, div = (1/J)(D_xi(J xi_x u) + D_eta(J eta_y v) + D_zeta(J zeta_z w))
 * (solver/divergencia.f90:225-354).  One line pass per direction. */
int emx_grad_div(emx_ctx* ctx, const double* u, const double* jbn, double* dvel, double* div,
                 double ihx, double ihy, double ihz);

/* Grid metrics from coordinates x (3 comps): JACOBEANO
 * (solver/jacob_malla.f90:1-171), with the context's scheme and
 * ih = n-1 per direction.  Returns EMX_ERR_METRIC if any |J| <= 1e-20. */
int emx_metrics(emx_ctx* ctx, const double* x, double* jbn);

/* SGS geometry: INISGDM (solver/sgdm_incomp.f90:2-114).  dmins may be
 * NULL; *deltamax (host) receives max spacing when non-NULL (synchronises). */
int emx_sgs_geometry(emx_ctx* ctx, const double* x, double* sgd,
                     double* dmins, double* deltamax);

/* Structure-function eddy viscosity and total viscosity: viscosidad_t
 * (solver/sgdm_incomp.f90:117-395) + viscosidad
 * (solver/viscosidad_incomp.f90:1-23).  vis5 may be NULL. */
int emx_sgdm(emx_ctx* ctx, const double* u, const double* dvel,
             const double* jbn, const double* sgd, double c4, double* amut,
             double* vis, double* vis5);

/* Pointwise fluxes of equation neq (1..5): flujos(neq)
 * (solver/flujos_compact.f90:1-220).  flux: 6 comps. */
int emx_flux(emx_ctx* ctx, int neq, const double* um, const double* u,
             const double* pres, const double* vis, const double* jbn,
             const double* dvel, double* flux);

/* Flux divergence: divergencia(neq) (solver/divergencia.f90:1-154) /
 * cuDivergence (solver/cuDivergence.cu:107-189).  flux is not modified. */
int emx_divergence(emx_ctx* ctx, const double* flux, const double* jbn,
                   double ihx, double ihy, double ihz, double* rs,
                   double* rsv);

/* flujos(neq) followed by divergencia(neq) without materialising the six
 * flux fields (the pair RUNGEK calls per equation, solver/rungek_ico.f90:
 * 83-95). */
int emx_flux_div(emx_ctx* ctx, int neq, const double* um, const double* u,
                 const double* pres, const double* vis, const double* jbn,
                 const double* dvel, double ihx, double ihy, double ihz,
                 double* rs, double* rsv);

/* The whole equation loop of a RUNGEK sub-step, `DO m=1,nd: flujos(m);
 * divergencia(m)` (solver/rungek_ico.f90:83-95), in three directional
 * passes: rs and rsv hold 5 components (one per equation).  rsv of the
 * continuity and scalar equations is zero, as in the reference. */
int emx_flux_div_all(emx_ctx* ctx, const double* um, const double* u,
                     const double* pres, const double* vis, const double* jbn,
                     const double* dvel, double ihx, double ihy, double ihz,
                     double* rs, double* rsv);

/* Dynamic Smagorinsky eddy viscosity (Germano identity, Lilly least
 * squares, box test filter of ratio 2 with weights (1/4,1/2,1/4) per
 * direction, clipped at 0).  No counterpart in the reference, whose model
 * is the structure-function sgdm; it is the target of the learned closure.
 * u: 3 comps; dvel: 9 comps (emx_grad layout); the physical gradient is
 * du_a/dx_b = jbn(1+4b) * dvel(3b+a) (diagonal metrics, as the reference's
 * hot path assumes); delta: filter width per point.  When vis is not NULL
 * it also receives vis = c4 + nu_t (viscosidad). */
int emx_dynsmag(emx_ctx* ctx, const double* u, const double* dvel, const double* jbn,
                const double* delta, double* nut, double c4, double* vis);

/* emx_dynsmag without a context and without any device allocation: no
 * counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its two scratch fields live in a static device array of
 * 2 * 2^21 doubles (32 MiB per device, held until the process ends), so
 * grids of at most 2^21 points (128^3): larger ones give EMX_ERR_SIZE
 * before any launch.  Same kernels and results as emx_dynsmag, launched on
 * stream (a hipStream_t; NULL = default stream).  Every call on one device
 * shares that device's array: calls that overlap on a device race. */
int emx_dynsmag_noalloc(int nx, int ny, int nz, const double* u, const double* dvel,
                        const double* jbn, const double* delta, double* nut, double c4,
                        double* vis, void* stream);

/* ---- batched multi-grid stage ----------------------------------------------
 * No counterpart in the reference.  Many independent grids ("boxes") of
 * uneven sizes advanced together: per stage the gradient fused with the
 * velocity divergence (emx_grad_div) and one WENO5 / SSP-RK3 level-set step
 * (emx_ls_advect_weno5) for every box, each phase one launch over all
 * boxes, the whole stage captured once in a HIP graph and replayed.  Each
 * box side 5..512 points.  A batch runs on its own stream and orders itself
 * after / before the caller's stream with events. */
typedef struct emx_batch emx_batch;

int emx_batch_create(emx_batch** out, int nbox, const int* dims, int scheme,
                     void* stream);
int emx_batch_destroy(emx_batch* b);
int emx_batch_set_stream(emx_batch* b, void* stream);
/* Device arrays of box i (column-major, as the single-grid entries): u (3
 * comps, in), jbn (11, in; read at every stage like u), phi (in/out),
 * dvel (9, out), div (out); inverse spacings of the derivatives; spacings
 * and time step of the level-set step. */
int emx_batch_bind(emx_batch* b, int i, const double* u, const double* jbn, double* phi,
                   double* dvel, double* div, double ihx, double ihy, double ihz,
                   double dx, double dy, double dz, double dt);
/* Line-solver shape of direction dir: lg lines per wavefront (8, 16, 32,
 * 64) and C points per work-item; built pairs (8,4) (8,8) (16,4) (16,8)
 * (32,4) (32,8) (64,8) (64,16).  lg = 0 (direction 0 only) is one wavefront per line,
 * each box with the shortest of 1, 2, 4, 8 points per lane up to C that
 * fits its lines (C = 1, 2, 4, 8; every line at most 64*C points): the
 * default for direction 0. */
int emx_batch_set_plan(emx_batch* b, int dir, int lg, int C);
int emx_batch_get_plan(const emx_batch* b, int dir, int* lg, int* C);
/* Autotuner: times every built shape of each direction's line pass on the
 * bound data (reps launches each) and keeps the fastest.  It runs after the
 * work queued on the batch's stream argument and overwrites the bound dvel
 * and div (phi is untouched); later work on that stream sees its writes. */
int emx_batch_tune(emx_batch* b, int reps);
/* One stage for all boxes; use_graph != 0 captures the launch sequence on
 * the first call (after a bind or plan change) and replays it. */
int emx_batch_stage(emx_batch* b, int use_graph);

/* ---- learned closure -------------------------------------------------------
 * No counterpart in the reference.  A 10 -> 64 -> 64 -> 64 -> 1 network
 * (SiLU hidden, softplus output) maps the physical velocity gradient
 * du_a/dx_b = jbn(1+4b) * dvel(3b+a) (feature 3a+b) and the filter width
 * delta (feature 9) to nu_t >= 0.  params: FP32[EMX_MLP_NPARAM] =
 * W1[64][10] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4;
 * norm: FP32[EMX_MLP_NNORM] = in_shift[10] in_scale[10] out_scale, with
 * x = (f - in_shift) * in_scale and nu_t = out_scale * softplus(z). */
#define EMX_MLP_NPARAM 9089
#define EMX_MLP_NNORM 21

/* nu_t at every grid point; hidden layers on the matrix cores (BF16
 * operands, FP32 accumulation), FP64 output; vis = c4 + nu_t when vis is
 * not NULL. */
int emx_mlp_nut(emx_ctx* ctx, const float* params, const float* norm, const double* dvel,
                const double* jbn, const double* delta, double* nut, double c4, double* vis);

/* Training.  Features of the n points idx (linear indices), FP32 [n][10]. */
int emx_mlp_gather(emx_ctx* ctx, const double* dvel, const double* jbn, const double* delta,
                   const long long* idx, int n, float* feat);
/* FP32 forward of n feature rows: act [n][193] (pre-activations of the
 * three hidden layers and the output), nu [n]. */
int emx_mlp_forward_f32(const float* params, const float* norm, const float* feat, int n,
                        float* act, float* nu, void* stream);
/* Floats of work emx_mlp_backward_f32 needs for n rows. */
size_t emx_mlp_backward_work(int n);
/* grad[EMX_MLP_NPARAM] = sum over the n rows of dnu * d nu / d params. */
int emx_mlp_backward_f32(const float* params, const float* norm, const float* feat,
                         const float* act, const float* dnu, int n, float* grad, float* work,
                         size_t work_floats, void* stream);
/* One Adam step (bias-corrected; step counts from 1). */
int emx_adam_f32(float* params, const float* grad, float* m, float* v, int n, float lr,
                 float beta1, float beta2, float eps, int step, void* stream);

/* ---- adjoints (transposes) ------------------------------------------------
 * No counterpart in the reference, which has no adjoint.  Each entry is the
 * exact transpose of the forward entry it names (same scheme and metric
 * projection), for gradient-based use and the learned closure's training.
 * Lines of at most 512 points in every direction (EMX_ERR_SIZE otherwise).
 * Outputs are overwritten. */

/* out = D^T g for D = emx_deriv along dir with inverse spacing inv_h. */
int emx_deriv_adj(emx_ctx* ctx, int dir, const double* g, double* out,
                  double inv_h);

/* Transpose of emx_grad: u_bar(c) = sum_d D_d^T dvel_bar(3d+c) (0-based
 * components), temp_bar = sum_d D_d^T dtemp_bar(d); the temperature pair
 * may both be NULL. */
int emx_grad_adj(emx_ctx* ctx, const double* dvel_bar, const double* dtemp_bar,
                 double* u_bar, double* temp_bar, double ihx, double ihy,
                 double ihz);

/* Transpose of emx_divergence with respect to its six flux fields. */
int emx_divergence_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                       const double* jbn, double ihx, double ihy, double ihz,
                       double* flux_bar);

/* Transpose of the linearisation of emx_flux_div_all about (um, u, vis,
 * dvel) (pres enters linearly; jbn is fixed): from the adjoints of rs (5
 * fields) and rsv (5 fields; 1 and 5 unused) to those of um (5, the first
 * is zero), u (3), pres, vis and dvel (9). */
int emx_flux_div_all_adj(emx_ctx* ctx, const double* rs_bar, const double* rsv_bar,
                         const double* um, const double* u, const double* vis,
                         const double* jbn, const double* dvel, double ihx,
                         double ihy, double ihz, double* um_bar, double* u_bar,
                         double* pres_bar, double* vis_bar, double* dvel_bar);

/* SSP-RK3 update of the five conservative variables after the flux
 * divergence of sub-step sub (1..3), solver/rungek_ico.f90:87-91, 152-153,
 * 216-219: um1 = um0 + dt*(rs+rsv); 0.75*um0 + 0.25*um + 0.25*dt*(rs+rsv);
 * um0/3 + 2*um/3 + (2/3)*dt*(rs+rsv).  All arrays 5 comps; um may be NULL
 * for sub-step 1.  (The impeller forcing that follows in RUNGEK, rushton of
 * src/Geometry/rush_ruth_me.f90, is not part of this library.) */
int emx_rk_update(emx_ctx* ctx, int sub, double* um1, const double* um0, const double* um,
                  const double* rs, const double* rsv, double dt);

/* Boundary values of um1 (5 comps) in RUNGEK's order (rungek_ico.f90:96-98):
 * frontera_z_val (src/MSVSingle/frontera_z.f90:219-262), frontera_y_val
 * (frontera_y.f90:220-247), frontera_x_val (frontera_x.f90:223-256).  In
 * place; three boundary-plane launches. */
int emx_frontera_val(emx_ctx* ctx, double* um1);
/* varc_to_varnc (src/MSVSingle/herramientas_incomp2.f90:2-30): temp (and
 * conc, may be NULL) = um1(5), pres = um1(1)/(Mach*Mach), u(c) =
 * um1(1+c)/um1(1).  mach > 0 (EMX_ERR_ARG otherwise). */
int emx_varc_to_varnc(emx_ctx* ctx, const double* um1, double mach, double* u, double* pres,
                      double* temp, double* conc);
/* varnc_to_varc (herramientas_incomp2.f90:32-58): um0(1) = pres*Mach*Mach,
 * um0(5) = temp, um0(1+c) = u(c)*um0(1); pres0 (may be NULL) = pres. */
int emx_varnc_to_varc(emx_ctx* ctx, const double* u, const double* pres, const double* temp,
                      double mach, double* um0, double* pres0);
/* The velocity clip of RUNGEK (solver/rungek_ico.f90:102-123), fused per
 * point: varc_to_varnc of um1, u clipped to [-6, 6], um1(1+c) =
 * u(c)*um1(1)*mask, varc_to_varnc again into u, pres, temp (conc may be
 * NULL).  um1 in place. */
int emx_rk_clip(emx_ctx* ctx, double* um1, const double* mask, double mach, double* u,
                double* pres, double* temp, double* conc);
/* um1 rebuilt after the level-set step (rungek_ico.f90:126-135):
 * um1(1) = pres*Mach*Mach, um1(1+c) = u(c)*um1(1); um1(5) untouched. */
int emx_rk_rebuild(emx_ctx* ctx, double* um1, const double* u, const double* pres, double mach);

/* Compact filter FILTRADO(itf) (src/MSVSingle/herramientas_incomp2.f90:
 * 227-410) with the arrays INIFILTRO builds for cfilt
 * (src/MSVSingle/filtro_incomp.f90:2-166): each of the five fields of um1,
 * times J (jbn 10), is filtered along the three directions in the order
 * itf (1..6) selects -- 1: xi,eta,zeta  2: xi,zeta,eta  3: eta,xi,zeta
 * 4: eta,zeta,xi  5: zeta,xi,eta  6: zeta,eta,xi -- one tridiagonal solve
 * per line with the 11-point right-hand side of FILTRONPER
 * (filtro_incomp.f90:307-362), xi and eta with the non-periodic arrays,
 * zeta with the periodic ones (as the reference passes them), then times
 * 1/J (jbn 11); um1 in place, then u, temp and pres recomputed (u =
 * um1(1+c)/um1(1), temp = um1(5), pres = um1(1)/(Mach*Mach)).
 * 0 <= cfilt < 0.5, mach > 0 (EMX_ERR_ARG otherwise).  The first call per
 * (line length, cfilt) builds the line tables (device allocation). */
int emx_filtrado(emx_ctx* ctx, int itf, double* um1, const double* jbn, double cfilt,
                 double mach, double* u, double* pres, double* temp);

/* Pressure extrapolation into the solid (sobject > 0): advect_CUDA
 * (src/LevelSet/lsTransportEc.cu:24-197 -> interface.cu:341-436).
 * dx,dy,dz are the computational spacings the reference passes
 * (1/deltax).  pres is updated in place. */
int emx_ls_extrapolate(emx_ctx* ctx, double* pres, const double* jbn,
                       const double* sobject, double dx, double dy,
                       double dz, double deltamin);

/* One SSP-RK3 step of WENO5 level-set advection, cuAdvectLsJB
 * (src/LevelSet/lsTransportEc.cu:201-418) on the 5-cell mirror padding
 * that advect_CUDA builds (lsTransportEc.cu:100-166).  phi in place;
 * vel: 3 comps. */
int emx_ls_advect_weno5(emx_ctx* ctx, double* phi, const double* vel,
                        const double* jbn, double dx, double dy, double dz,
                        double dt);

/* Level-set reinitialisation, cuReinitLsJB (src/LevelSet/lsTransportEc.cu:
 * 422-676): on the 5-cell mirror padding, ten iterations of ghost
 * mirroring (cuGhostCellsMirror3d{Z,Y,X}, lsTools.cu:482-585, over the
 * full padded ranges), phi0 = phi, and three SSP-RK3 stages of
 * phi_t = S(phi0)(1 - |grad phi|) with the JB WENO5 derivatives and the
 * Godunov gradient of reini_RS_WENOJB (lsTools.cu:824-900); S = phi0 /
 * sqrt(phi0^2 + dmin^2), dt = deltamin/2, dx..dz the computational spacings.
 * phi in place; dmin a field (the reference's d_dmins). */
int emx_ls_reinit(emx_ctx* ctx, double* phi, const double* jbn, const double* dmin, double dx,
                  double dy, double dz, double deltamin);

/* Smoothed Heaviside of the level set, heviside (solver/advectLS.f90:29-59):
 * with eps = 1.5 deltamax, hevi = 0 where phi < -eps, 1 where phi > eps,
 * and 0.5 + phi/(2 eps) + sin(pi phi/eps)/(2 pi) in between; a NaN phi
 * leaves its hevi unchanged, as in the reference.  deltamax is INISGDM's
 * largest cell size (emx_sgs_geometry returns it); it must be positive and
 * finite (EMX_ERR_ARG).  phi == hevi (in place) is allowed. */
int emx_heviside(emx_ctx* ctx, const double* phi, double deltamax, double* hevi);

/* ---- the reference's on-disk formats (host memory, no device use) -------
 * Fortran unformatted sequential files: each record a 4-byte length, the
 * data, the length again; 4-byte integers, 8-byte reals, arrays
 * column-major.  Records above 2^31-1 bytes (the compilers' continued
 * records) are refused: EMX_ERR_SIZE on write, EMX_ERR_IO on read.
 * Malformed or missing files give EMX_ERR_IO, header mismatches
 * EMX_ERR_SIZE. */
/* field_SS.NNN of GRAVA (solver/gravar_incomp.f90:20-35), read by
 * LEER_CAMPOS (solver/leer_incomp.f90:45-96): (nd) (nx,ny,nz) u1 u2 u3 temp
 * pres.  Reading requires nd = 5 and the given dimensions. */
int emx_io_write_field(const char* path, int nd, int nx, int ny, int nz, const double* u,
                       const double* temp, const double* pres);
int emx_io_read_field(const char* path, int nx, int ny, int nz, double* u, double* temp,
                      double* pres);
int emx_io_field_dims(const char* path, int* nd, int* nx, int* ny, int* nz);
/* The grid file of LEER_MALLA (leer_incomp.f90:1-41): (nvar) (nx,ny,nz)
 * x y z; nvar is written as given and not checked on reading. */
int emx_io_write_grid(const char* path, int nvar, int nx, int ny, int nz, const double* x);
int emx_io_read_grid(const char* path, int nx, int ny, int nz, double* x);
/* time_SS.NNN (gravar_incomp.f90:37-41, LEER_TIME leer_incomp.f90:245-262):
 * one record (dto, it). */
int emx_io_write_time(const char* path, double dto, int it);
int emx_io_read_time(const char* path, double* dto, int* it);
/* One scalar field: header 0 = the field record alone (impellermask_SS.NNN,
 * gravar_incomp.f90:43-47); 1 = (nx,ny,nz) then the field (sobject.field,
 * leer_incomp.f90:98-105); 2 = (count) (nx,ny,nz) then the field (mez.mask
 * of LEER_MASCARA, leer_incomp.f90:215-232; count written as 1, ignored on
 * reading). */
int emx_io_write_scalar(const char* path, int header, int nx, int ny, int nz, const double* a);
int emx_io_read_scalar(const char* path, int header, int nx, int ny, int nz, double* a);

/* ---- host-pointer drop-ins with the reference's bind(C) signatures ------ */

/* solver/deriv_vel_cuda.f90:19-39 -> solver/cuDeriv.cu:529-540.
 * amut, dxyzsgd and jbn are accepted and unused, as in the reference.
 * Scheme C6C.  On error prints and exits(-1), as the reference does. */
void cudaDeriv(double* dvel, double* dtemp, double* u, double* temp,
               double* amut, double* dxyzsgd, double* jbn, double dx,
               double dy, double dz, int Nx, int Ny, int Nz);

/* solver/divergencia_cuda.f90:21-42 -> solver/cuDivergence.cu:107-119.
 * Scheme C6C.  e..gv are not modified. */
void cuDivergence(double* rs, double* rsv, double* e, double* f, double* g,
                  double* ev, double* fv, double* gv, double* jbn, double dx,
                  double dy, double dz, int Nx, int Ny, int Nz);

/* include/LevelSet/interfaceAdvect.h:3-16 -> lsTransportEc.cu:24-35. */
void advect_CUDA(double* pres, double* jbn, double* sobject, double dx,
                 double dy, double dz, int Nx, int Ny, int Nz,
                 double deltamin);

/* Host-pointer forms of the two hot-path routines the reference runs in
 * Fortran on the CPU, for the Fortran shim: flujos(neq)
 * (solver/flujos_compact.f90:1-220) and viscosidad_t+viscosidad
 * (solver/sgdm_incomp.f90:117-395, solver/viscosidad_incomp.f90:1-23).
 * Return a status like the emx_* functions. */
int emx_host_flujos(int neq, const double* um, const double* u,
                    const double* pres, const double* vis, const double* jbn,
                    const double* dvel, double* e, double* f, double* g,
                    double* ev, double* fv, double* gv, int Nx, int Ny,
                    int Nz);
int emx_host_sgdm(const double* u, const double* dvel, const double* jbn,
                  const double* sgd, double c4, double* amut, double* vis,
                  double* vis5, int Nx, int Ny, int Nz);

#ifdef __cplusplus
}
#endif
#endif
