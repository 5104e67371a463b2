// Batched multi-grid stage: many independent grids ("boxes") of uneven
// sizes advanced together.  One stage = the velocity gradient fused with
// the velocity divergence (ProbGradDiv, three directional line passes) and
// one WENO5 / SSP-RK3 level-set step (pad, three stages, the last writing
// the interior straight into each box's phi).  Every
// phase is ONE launch over all boxes: workgroups are dealt out box by box
// from per-phase prefix tables (blk_start), each box with its own line
// tables, geometry and data.  The phase sequence (seven launches) is
// captured once in a HIP graph and replayed.  The line-solver shape of
// each direction (lines per wavefront LG and points per work-item C, i.e.
// the split between the serial Thomas sweep inside a chunk and the
// parallel carry combination across chunks) is chosen by timing the
// candidates on the bound data (emx_batch_tune).
//
// The reference has no batched form; the per-box operators it batches are
// DERIV_VEL (solver/deriv_vel.f90), divergencia (solver/divergencia.f90)
// and cuAdvectLsJB (src/LevelSet/lsTransportEc.cu:201-418).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ops_common.hpp"

namespace emx {

struct BatchShape {
  int lg, C;
};

// Shapes the batched line pass is built for (NV = 4 fields per point):
// lg lines per wavefront of the strided kind, or lg = 0 for the wavefront
// kind (one wavefront per line, direction 0 only: its lines are contiguous,
// so each load instruction of a wave reads consecutive points), where each
// box takes the shortest of the chunk lengths 1, 2, 4, 8 up to C that fits.
static const BatchShape kBatchShapes[] = {{8, 4},  {8, 8},   {16, 4}, {16, 8},
                                          {32, 4}, {32, 8},  {64, 8}, {64, 16},
                                          {0, 1},  {0, 2},   {0, 4},  {0, 8}};
constexpr int kWaveLinesPerBlock = 16;  // four per wave

static int lg_kind(int lg) {
  return lg == 0 ? KIND_WAVE
                 : (lg == 64 ? KIND_STRIDED
                             : (lg == 32 ? KIND_STRIDED32
                                         : (lg == 16 ? KIND_STRIDED16 : KIND_STRIDED8)));
}

static size_t wave_lds(int C) {
  return (size_t)(5 * 64 * C + 768 + 4 * 64 * (C + 1)) * sizeof(double);
}

// small: every workgroup of the pass has at most 256 work-items.
template <int S, int C, int LG>
static void* batched_kernel(bool small) {
  return small ? (void*)k_lines_batched<S, C, LG, ProbGradDiv, 256>
               : (void*)k_lines_batched<S, C, LG, ProbGradDiv, 1024>;
}

template <int S>
static void* batched_wave_kernel(int C) {
  switch (C) {
    case 1: return (void*)k_lines_batched_wave<S, 1, ProbGradDiv>;
    case 2: return (void*)k_lines_batched_wave<S, 2, ProbGradDiv>;
    case 4: return (void*)k_lines_batched_wave<S, 4, ProbGradDiv>;
    default: return (void*)k_lines_batched_wave<S, 8, ProbGradDiv>;
  }
}

template <int S>
static void* batched_kernel_s(int lg, int C, bool small) {
  if (lg == 0) return batched_wave_kernel<S>(C);
  if (lg == 8) return C == 4 ? batched_kernel<S, 4, 8>(small) : batched_kernel<S, 8, 8>(small);
  if (lg == 16) return C == 4 ? batched_kernel<S, 4, 16>(small) : batched_kernel<S, 8, 16>(small);
  if (lg == 32) return C == 4 ? batched_kernel<S, 4, 32>(small) : batched_kernel<S, 8, 32>(small);
  return C == 8 ? batched_kernel<S, 8, 64>(small) : batched_kernel<S, 16, 64>(small);
}

// threads: the pass's workgroup size (the strided kinds' bound: at most
// 256 selects the kernel compiled for 256 work-items).
static void* batched_kernel_for(int scheme, int lg, int C, int threads) {
  const bool small = threads <= 256;
  switch (scheme) {
    case EMX_SCHEME_C4: return batched_kernel_s<EMX_SCHEME_C4>(lg, C, small);
    case EMX_SCHEME_C6F: return batched_kernel_s<EMX_SCHEME_C6F>(lg, C, small);
    default: return batched_kernel_s<EMX_SCHEME_C6C>(lg, C, small);
  }
}

}  // namespace emx

struct emx_batch {
  int device = 0;
  int scheme = EMX_SCHEME_C6C;
  int nbox = 0;
  std::vector<emx::Dims> dims;
  hipStream_t user_stream = nullptr;
  hipStream_t stream = nullptr;  // own stream: graph capture needs one
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  struct Box {
    const double* u = nullptr;
    const double* jbn = nullptr;
    double* phi = nullptr;
    double* dvel = nullptr;
    double* div = nullptr;
    double ih[3] = {0, 0, 0};
    double dxyz[3] = {0, 0, 0};
    double dt = 0.0;
    bool bound = false;
  };
  std::vector<Box> box;
  // scratch
  double* acc = nullptr;        // sum of N
  double* ls = nullptr;         // sum of 3 * NG
  std::vector<long long> acc_off, ls_off;
  // line passes, per direction
  emx::BatchShape shape[3];
  std::vector<emx::BatchLines<emx::ProbGradDiv>> lines_h[3];
  emx::BatchLines<emx::ProbGradDiv>* lines_d[3] = {nullptr, nullptr, nullptr};
  std::vector<emx::BatchWaveLines<emx::ProbGradDiv>> waves_h;  // direction 0
  emx::BatchWaveLines<emx::ProbGradDiv>* waves_d = nullptr;
  int* blk_d[3] = {nullptr, nullptr, nullptr};
  int nblk[3] = {0, 0, 0};
  int threads[3] = {0, 0, 0};
  size_t lds[3] = {0, 0, 0};
  // level-set phases
  std::vector<emx::LsBatchGrid> ls_h;
  emx::LsBatchGrid* ls_d = nullptr;
  int* blk_pad_d = nullptr;
  int nblk_pad = 0;
  // state
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace emx {
namespace {

void drop_graph(emx_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void free_batch(emx_batch* b) {
  if (!b) return;
  drop_graph(b);
  (void)hipFree(b->acc);
  (void)hipFree(b->ls);
  for (int d = 0; d < 3; d++) {
    (void)hipFree(b->lines_d[d]);
    (void)hipFree(b->blk_d[d]);
  }
  (void)hipFree(b->waves_d);
  (void)hipFree(b->ls_d);
  (void)hipFree(b->blk_pad_d);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Orders the batch's stream after everything queued so far on the caller's
// stream, then waits until the batch's stream is idle.  The descriptor
// tables are rewritten with synchronous copies on the null stream, which
// does not wait for the batch's non-blocking stream: no kernel of the batch
// may be reading them, and the caller's inputs must be complete before a
// pass of the autotuner reads them.
int quiesce(emx_batch* b) {
  EMX_TRY_HIP(hipEventRecord(b->ev_in, b->user_stream));
  EMX_TRY_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  EMX_TRY_HIP(hipStreamSynchronize(b->stream));
  return EMX_OK;
}

int box_n(const Dims& d, int dir) { return dir == 0 ? d.nx : (dir == 1 ? d.ny : d.nz); }

bool shape_fits(const emx_batch* b, int dir, BatchShape s) {
  if (s.lg == 0 && dir != 0) return false;
  for (const Dims& d : b->dims) {
    const int n = box_n(d, dir);
    if ((n + s.C - 1) / s.C > (s.lg == 0 ? 64 : 1024 / s.lg)) return false;
  }
  return true;
}

// Direction 0 on the wavefront kind: kWaveLinesPerBlock lines per workgroup,
// each box with its shortest chunk (LDS sized for the longest, s.C).
int plan_wave_lines(emx_batch* b, BatchShape s) {
  auto& H = b->waves_h;
  H.assign(b->nbox, BatchWaveLines<ProbGradDiv>{});
  std::vector<int> blk(b->nbox + 1, 0);
  for (int i = 0; i < b->nbox; i++) {
    const Dims& d = b->dims[i];
    const DirGeom g = dir_geom(d, 0);
    BatchWaveLines<ProbGradDiv>& L = H[i];
    // the shortest chunk of 1, 2, 3, 4, 8 (at most s.C) whose 64 lanes hold the line
    int C = 1;
    while (C < s.C && g.n > 64 * C) C = C == 2 ? 3 : (C == 3 ? 4 : 2 * C);
    EMX_TRY(get_line_tab(b->scheme, g.n, C, KIND_WAVE, &L.T));
    L.nlines = g.nlines;
    L.nyl = g.nyl;
    L.sj = g.sj;
    L.sk = g.sk;
    L.sl = g.sl;
    L.nblk = (int)((g.nlines + kWaveLinesPerBlock - 1) / kWaveLinesPerBlock);
    const emx_batch::Box& B = b->box[i];
    L.ih = B.ih[0];
    L.prob = ProbGradDiv{B.u, B.jbn, B.dvel, b->acc + b->acc_off[i], B.div, d.N(), 0};
    blk[i + 1] = blk[i] + L.nblk;
  }
  b->shape[0] = s;
  b->nblk[0] = blk[b->nbox];
  b->threads[0] = 256;
  b->lds[0] = wave_lds(s.C);
  void* k = batched_kernel_for(b->scheme, 0, s.C, 256);
  if (b->lds[0] > 65536)
    EMX_TRY_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)b->lds[0]));
  EMX_TRY_HIP(hipMemcpy(b->waves_d, H.data(), H.size() * sizeof(H[0]), hipMemcpyHostToDevice));
  EMX_TRY_HIP(hipMemcpy(b->blk_d[0], blk.data(), blk.size() * sizeof(int),
                        hipMemcpyHostToDevice));
  drop_graph(b);
  return EMX_OK;
}

// Tables, descriptors and launch geometry of direction dir under shape s.
int plan_lines(emx_batch* b, int dir, BatchShape s) {
  if (!shape_fits(b, dir, s)) {
    set_error("batch: lines along direction %d do not fit %d lines x %d points per item",
              dir, s.lg, s.C);
    return EMX_ERR_SIZE;
  }
  if (s.lg == 0) return plan_wave_lines(b, s);
  auto& H = b->lines_h[dir];
  H.assign(b->nbox, BatchLines<ProbGradDiv>{});
  std::vector<int> blk(b->nbox + 1, 0);
  int maxthreads = 64;
  size_t maxlds = 0;
  for (int i = 0; i < b->nbox; i++) {
    const Dims& d = b->dims[i];
    const DirGeom g = dir_geom(d, dir);
    BatchLines<ProbGradDiv>& L = H[i];
    EMX_TRY(get_line_tab(b->scheme, g.n, s.C, lg_kind(s.lg), &L.T));
    L.nlanes = g.nlanes;
    L.gx = (g.nlanes + s.lg - 1) / s.lg;
    L.si = g.si;
    L.so = g.so;
    L.sl = g.sl;
    const emx_batch::Box& B = b->box[i];
    L.ih = B.ih[dir];
    L.prob = ProbGradDiv{B.u, B.jbn, B.dvel, b->acc + b->acc_off[i], B.div, d.N(), dir};
    blk[i + 1] = blk[i] + L.gx * g.nouter;
    maxthreads = std::max(maxthreads, ((L.T.P * s.lg + 63) / 64) * 64);
    const size_t lds = (size_t)(5 * L.T.n + 1 + 2 * L.T.P * L.T.KB +
                                ProbGradDiv::NV * L.T.P * s.lg) * sizeof(double);
    maxlds = std::max(maxlds, lds);
  }
  b->shape[dir] = s;
  b->nblk[dir] = blk[b->nbox];
  b->threads[dir] = maxthreads;
  b->lds[dir] = maxlds;
  void* k = batched_kernel_for(b->scheme, s.lg, s.C, maxthreads);
  if (maxlds > 65536)
    EMX_TRY_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
  EMX_TRY_HIP(hipMemcpy(b->lines_d[dir], H.data(), H.size() * sizeof(H[0]),
                        hipMemcpyHostToDevice));
  EMX_TRY_HIP(hipMemcpy(b->blk_d[dir], blk.data(), blk.size() * sizeof(int),
                        hipMemcpyHostToDevice));
  drop_graph(b);
  return EMX_OK;
}

int launch_pass(emx_batch* b, int dir) {
  void* k = batched_kernel_for(b->scheme, b->shape[dir].lg, b->shape[dir].C,
                               b->threads[dir]);
  const void* grids = b->shape[dir].lg == 0 ? (const void*)b->waves_d : (const void*)b->lines_d[dir];
  const int* blk = b->blk_d[dir];
  int ngrid = b->nbox;
  void* args[] = {(void*)&grids, (void*)&blk, (void*)&ngrid};
  EMX_TRY_HIP(hipLaunchKernel(k, dim3(b->nblk[dir]), dim3(b->threads[dir]), args, b->lds[dir],
                              b->stream));
  return EMX_OK;
}

int launch_stage(emx_batch* b) {
  for (int d = 0; d < 3; d++) EMX_TRY(launch_pass(b, d));
  return op_batch_ls_step(b->stream, b->ls_d, b->blk_pad_d, b->nblk_pad, b->nbox);
}

// Descriptors of every phase from the bound boxes.
int commit(emx_batch* b) {
  for (int i = 0; i < b->nbox; i++)
    if (!b->box[i].bound) {
      set_error("batch: box %d has no bound data", i);
      return EMX_ERR_ARG;
    }
  EMX_TRY(quiesce(b));
  for (int d = 0; d < 3; d++) EMX_TRY(plan_lines(b, d, b->shape[d]));
  std::vector<int> bp(b->nbox + 1, 0);
  b->ls_h.resize(b->nbox);
  for (int i = 0; i < b->nbox; i++) {
    const Dims& d = b->dims[i];
    const emx_batch::Box& B = b->box[i];
    const long long NG = ls_padded_points(d.nx, d.ny, d.nz);
    double* base = b->ls + b->ls_off[i];
    LsBatchGrid& G = b->ls_h[i];
    G.nx = d.nx;
    G.ny = d.ny;
    G.nz = d.nz;
    G.phi = B.phi;
    G.vel = B.u;
    G.jbn = B.jbn;
    G.P = base;
    G.T1 = base + NG;
    G.T2 = base + 2 * NG;
    G.iX = 1.0 / B.dxyz[0];
    G.iY = 1.0 / B.dxyz[1];
    G.iZ = 1.0 / B.dxyz[2];
    G.dt = B.dt;
    bp[i + 1] = bp[i] + (int)((NG + 255) / 256);
  }
  b->nblk_pad = bp[b->nbox];
  EMX_TRY_HIP(hipMemcpy(b->ls_d, b->ls_h.data(), b->nbox * sizeof(LsBatchGrid),
                        hipMemcpyHostToDevice));
  EMX_TRY_HIP(hipMemcpy(b->blk_pad_d, bp.data(), bp.size() * sizeof(int), hipMemcpyHostToDevice));
  EMX_TRY_HIP(hipStreamSynchronize(b->stream));
  b->dirty = false;
  return EMX_OK;
}

}  // namespace
}  // namespace emx

using emx::set_error;

extern "C" {

int emx_batch_create(emx_batch** out, int nbox, const int* dims, int scheme, void* stream) {
  if (!out || !dims || nbox <= 0) {
    set_error("emx_batch_create: bad arguments");
    return EMX_ERR_ARG;
  }
  *out = nullptr;
  if (scheme < 0 || scheme > 2) {
    set_error("emx_batch_create: unknown scheme %d", scheme);
    return EMX_ERR_ARG;
  }
  auto* b = new emx_batch();
  b->scheme = scheme;
  b->nbox = nbox;
  b->user_stream = (hipStream_t)stream;
  long long accN = 0, lsN = 0;
  for (int i = 0; i < nbox; i++) {
    const emx::Dims d{dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]};
    const int nmin = emx::scheme_min_n(scheme);
    if (d.nx < nmin || d.ny < nmin || d.nz < nmin || d.nx > 512 || d.ny > 512 || d.nz > 512) {
      set_error("emx_batch_create: box %d is %d x %d x %d (each side %d..512)", i, d.nx, d.ny,
                d.nz, nmin);
      delete b;
      return EMX_ERR_SIZE;
    }
    b->dims.push_back(d);
    b->acc_off.push_back(accN);
    b->ls_off.push_back(lsN);
    accN += d.N();
    lsN += 3 * emx::ls_padded_points(d.nx, d.ny, d.nz);
  }
  b->box.resize(nbox);
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
    set_error("emx_batch_create: stream or event creation failed");
    emx::free_batch(b);
    return EMX_ERR_HIP;
  }
  bool ok = hipMalloc(&b->acc, accN * sizeof(double)) == hipSuccess &&
            hipMalloc(&b->ls, lsN * sizeof(double)) == hipSuccess &&
            hipMalloc(&b->ls_d, nbox * sizeof(emx::LsBatchGrid)) == hipSuccess &&
            hipMalloc(&b->blk_pad_d, (nbox + 1) * sizeof(int)) == hipSuccess &&
            hipMalloc(&b->waves_d, nbox * sizeof(emx::BatchWaveLines<emx::ProbGradDiv>)) ==
                hipSuccess;
  for (int d = 0; d < 3 && ok; d++)
    ok = hipMalloc(&b->lines_d[d], nbox * sizeof(emx::BatchLines<emx::ProbGradDiv>)) ==
             hipSuccess &&
         hipMalloc(&b->blk_d[d], (nbox + 1) * sizeof(int)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    set_error("emx_batch_create: device allocation failed (%lld + %lld doubles)", accN, lsN);
    emx::free_batch(b);
    return EMX_ERR_NOMEM;
  }
  // default shape: 16 lines per wavefront, the shortest chunk that fits;
  // direction 0 on the wavefront kind with the shortest chunk that fits
  for (int d = 0; d < 3; d++) {
    b->shape[d] = emx::BatchShape{16, 4};
    if (!emx::shape_fits(b, d, b->shape[d])) b->shape[d] = emx::BatchShape{16, 8};
  }
  for (int C : {8, 4, 2, 1})
    if (emx::shape_fits(b, 0, emx::BatchShape{0, C})) b->shape[0] = emx::BatchShape{0, C};
  *out = b;
  return EMX_OK;
}

int emx_batch_destroy(emx_batch* b) {
  emx::free_batch(b);
  return EMX_OK;
}

int emx_batch_bind(emx_batch* b, int i, const double* u, const double* jbn, double* phi,
                   double* dvel, double* div, double ihx, double ihy, double ihz, double dx,
                   double dy, double dz, double dt) {
  if (!b || i < 0 || i >= b->nbox) {
    set_error("emx_batch_bind: bad batch or box index %d", i);
    return EMX_ERR_ARG;
  }
  if (!u || !jbn || !phi || !dvel || !div) {
    set_error("emx_batch_bind: null array for box %d", i);
    return EMX_ERR_ARG;
  }
  emx_batch::Box& B = b->box[i];
  B.u = u;
  B.jbn = jbn;
  B.phi = phi;
  B.dvel = dvel;
  B.div = div;
  B.ih[0] = ihx;
  B.ih[1] = ihy;
  B.ih[2] = ihz;
  B.dxyz[0] = dx;
  B.dxyz[1] = dy;
  B.dxyz[2] = dz;
  B.dt = dt;
  B.bound = true;
  b->dirty = true;
  emx::drop_graph(b);
  return EMX_OK;
}

int emx_batch_set_plan(emx_batch* b, int dir, int lg, int C) {
  if (!b || dir < 0 || dir > 2) {
    set_error("emx_batch_set_plan: bad arguments");
    return EMX_ERR_ARG;
  }
  bool known = false;
  for (const auto& s : emx::kBatchShapes) known |= (s.lg == lg && s.C == C);
  if (!known) {
    set_error("emx_batch_set_plan: shape %d lines x %d points is not built", lg, C);
    return EMX_ERR_ARG;
  }
  if (!emx::shape_fits(b, dir, emx::BatchShape{lg, C})) {
    set_error("emx_batch_set_plan: lines along %d do not fit %d x %d", dir, lg, C);
    return EMX_ERR_SIZE;
  }
  b->shape[dir] = emx::BatchShape{lg, C};
  b->dirty = true;
  emx::drop_graph(b);
  return EMX_OK;
}

int emx_batch_get_plan(const emx_batch* b, int dir, int* lg, int* C) {
  if (!b || dir < 0 || dir > 2 || !lg || !C) {
    set_error("emx_batch_get_plan: bad arguments");
    return EMX_ERR_ARG;
  }
  *lg = b->shape[dir].lg;
  *C = b->shape[dir].C;
  return EMX_OK;
}

int emx_batch_tune(emx_batch* b, int reps) {
  if (!b) {
    set_error("emx_batch_tune: null batch");
    return EMX_ERR_ARG;
  }
  if (reps < 1) reps = 5;
  if (b->dirty) EMX_TRY(emx::commit(b));
  EMX_TRY(emx::quiesce(b));
  hipEvent_t e0, e1;
  EMX_TRY_HIP(hipEventCreate(&e0));
  EMX_TRY_HIP(hipEventCreate(&e1));
  int rc = EMX_OK;
  for (int d = 0; d < 3 && rc == EMX_OK; d++) {
    float best = 1e30f;
    emx::BatchShape pick = b->shape[d];
    for (const auto& s : emx::kBatchShapes) {
      if (!emx::shape_fits(b, d, s)) continue;
      rc = emx::quiesce(b);  // the previous shape's launches read the tables
      if (rc == EMX_OK) rc = emx::plan_lines(b, d, s);
      if (rc != EMX_OK) break;
      // passes 0..d so that the accumulator a pass reads is written
      for (int w = 0; w < 2 && rc == EMX_OK; w++)
        for (int q = 0; q <= d && rc == EMX_OK; q++) rc = emx::launch_pass(b, q);
      if (rc != EMX_OK) break;
      (void)hipEventRecord(e0, b->stream);
      for (int r = 0; r < reps && rc == EMX_OK; r++) rc = emx::launch_pass(b, d);
      (void)hipEventRecord(e1, b->stream);
      if (hipEventSynchronize(e1) != hipSuccess) {
        set_error("emx_batch_tune: timing failed");
        rc = EMX_ERR_HIP;
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        pick = s;
      }
    }
    if (rc == EMX_OK) rc = emx::quiesce(b);
    if (rc == EMX_OK) rc = emx::plan_lines(b, d, pick);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the timed passes wrote dvel and div: the caller's stream sees them after
  if (rc == EMX_OK && hipEventRecord(b->ev_out, b->stream) == hipSuccess)
    (void)hipStreamWaitEvent(b->user_stream, b->ev_out, 0);
  return rc;
}

int emx_batch_stage(emx_batch* b, int use_graph) {
  if (!b) {
    set_error("emx_batch_stage: null batch");
    return EMX_ERR_ARG;
  }
  if (b->dirty) EMX_TRY(emx::commit(b));
  EMX_TRY_HIP(hipEventRecord(b->ev_in, b->user_stream));
  EMX_TRY_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  if (use_graph) {
    if (!b->exec) {
      EMX_TRY_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = emx::launch_stage(b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc != EMX_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc == EMX_OK) set_error("emx_batch_stage: graph capture failed: %s", hipGetErrorString(e));
        return rc != EMX_OK ? rc : EMX_ERR_HIP;
      }
      b->graph = g;
      EMX_TRY_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    EMX_TRY_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    EMX_TRY(emx::launch_stage(b));
  }
  EMX_TRY_HIP(hipEventRecord(b->ev_out, b->stream));
  EMX_TRY_HIP(hipStreamWaitEvent(b->user_stream, b->ev_out, 0));
  return EMX_OK;
}

int emx_batch_set_stream(emx_batch* b, void* stream) {
  if (!b) {
    set_error("emx_batch_set_stream: null batch");
    return EMX_ERR_ARG;
  }
  b->user_stream = (hipStream_t)stream;
  return EMX_OK;
}

}  // extern "C"
