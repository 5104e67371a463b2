// Internal declarations of libemix3d_hip: status handling, the per-context
// state and the line-solver tables.  Everything here is FP64 and uses the
// reference's Fortran column-major layout: element (i,j,k,c) of an
// (nx,ny,nz,nc) array sits at i + nx*(j + ny*(k + nz*c))
// (src/Engine/cuArray.cu:119-124, solver/cuDeriv.cu:588-590).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/emix3d.h"

namespace emx {

// A pointer read from a descriptor in memory (not a kernel argument) is a
// generic one, and accesses through it compile to FLAT instructions, which
// count against the LDS wait counter too: every LDS wait of a line solve
// then also waits for them.  gref(p) is *p addressed as global memory (p
// must point to device memory), so the access is a global load or store.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T& gref(T* p) {
  return *(__attribute__((address_space(1))) T*)p;
}


void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define EMX_TRY_HIP(call)                                                   \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::emx::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,   \
                       hipGetErrorString(e_));                              \
      return EMX_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

#define EMX_TRY(call)                 \
  do {                                \
    int rc_ = (call);                 \
    if (rc_ != EMX_OK) return rc_;    \
  } while (0)

// Workgroups go to the eight XCDs round robin (XCD = id mod 8), each XCD
// with its own L2.  Renumbered by xcd_block, each XCD works through one
// contiguous eighth of the block index space, so the stencil neighbours of
// its points (rows and planes away, or the neighbouring tiles) sit in the
// same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned q = nb / 8;
  return b < 8 * q ? (b % 8) * q + b / 8 : b;
}

// Kinds of line solver launch (see compact.hpp).
enum LineKind {
  KIND_WAVE = 0,       // one wavefront per line, lanes along the line
  KIND_STRIDED = 1,    // 64 lines per wave, chunks over the waves
  KIND_STRIDED32 = 2,  // 32 lines per wave, two chunk slots per wave
  KIND_STRIDED16 = 3,  // 16 lines per wave, four chunk slots per wave
  KIND_STRIDED8 = 4,   // 8 lines per wave, eight chunk slots per wave
  // flux divergence only: the projected fluxes of the direction formed by
  // one pointwise kernel into scratch, then one single-field solve per flux
  // with the direction's derivative plan
  KIND_MATERIALIZED = 5,
};

// Device tables of one (scheme, n, chunk) combination.  All arrays live in
// one device allocation; pointers are device addresses.  With the Thomas
// factors of the constant tridiagonal matrix (bet_l, gam_l) the solve of a
// line chunked into P pieces of C points is:
//   forward  y_l = r_l*ib_l - m_l*y_{l-1}, computed per chunk from a zero
//            carry and corrected by y_l += H_l * Y_{c-1};
//   backward x_l = y_l - gam_{l+1}*x_{l+1}, per chunk from a zero carry,
//            corrected by x_l += G_l * X_{c+1};
// the chunk carries Y/X come from the weights Wf/Wb (strided kind) or from
// log2 wave-scan steps with multipliers Af/Ab (wave kind).
struct LineTab {
  int n;      // line length
  int C;      // points per chunk
  int P;      // chunks per line
  int kind;   // LineKind
  int K;      // chunk carries further than K chunks away are below 1e-24
  int KB;     // max(K, 1): width of the banded carry weights
  // all arrays below live in one device block (blob), laid out as
  //   ib[n] m[n] gam[n+1] H[n] G[n] | Af[384] Ab[384] | WfB[P*KB] WbB[P*KB]
  //   | Wf[P*P] Wb[P*P]
  // so that a kernel stages [0, 5n+1) plus its kind's weights into LDS.
  const double* blob;
  const double* ib;   // [n]    1/bet_l
  const double* m;    // [n]    a_l/bet_l
  const double* gam;  // [n+1]  gam_l (gam_0 = gam_n = 0)
  const double* H;    // [n]    prod_{t=s_c..l} (-m_t)
  const double* G;    // [n]    prod_{t=l+1..e_c} (-gam_t)
  const double* Af;   // [6*64] wave kind: forward scan multipliers
  const double* Ab;   // [6*64] wave kind: backward scan multipliers
  const double* WfB;  // [P*KB] strided kinds: WfB[c*KB+k] = Wf[c][c-1-k]
  const double* WbB;  // [P*KB] strided kinds: WbB[c*KB+k] = Wb[c][c+1+k]
  const double* Wf;   // [P*P]  forward carry weights (full)
  const double* Wb;   // [P*P]  backward carry weights (full)
  // tables of the transposed matrix only (TAB_TRANSPOSED): B^T as seven
  // diagonals, BT[7l+k] = B[l+k-3][l] (B = the right-hand-side matrix,
  // inverse spacing excluded); nullptr otherwise
  const double* BT;   // [7n]
};

// Scheme flag of get_line_tab: the factors and carry weights of A^T
// (sub- and super-diagonals exchanged) plus the B^T diagonals, for the
// transposed (adjoint) derivative.
constexpr int TAB_TRANSPOSED = 16;

// Line "schemes" of the compact filter (not public scheme ids): the
// tridiagonal matrix of FILTRONPER with INIFILTRO's non-periodic arrays
// (xi, eta) or periodic arrays (zeta), and its 11-point right-hand side.
// Their tables depend on cfilt as well.
constexpr int FILTER_NP = 3;
constexpr int FILTER_P = 4;

// Builds (once per device) and returns the table.  Thread-safe.  scheme
// may carry TAB_TRANSPOSED; cfilt is the filter parameter of the filter
// schemes (ignored otherwise).
int get_line_tab(int scheme, int n, int C, int kind, LineTab* out, double cfilt = 0.0);
// Host copy of the Thomas factors (for tests and the adjoint).
int host_thomas(int scheme, int n, std::vector<double>& a,
                std::vector<double>& b, std::vector<double>& c, double cfilt = 0.0);
int scheme_min_n(int scheme);

// Launch configuration of one direction of one context.
// The line-solver shape is chosen per operation class and direction.
enum PlanOp { PLAN_DERIV = 0, PLAN_FLUX = 1, PLAN_OPS = 2 };

struct DirPlan {
  LineTab tab;
  int kind;
  int C;
};

struct Dims {
  int nx, ny, nz;
  __host__ __device__ long long N() const { return (long long)nx * ny * nz; }
};

// Per-launch timing (emx_timing_begin / emx_timing_end): two HIP events on
// the context stream around each launch, with the launch's algorithmic
// bytes (every field it must read or write once, 8 B per point).
struct TimedLaunch {
  const char* label;
  int dir;
  double bytes;
};
struct TimingEntry {
  const char* label;
  int dir;
  int launches;
  double ms;
  double bytes;
};

}  // namespace emx

// The opaque context of the C ABI.
struct emx_ctx {
  int device = 0;
  int scheme = EMX_SCHEME_C6C;
  emx::Dims d{0, 0, 0};
  hipStream_t stream = nullptr;
  emx::DirPlan plan[emx::PLAN_OPS][3];  // [operation class][direction]
  // transposed-operator tables per direction (empty when a line is longer
  // than the transposed solve supports)
  emx::LineTab adj[3] = {};
  // persistent scratch (device): eight accumulators for the fused
  // flux-divergence, two fields for the dynamic eddy-viscosity model.
  double* acc = nullptr;       // 8*N
  double* fluxbuf = nullptr;   // 8*N per materialized direction: projected fluxes
  int fluxbuf_dirs = 0;        // directions fluxbuf holds
  double* sgs_tmp = nullptr;   // 2*N (dynamic-model contractions, flags)
  void* mlp_pack = nullptr;    // BF16 operands of the closure network
  // level-set scratch (allocated lazily on first level-set call)
  double* ls_buf = nullptr;
  size_t ls_bytes = 0;
  // staging for the host-pointer entry points
  double* stage = nullptr;
  size_t stage_bytes = 0;
  // per-launch timing (off unless emx_timing_begin)
  bool timing = false;
  std::string timing_label;                 // record only this label (empty: all)
  std::vector<hipEvent_t> ev_pool;          // 2 per recorded launch
  std::vector<emx::TimedLaunch> timed;
  std::vector<emx::TimingEntry> timing_summary;
};

namespace emx {
// Brackets one kernel launch with timing events when the context records.
class LaunchTimer {
 public:
  LaunchTimer(emx_ctx* c, const char* label, int dir, double bytes);
  ~LaunchTimer();
  LaunchTimer(const LaunchTimer&) = delete;
  LaunchTimer& operator=(const LaunchTimer&) = delete;

 private:
  emx_ctx* c_;
  bool on_;
};
}  // namespace emx

// One pointwise launch statement, timed when the context records.
#define EMX_TIMED(ctx, label, bytes, ...)                      \
  do {                                                         \
    ::emx::LaunchTimer timer_((ctx), (label), -1, (bytes));    \
    __VA_ARGS__;                                               \
  } while (0)
