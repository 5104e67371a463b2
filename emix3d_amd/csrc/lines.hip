// Line-solver tables: the Thomas factors of the constant tridiagonal matrix
// of each compact scheme, and the chunk-carry constants that let a line be
// solved as P independent chunks plus a carry exchange.
//
// The matrix of every line of a direction is the same (it depends only on
// the scheme and the line length), so the factors the reference recomputes
// per line (solver/derivada.f90:111-122, src/Schemes/derivCompact.cu:199-212)
// are computed once here, in FP64 on the host, and kept on the device.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "compact.hpp"

namespace emx {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const char* last_error() { return g_last_error.c_str(); }

// n >= 5 for every scheme: the C4 matrix is singular at n = 4 (its last
// pivot bet_3 = 1 - 3*(1/3) vanishes), where the reference divides by zero.
int scheme_min_n(int scheme) {
  (void)scheme;
  return 5;
}

// Coefficients a (sub), b (diag), c (super):
//  C4 : solver/derivada.f90:10-62
//  C6F: src/MSVSingle/derivada_incomp.f90:10-62
//  C6C: src/Schemes/derivCompact.cu:165-189
//  filter: src/MSVSingle/filtro_incomp.f90:30-139 (INIFILTRO sets cfilt2..5
//  to zero on both of its branches, so the non-periodic arrays are zero on
//  the five rows at each end and cfilt on rows 6..n-5, 1-based)
int host_thomas(int scheme, int n, std::vector<double>& a,
                std::vector<double>& b, std::vector<double>& c, double cfilt) {
  if (scheme < 0 || scheme > FILTER_P) return EMX_ERR_ARG;
  if (n < scheme_min_n(scheme)) return EMX_ERR_SIZE;
  a.assign(n, 0.0);
  b.assign(n, 1.0);
  c.assign(n, 0.0);
  if (scheme == FILTER_NP || scheme == FILTER_P) {
    for (int l = 0; l < n; l++)
      if (scheme == FILTER_P || (l + 1 >= 6 && l + 1 <= n - 5)) a[l] = c[l] = cfilt;
  } else if (scheme == EMX_SCHEME_C4) {
    for (int l = 1; l < n - 1; l++) a[l] = c[l] = 1. / 4.;
    a[0] = 0.0; a[n - 1] = 3.0;
    c[0] = 3.0; c[n - 1] = 0.0;
  } else {
    const double edge = scheme == EMX_SCHEME_C6F ? 2.0 : 3.0;
    for (int l = 2; l < n - 2; l++) a[l] = c[l] = 1.0 / 3.0;
    a[0] = 0.0; a[1] = 0.25; a[n - 2] = 0.25; a[n - 1] = edge;
    c[0] = edge; c[1] = 0.25; c[n - 2] = 0.25; c[n - 1] = 0.0;
  }
  return EMX_OK;
}

namespace {

struct TabKey {
  int dev, scheme, n, C, kind;
  uint64_t aux;  // bits of cfilt (filter schemes)
  bool operator<(const TabKey& o) const {
    return std::tie(dev, scheme, n, C, kind, aux) <
           std::tie(o.dev, o.scheme, o.n, o.C, o.kind, o.aux);
  }
};

std::mutex g_tab_mu;
std::map<TabKey, LineTab> g_tabs;  // device memory held until process exit

// B^T diagonals of scheme S: column l of B from the right-hand-side rule
// applied to unit vectors.
template <int S>
void rhs_transposed(int n, std::vector<double>& BT) {
  BT.assign((size_t)7 * n, 0.0);
  for (int l = 0; l < n; l++)
    for (int k = 0; k < 7; k++) {
      const int r = l + k - 3;  // row of B
      if (r < 0 || r >= n) continue;
      double w[7] = {0, 0, 0, 0, 0, 0, 0};
      w[3 + (l - r)] = 1.0;  // f_l is w[l-r] of row r
      BT[(size_t)7 * l + k] = rhs_row<S>(w + 3, r, n, 1.0);
    }
}

int build_tab(int scheme_flags, int n, int C, int kind, LineTab* t, double cfilt) {
  const int scheme = scheme_flags & ~TAB_TRANSPOSED;
  const bool transposed = (scheme_flags & TAB_TRANSPOSED) != 0;
  if (transposed && scheme > EMX_SCHEME_C6C) return EMX_ERR_ARG;
  std::vector<double> a, b, c;
  EMX_TRY(host_thomas(scheme, n, a, b, c, cfilt));
  std::vector<double> BT;
  if (transposed) {
    // row l of A^T: A[l-1][l], A[l][l], A[l+1][l] = c_{l-1}, b_l, a_{l+1}
    std::vector<double> at(n, 0.0), ct(n, 0.0);
    for (int l = 1; l < n; l++) at[l] = c[l - 1];
    for (int l = 0; l + 1 < n; l++) ct[l] = a[l + 1];
    a.swap(at);
    c.swap(ct);
    switch (scheme) {
      case EMX_SCHEME_C4: rhs_transposed<EMX_SCHEME_C4>(n, BT); break;
      case EMX_SCHEME_C6F: rhs_transposed<EMX_SCHEME_C6F>(n, BT); break;
      default: rhs_transposed<EMX_SCHEME_C6C>(n, BT); break;
    }
  }
  const int P = (n + C - 1) / C;
  const int pmax = kind == KIND_WAVE ? 64 : kind == KIND_STRIDED ? 16
                 : kind == KIND_STRIDED32 ? 32 : kind == KIND_STRIDED8 ? 128 : 64;
  if (P > pmax) {
    set_error("line of %d points needs %d chunks of %d, more than the %d "
              "this launch shape (kind %d) holds", n, P, C, pmax, kind);
    return EMX_ERR_SIZE;
  }
  std::vector<double> bet(n), gam(n + 1, 0.0), ib(n), m(n), H(n), G(n);
  bet[0] = b[0];
  for (int l = 1; l < n; l++) {
    gam[l] = c[l - 1] / bet[l - 1];
    bet[l] = b[l] - a[l] * gam[l];
  }
  for (int l = 0; l < n; l++) {
    ib[l] = 1.0 / bet[l];
    m[l] = a[l] * ib[l];
  }
  std::vector<double> Hend(P), Gs(P);
  for (int ch = 0; ch < P; ch++) {
    int s = ch * C, e = std::min(s + C, n);
    double h = 1.0;
    for (int l = s; l < e; l++) { h *= -m[l]; H[l] = h; }
    for (int l = s; l < e; l++) {
      double g = 1.0;
      for (int t = l + 1; t <= e; t++) g *= -gam[t];  // gam[n] = 0
      G[l] = g;
    }
    Hend[ch] = H[e - 1];
    Gs[ch] = G[s];
  }
  std::vector<double> Wf((size_t)P * P, 0.0), Wb((size_t)P * P, 0.0);
  for (int ch = 0; ch < P; ch++) {
    for (int t = 0; t < ch; t++) {
      double w = 1.0;
      for (int u = t + 1; u <= ch - 1; u++) w *= Hend[u];
      Wf[(size_t)ch * P + t] = w;
    }
    for (int t = ch + 1; t < P; t++) {
      double w = 1.0;
      for (int u = ch + 1; u <= t - 1; u++) w *= Gs[u];
      Wb[(size_t)ch * P + t] = w;
    }
  }
  // carries from chunks further than K away weigh less than 1e-24
  int K = 0;
  for (int ch = 0; ch < P; ch++)
    for (int t = 0; t < P; t++) {
      const double w = t < ch ? Wf[(size_t)ch * P + t] : (t > ch ? Wb[(size_t)ch * P + t] : 0.0);
      if (std::fabs(w) >= 1e-24) K = std::max(K, std::abs(ch - t));
    }
  std::vector<double> Af(6 * 64, 0.0), Ab(6 * 64, 0.0);
  for (int k = 0; k < 6; k++) {
    int d = 1 << k;
    for (int ch = 0; ch < 64; ch++) {
      if (ch < P && ch >= d) {
        double w = 1.0;
        for (int u = ch - d + 1; u <= ch; u++) w *= Hend[u];
        Af[k * 64 + ch] = w;
      }
      if (ch < P && ch + d < P) {
        double w = 1.0;
        for (int u = ch; u <= ch + d - 1; u++) w *= Gs[u];
        Ab[k * 64 + ch] = w;
      }
    }
  }
  // banded carry weights of the strided kinds: WfB[c*KB+k] = Wf[c][c-1-k],
  // WbB[c*KB+k] = Wb[c][c+1+k], k < KB
  const int KB = std::max(K, 1);
  std::vector<double> WfB((size_t)P * KB, 0.0), WbB((size_t)P * KB, 0.0);
  for (int ch = 0; ch < P; ch++)
    for (int k = 0; k < KB; k++) {
      if (ch - 1 - k >= 0) WfB[(size_t)ch * KB + k] = Wf[(size_t)ch * P + ch - 1 - k];
      if (ch + 1 + k < P) WbB[(size_t)ch * KB + k] = Wb[(size_t)ch * P + ch + 1 + k];
    }
  // one device block, in the order the kernels stage it into LDS:
  //   ib m gam(n+1) H G | Af Ab (wave kind) | WfB WbB (strided kinds) | Wf Wb
  size_t off[13];
  off[0] = 0;
  off[1] = off[0] + n;                 // m
  off[2] = off[1] + n;                 // gam (n+1)
  off[3] = off[2] + n + 1;             // H
  off[4] = off[3] + n;                 // G
  off[5] = off[4] + n;                 // Af
  off[6] = off[5] + 6 * 64;            // Ab
  off[7] = off[6] + 6 * 64;            // WfB
  off[8] = off[7] + (size_t)P * KB;    // WbB
  off[9] = off[8] + (size_t)P * KB;    // Wf
  off[10] = off[9] + (size_t)P * P;    // Wb
  off[11] = off[10] + (size_t)P * P;  // BT (transposed tables)
  off[12] = off[11] + BT.size();
  std::vector<double> host(off[12]);
  memcpy(&host[off[0]], ib.data(), n * sizeof(double));
  memcpy(&host[off[1]], m.data(), n * sizeof(double));
  memcpy(&host[off[2]], gam.data(), (n + 1) * sizeof(double));
  memcpy(&host[off[3]], H.data(), n * sizeof(double));
  memcpy(&host[off[4]], G.data(), n * sizeof(double));
  memcpy(&host[off[5]], Af.data(), 6 * 64 * sizeof(double));
  memcpy(&host[off[6]], Ab.data(), 6 * 64 * sizeof(double));
  memcpy(&host[off[7]], WfB.data(), (size_t)P * KB * sizeof(double));
  memcpy(&host[off[8]], WbB.data(), (size_t)P * KB * sizeof(double));
  memcpy(&host[off[9]], Wf.data(), (size_t)P * P * sizeof(double));
  memcpy(&host[off[10]], Wb.data(), (size_t)P * P * sizeof(double));
  if (!BT.empty()) memcpy(&host[off[11]], BT.data(), BT.size() * sizeof(double));
  double* dev = nullptr;
  if (hipMalloc(&dev, off[12] * sizeof(double)) != hipSuccess) {
    set_error("hipMalloc of a %zu-byte line table failed", off[12] * sizeof(double));
    return EMX_ERR_NOMEM;
  }
  EMX_TRY_HIP(hipMemcpy(dev, host.data(), off[12] * sizeof(double),
                        hipMemcpyHostToDevice));
  t->n = n;
  t->C = C;
  t->P = P;
  t->kind = kind;
  t->K = K;
  t->KB = KB;
  t->blob = dev;
  t->ib = dev + off[0];
  t->m = dev + off[1];
  t->gam = dev + off[2];
  t->H = dev + off[3];
  t->G = dev + off[4];
  t->Af = dev + off[5];
  t->Ab = dev + off[6];
  t->WfB = dev + off[7];
  t->WbB = dev + off[8];
  t->Wf = dev + off[9];
  t->Wb = dev + off[10];
  t->BT = BT.empty() ? nullptr : dev + off[11];
  return EMX_OK;
}

}  // namespace

int get_line_tab(int scheme, int n, int C, int kind, LineTab* out, double cfilt) {
  int dev = 0;
  EMX_TRY_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tab_mu);
  uint64_t bits = 0;
  if ((scheme & ~TAB_TRANSPOSED) >= FILTER_NP) memcpy(&bits, &cfilt, sizeof bits);
  TabKey key{dev, scheme, n, C, kind, bits};
  auto it = g_tabs.find(key);
  if (it != g_tabs.end()) {
    *out = it->second;
    return EMX_OK;
  }
  LineTab t{};
  EMX_TRY(build_tab(scheme, n, C, kind, &t, cfilt));
  g_tabs[key] = t;
  *out = t;
  return EMX_OK;
}

// ---------------------------------------------------------------------------
// Per-launch timing.  Events are created on demand while the context
// records and reused by later recordings; a failure to create or record
// one stops the recording (timing is a diagnostic, never an error of the
// operator).
LaunchTimer::LaunchTimer(emx_ctx* c, const char* label, int dir, double bytes)
    : c_(c), on_(c && c->timing &&
                 (c->timing_label.empty() || c->timing_label == label)) {
  if (!on_) return;
  const size_t need = 2 * (c->timed.size() + 1);
  while (c->ev_pool.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      c->timing = false;
      on_ = false;
      return;
    }
    c->ev_pool.push_back(e);
  }
  if (hipEventRecord(c->ev_pool[need - 2], c->stream) != hipSuccess) {
    c->timing = false;
    on_ = false;
    return;
  }
  c->timed.push_back(TimedLaunch{label, dir, bytes});
}

LaunchTimer::~LaunchTimer() {
  if (!on_) return;
  if (hipEventRecord(c_->ev_pool[2 * c_->timed.size() - 1], c_->stream) != hipSuccess) {
    c_->timed.pop_back();
    c_->timing = false;
  }
}

}  // namespace emx
