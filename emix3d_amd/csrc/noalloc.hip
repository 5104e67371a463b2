// emx_dynsmag_noalloc: the dynamic Smagorinsky model (emx_dynsmag's two
// kernels) for callers that cannot allocate device memory -- code under
// HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  Its two scratch fields (L:M and M:M) live in a file-scope
// device array of fixed size instead of context memory; the caller passes
// the inputs and outputs and a stream.
//
// The array's host-side name is not a device address: its address on each
// device comes from hipGetSymbolAddress, once per device (with that device
// current) on the first call there, and is kept per device.  Every call on
// a device shares that device's array: calls that overlap on one device
// (two streams, two host threads) race.  Both scratch fields are written
// in full by the first kernel before the second reads them, so nothing
// needs clearing between calls.
#include <mutex>

#include "ops_common.hpp"

namespace {

constexpr long long kMaxPoints = 1LL << 21;  // 128^3 grid points
constexpr int kMaxDevices = 64;

// L:M then M:M, kMaxPoints doubles each (32 MiB per device).
static __device__ __align__(256) double scratch[2 * kMaxPoints];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_of_current_device(double** out) {
  int dev = 0;
  EMX_TRY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    emx::set_error("emx_dynsmag_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    return EMX_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      emx::set_error("emx_dynsmag_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                     hipGetErrorString(e));
      return EMX_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return EMX_OK;
}

}  // namespace

extern "C" int emx_dynsmag_noalloc(int nx, int ny, int nz, const double* u, const double* dvel,
                                   const double* jbn, const double* delta, double* nut,
                                   double c4, double* vis, void* stream) {
  if (!u || !dvel || !jbn || !delta || !nut) {
    emx::set_error("emx_dynsmag_noalloc: null input or output pointer");
    return EMX_ERR_ARG;
  }
  if (nx < 1 || ny < 1 || nz < 1) {
    emx::set_error("emx_dynsmag_noalloc: bad grid %d x %d x %d", nx, ny, nz);
    return EMX_ERR_SIZE;
  }
  const long long N = (long long)nx * ny * nz;
  if (N > kMaxPoints) {
    emx::set_error("emx_dynsmag_noalloc: %d x %d x %d = %lld points exceed the static scratch "
                   "of %lld points (use emx_dynsmag)", nx, ny, nz, N, kMaxPoints);
    return EMX_ERR_SIZE;
  }
  double* s = nullptr;
  EMX_TRY(scratch_of_current_device(&s));
  return emx::launch_dynsmag(nullptr, static_cast<hipStream_t>(stream), emx::Dims{nx, ny, nz}, u,
                             dvel, jbn, delta, nut, c4, vis, s, s + N);
}
