// m3drt_sc_sweep_noalloc: the sweep of m3drt_sc_sweep with its scratch (the
// per-cell step lengths the sweep kernels take, e.g. the (dt500, 1/chi500) pairs
// of the straight-line kernel) in a fixed device array instead of a buffer the
// library allocates, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).  The kernels are m3drt_sc_sweep's.
//
// The array is a file-scope __device__ variable: every device that loads the
// library holds its own copy until the process ends.  In host code its name is
// only a stand-in for the symbol, so its address on a device is resolved once, with
// that device current (noalloc_setup, called by m3drt_create), and only that
// address is handed to kernels.  Every call on one device shares the array and
// leaves its contents behind; the sweep kernels write all of the scratch they read
// before reading it, so nothing must start at zero.  Calls that overlap on one
// device (two streams, two threads) race on it.
#include "m3drt_internal.hpp"
#include <mutex>

namespace {
constexpr size_t kNoallocBytes = size_t(32) << 20;  // 32 MiB: the 120 x 120 x 230 box takes 26.6 MB
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace m3drt {
namespace {
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(M3DRT_ELIMIT, "m3drt: device index beyond " + std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return M3DRT_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
  if (!p) return set_error(M3DRT_EDEVICE, "hipGetSymbolAddress(scratch) returned no address");
  g_ptr[device] = p;
  return M3DRT_OK;
}

void *noalloc_scratch(int device, size_t *bytes) {
  *bytes = kNoallocBytes;
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace m3drt

extern "C" int m3drt_sc_sweep_noalloc(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (!ctx) return m3drt::set_error(M3DRT_EINVAL, "m3drt_sc_sweep_noalloc: ctx is NULL");
  size_t cap = 0;
  void *p = m3drt::noalloc_scratch(ctx->device, &cap);
  if (!p) return m3drt::set_error(M3DRT_EDEVICE, "m3drt_sc_sweep_noalloc: the fixed scratch of this device is unresolved");
  // this call's scratch is the fixed array; the context's own buffer is left as it is
  void *own = ctx->scratch;
  const size_t own_bytes = ctx->scratch_bytes;
  const bool own_captured = ctx->scratch_captured;
  ctx->scratch = p;
  ctx->scratch_bytes = cap;
  ctx->scratch_fixed = true;
  const int rc = m3drt::sc_sweep_checked(ctx, d, b, "m3drt_sc_sweep_noalloc");
  ctx->scratch = own;
  ctx->scratch_bytes = own_bytes;
  ctx->scratch_captured = own_captured;
  ctx->scratch_fixed = false;
  return rc;
}
