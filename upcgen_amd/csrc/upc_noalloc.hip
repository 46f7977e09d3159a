// upc_flux_form_noalloc: the per-point form-factor flux (upc_flux_form) on a
// file-scope device scratch array instead of the context's work buffers, for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// The scratch array is one per device (each device that runs the library
// holds kScratchBytes until the process ends).  Its device address is taken
// with hipGetSymbolAddress once per device, on that device, and that address
// (never the array's host-side name) is what the kernels receive.  Every call
// on a device uses the same array: calls that overlap on one device race.
#include "upc_internal.h"
#include "upc_qags.h"

#include <mutex>

namespace upc {
namespace {

constexpr int kMaxPoints = 1 << 18;  // points per call (262,144)

// Layout of the scratch array (offsets in bytes, each 256-aligned).
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kOffIdx = 0;
constexpr size_t kOffOverflow = align256(kOffIdx + sizeof(int) * kMaxPoints);
constexpr size_t kOffCounters = align256(kOffOverflow + sizeof(long long) * kMaxPoints);
constexpr size_t kScratchBytes = align256(kOffCounters + sizeof(unsigned long long) * 8);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (taken on first use).
int scratch_ptr(int dev, unsigned char** out)
{
  if (dev < 0 || dev >= kMaxDevices) return set_error("device index out of range"), UPC_EDEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error("hipGetSymbolAddress(scratch) returned null"), UPC_EDEVICE;
    g_ptr[dev] = p;
  }
  *out = (unsigned char*)g_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace upc

using namespace upc;

extern "C" int upc_flux_form_noalloc(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out,
                                     int* d_ier, void* stream)
{
  if (!c || n < 0 || (n && (!d_b || !d_k || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  if (n > kMaxPoints)
    return set_error("upc_flux_form_noalloc takes at most " + std::to_string(kMaxPoints) + " points per call (" +
                     std::to_string(n) + " given)"),
           UPC_EINVAL;
  if (n == 0) return 0;
  int dev = -1;
  UPC_HIP(hipGetDevice(&dev));
  if (dev != c->device)
    return set_error("the current device is not the context's device (" + std::to_string(dev) + " vs " +
                     std::to_string(c->device) + ")"),
           UPC_EINVAL;
  unsigned char* base = nullptr;
  if (int rc = scratch_ptr(dev, &base)) return rc;
  const FluxFormBufs w{(int*)(base + kOffIdx),
                       (long long*)(base + kOffOverflow),
                       (unsigned long long*)(base + kOffCounters)};
  return enqueue_flux_form(c, d_b, d_k, n, d_out, d_ier, w, (hipStream_t)stream);
}
