// fvh_static.hip -- the library-owned workspace of the allocation-free encode
// entry point (fvh_encode_launch_static).
//
// One file-scope device array per GPU, holding the encoder's workspace
// (frame and tile descriptors, residuals, order-0 scratch, rank bytes, work
// counters).  Host code never passes the array's name to a kernel: in host
// code that name is a host-side stand-in, not a device address.  The device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and cached; every later call uses the cached pointer.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "fvh_kernels.h"

namespace fvh {

static __device__ __align__(256) uint8_t scratch[kStaticWorkspaceBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t static_workspace(hipStream_t stream, void** out) {
  *out = nullptr;
  int device = 0;
  hipError_t err = hipStreamGetDevice(stream, &device);
  if (err != hipSuccess) return err;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device] == nullptr) {
    int current = 0;
    err = hipGetDevice(&current);
    if (err != hipSuccess) return err;
    if (current != device && (err = hipSetDevice(device)) != hipSuccess) return err;
    void* ptr = nullptr;
    err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) {
      const hipError_t back = hipSetDevice(current);
      if (err == hipSuccess) err = back;
    }
    if (err != hipSuccess) return err;
    if (ptr == nullptr || ((uintptr_t)ptr & 255) != 0) return hipErrorInvalidSymbol;
    g_scratch[device] = ptr;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

}  // namespace fvh
