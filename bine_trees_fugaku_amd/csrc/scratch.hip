// scratch.hip -- the device workspace of the allocation-free entry point
// (bine_allreduce_noalloc).
//
// One file-scope device array per GPU the library runs on, held until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: in host code that name is only a stand-in for the device symbol.  The
// device address is resolved once per device, with that device current, by
// hipGetSymbolAddress, and kept per device; kernels receive that address.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace bine {
constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB per device
constexpr int kScratchDevices = 64;
}  // namespace bine

static __device__ __align__(256) unsigned char scratch[bine::kScratchBytes];

namespace bine {

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

size_t scratch_bytes() { return kScratchBytes; }

// Resolve the device address of `scratch` on `device` (once per device).  The caller's current device is
// restored.  Returns the HIP error of the lookup; the address stays unset on failure.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    if ((e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

// The device address resolved by scratch_init, or nullptr.
void* scratch_ptr(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace bine
