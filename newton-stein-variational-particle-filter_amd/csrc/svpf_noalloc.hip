// svpf_noalloc.hip -- svpf_run_sequence_device_noalloc: the device loop with its
// per-tick records staged in a fixed array compiled into the library instead of
// a per-state buffer grown with hipMalloc, for callers that cannot allocate.
//
// `scratch` is a file-scope device array; each GPU the library runs on holds its
// own copy at its own address, until the process ends.  In host code its name
// is only a stand-in: the device address comes from hipGetSymbolAddress, looked
// up once per device with that device current, and only that address is passed
// to kernels and copies.  Every call on a GPU shares the array, which keeps the
// last call's records: the records of ticks [0, T) are written by the call's own
// kernel before anything reads them, and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/svpf.h"
#include "svpf_internal.h"
#include "svpf_types.h"

static __device__ __align__(256) float scratch[(size_t)SVPF_NOALLOC_MAX_TICKS * SVPF_OUTW];

namespace {
constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// The current device's copy of `scratch` (looked up on the first call on that device).
hipError_t scratch_address(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}
}  // namespace

extern "C" void svpf_run_sequence_device_noalloc(SVPFState* s, const float* d_obs, int T, const SVPFParams* p,
                                                 float* d_ll, float* d_vol) {
    if (!s || !s->impl) return;
    if (T > SVPF_NOALLOC_MAX_TICKS) {
        svpf_internal_set_error(s, SVPF_RC_TOO_LARGE, hipSuccess);
        return;
    }
    float* records = nullptr;
    hipError_t e = scratch_address(&records);
    if (e != hipSuccess) {
        svpf_internal_set_error(s, SVPF_RC_HIP, e);
        return;
    }
    svpf_internal_run_sequence_device(s, d_obs, T, p, d_ll, d_vol, records);
}
