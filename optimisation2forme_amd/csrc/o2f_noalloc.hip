// o2f_hllc_flux_batch without device allocation: the faces are staged
// through a fixed device array of this library instead of hipMalloc'd
// memory, for callers that must not allocate (serving loops, harnesses that
// forbid hipMalloc).
//
// `scratch` holds one batch of at most kNoAllocMaxFaces faces: qL (4 n),
// qR (4 n), the normals (2 n) and the fluxes (4 n), SoA, 14 doubles per face
// (1.75 MiB).  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device memory: its device address is obtained once per device with
// hipGetSymbolAddress and only that address is passed on.  Calls on one GPU
// share the array; a per-device lock serialises calls from several host
// threads (without it they would race).  Every flux the call copies back is
// written by the kernel in the same call, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "o2f.h"
#include "o2f_kernels.hpp"

namespace o2f {
extern thread_local std::string g_error;
}

namespace {

constexpr int kNoAllocMaxFaces = O2F_NOALLOC_MAX_FACES;
constexpr int kDoublesPerFace = 14;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kDoublesPerFace * kNoAllocMaxFaces];

struct DeviceScratch {
    std::once_flag once;
    hipError_t status = hipSuccess;
    double *ptr = nullptr;  // device address of this device's `scratch`
    std::mutex lock;
};
DeviceScratch g_dev[kMaxDevices];

// The current device's copy of `scratch`, looked up on its first use.
hipError_t device_scratch(DeviceScratch *&d) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    d = &g_dev[dev];
    std::call_once(d->once, [d]() {
        void *p = nullptr;
        d->status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d->status == hipSuccess) d->ptr = static_cast<double *>(p);
    });
    return d->status;
}

}  // namespace

extern "C" int o2f_hllc_flux_batch_noalloc(const double *qL, const double *qR, const double *normal, double gamma,
                                           double *F, int n) {
    using o2f::g_error;
    if (!qL || !qR || !normal || !F || n < 0) {
        g_error = "o2f_hllc_flux_batch_noalloc: bad argument";
        return O2F_ERR_ARG;
    }
    if (n > kNoAllocMaxFaces) {
        g_error = "o2f_hllc_flux_batch_noalloc: " + std::to_string(n) + " faces exceed the maximum of " +
                  std::to_string(kNoAllocMaxFaces) + " (O2F_NOALLOC_MAX_FACES); use o2f_hllc_flux_batch";
        return O2F_ERR_ARG;
    }
    if (n == 0) return O2F_OK;
    DeviceScratch *d = nullptr;
    hipError_t e = device_scratch(d);
    if (e != hipSuccess) {
        g_error = std::string("o2f_hllc_flux_batch_noalloc: scratch address: ") + hipGetErrorString(e);
        return O2F_ERR_HIP;
    }
    std::lock_guard<std::mutex> hold(d->lock);
    double *s = d->ptr;
    const size_t bytes = (size_t)n * sizeof(double);
    e = hipMemcpy(s, qL, 4 * bytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s + 4 * (size_t)n, qR, 4 * bytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s + 8 * (size_t)n, normal, 2 * bytes, hipMemcpyHostToDevice);
    if (!e) e = o2f::launch_hllc_batch(s, s + 4 * (size_t)n, s + 8 * (size_t)n, gamma, s + 10 * (size_t)n, n, nullptr);
    if (!e) e = hipMemcpy(F, s + 10 * (size_t)n, 4 * bytes, hipMemcpyDeviceToHost);
    if (e) {
        g_error = std::string("o2f_hllc_flux_batch_noalloc: ") + hipGetErrorString(e);
        return O2F_ERR_HIP;
    }
    return O2F_OK;
}
