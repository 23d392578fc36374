// lag2d_division_check_noalloc: lag2d_division_check without any device
// allocation.  The device copies of a, b and q (the staging buffer the
// allocating entry point gets from hipMalloc) live in `scratch`, a
// file-scope device array of LAG2D_NOALLOC_MAX_N * 3 doubles that every GPU
// the library runs on holds until the process ends.
//
// In host code the name `scratch` is only a host-side stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, called once per
// device with that device current, and only that address is passed to
// copies and kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lag2d.h"
#include "lag2d_dev.h"
#include "lag2d_kernels.h"

using namespace lag2d;

static __device__ __align__(256) double scratch[3 * LAG2D_NOALLOC_MAX_N];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device.
int scratch_ptr(double **out, std::string *err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return LAG2D_EDEVICE;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        *err = "device ordinal out of range";
        return LAG2D_EDEVICE;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return LAG2D_EDEVICE;
        }
        g_ptr[dev] = (double *)p;
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace

// Error reporting shares lag2d_last_error with the rest of the library.
int lag2d_set_error(int code, const std::string &msg);

extern "C" int lag2d_division_check_noalloc(int n, const double *a, const double *b, double *q)
{
    if (n < 0 || (n > 0 && (!a || !b || !q))) return lag2d_set_error(LAG2D_EINVAL, "bad argument");
    if (n > LAG2D_NOALLOC_MAX_N)
        return lag2d_set_error(LAG2D_EINVAL, "n = " + std::to_string(n) + " exceeds the scratch limit of " +
                                                 std::to_string(LAG2D_NOALLOC_MAX_N) + " quotients");
    if (n == 0) return 0;
    double *s = nullptr;
    std::string err;
    int rc = scratch_ptr(&s, &err);
    if (rc) return lag2d_set_error(rc, err);
    // scratch = [a | b | q], each n doubles; q is written in full by the
    // kernel, so nothing needs clearing from an earlier call
    const size_t bytes = (size_t)n * sizeof(double);
    hipError_t e = hipMemcpy(s, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s + n, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_division_check, dim3((n + LAG2D_BLOCK - 1) / LAG2D_BLOCK), dim3(LAG2D_BLOCK), 0, 0,
                           (const double *)s, (const double *)(s + n), s + 2 * (size_t)n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(q, s + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return lag2d_set_error(LAG2D_EDEVICE, std::string("division check (no alloc): ") + hipGetErrorString(e));
    return 0;
}
