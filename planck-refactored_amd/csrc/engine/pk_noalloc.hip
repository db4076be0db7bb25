// pk_noalloc.hip — pk_fock_jk_noalloc: the J/K build for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The per-build scratch (scaled density, J and K accumulators, and the FP32
// copy of the density that mixed precision reads) lives in a file-scope device
// array of fixed size, 3.5 kMaxElems doubles: 28 MiB on every GPU the library
// runs on, held until the process ends.  Every call on a
// GPU shares that GPU's copy, so calls that overlap on one GPU race; each call
// clears the accumulators it uses (fill kernels, no memset).  On the host the
// array's name is only a stand-in: the kernels get the device address taken
// once per device by hipGetSymbolAddress, with that device current.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pk_fock.h"
#include "pk_internal.h"

namespace {

constexpr size_t kMaxElems = (size_t)1 << 20;  // packed matrix elements: one molecule of <= 1024 functions
constexpr int kMaxDevices = 64;

}  // namespace

// scaled density, J and K accumulators (doubles), FP32 density copy (floats)
static __device__ __align__(256) double scratch[3 * kMaxElems + kMaxElems / 2];

namespace {

std::mutex g_mu;
double *g_dev_scratch[kMaxDevices] = {};

// device address of this device's copy of `scratch`, taken on first use
int scratch_address(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pk_set_error(PK_ERR_HIP, (std::string("pk_fock_jk_noalloc: hipGetDevice: ") + hipGetErrorString(e)).c_str());
    if (dev < 0 || dev >= kMaxDevices) return pk_set_error(PK_ERR_UNSUPPORTED, "pk_fock_jk_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_dev_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pk_set_error(PK_ERR_HIP, (std::string("pk_fock_jk_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e)).c_str());
        g_dev_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_dev_scratch[dev];
    return PK_OK;
}

}  // namespace

extern "C" int pk_fock_jk_noalloc(pk_ctx *ctx, const double *P, double *J, double *K) {
    if (!ctx || !P) return pk_set_error(PK_ERR_ARG, "pk_fock_jk_noalloc: null context or density");
    double *s = nullptr;
    const int rc = scratch_address(&s);
    if (rc != PK_OK) return rc;
    return pk_build_with_scratch(ctx, P, J, K, s, kMaxElems, "pk_fock_jk_noalloc");
}
