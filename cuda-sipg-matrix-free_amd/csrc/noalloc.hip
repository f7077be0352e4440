// noalloc.hip -- sipg_vec_dot_noalloc: the operator-free dot product of
// sipg_vec_dot with its scratch in a fixed device array instead of a
// per-call allocation, for callers that cannot allocate device memory.
//
// The scratch holds the dot kernel's block sums (at most 1024: the grid of
// the vector kernels is capped there) and the reduced sum.  Every word the
// reduction reads is written by the dot kernel of the same call, so nothing
// has to be cleared between calls.  Each device holds its own copy of the
// array at its own address, looked up once per device with that device
// current; kernels get that address, never the array's host-side name.
// Calls on one device share the array: calls that overlap in time on one
// device (several host threads, or several streams) race on it.
#include "sipg_internal.hpp"

#include <mutex>

using namespace sipg_internal;

namespace {

constexpr int kMaxDotPartials = 1024;
constexpr int kScratchLen = kMaxDotPartials + 1;   // [0]: the sum, [1..]: block sums
static __device__ __align__(256) double scratch[kScratchLen];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
double *g_ptr[kMaxDevices];
hipError_t g_err[kMaxDevices];

// Device address of this device's scratch (looked up on the first call on
// the device, with it current).
int scratch_address(double **out) {
    int dev = 0;
    SIPG_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return SIPG_ERR_UNSUPPORTED;
    std::call_once(g_once[dev], [dev] {
        void *p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? static_cast<double *>(p) : nullptr;
    });
    if (g_err[dev] != hipSuccess || !g_ptr[dev]) {
        set_last_hip_error(g_err[dev] != hipSuccess ? g_err[dev] : hipErrorInvalidSymbol);
        return SIPG_ERR_HIP;
    }
    *out = g_ptr[dev];
    return SIPG_OK;
}

}  // namespace

extern "C" int sipg_vec_dot_noalloc(int dtype, int64_t n, const void *a, const void *b, double *result,
                                    void *stream) {
    if (!a || !b || !result || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    const int nb = vec_num_blocks(n);
    if (nb < 1 || nb > kMaxDotPartials) return SIPG_ERR_INVALID;     // cannot happen: the grid is capped
    double *s = nullptr;
    int rc = scratch_address(&s);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = launch_dot_partials(dtype, a, b, n, s + 1, nb, st))) return rc;
    if ((rc = launch_reduce_partials(s + 1, nb, s, st))) return rc;
    SIPG_HIP_TRY(hipMemcpyAsync(result, s, sizeof(double), hipMemcpyDeviceToHost, st));
    SIPG_HIP_TRY(hipStreamSynchronize(st));
    return SIPG_OK;
}
