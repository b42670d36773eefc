// vt_noalloc.hip -- the Roe flux on independent records without any device
// allocation: the records, fluxes and codes are staged through one static
// device array instead of per-call hipMalloc buffers (vt_numerical_flux).
//
// The array is held by every GPU the library runs on until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress on the
// first call made with that device current) and only that address is used on
// the host side: the array's name in host code is a host-side stand-in, not a
// device address.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vt_api.h"
#include "vt_ctx.hpp"
#include "vt_kernels.hpp"

namespace {

constexpr int kMaxRecords = VT_NOALLOC_MAX_RECORDS;
// records [kMaxRecords][10] doubles, then fluxes [kMaxRecords][4] doubles, then
// codes [kMaxRecords] ints (kMaxRecords / 2 doubles of space).
constexpr size_t kRecOff = 0;
constexpr size_t kFluxOff = kRecOff + 10 * static_cast<size_t>(kMaxRecords);
constexpr size_t kCodeOff = kFluxOff + 4 * static_cast<size_t>(kMaxRecords);
constexpr size_t kScratchDoubles = kCodeOff + (static_cast<size_t>(kMaxRecords) + 1) / 2;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};

// Device address of this device's copy of the array (looked up on first use).
double* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (!vti::hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch")) return nullptr;
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int vt_numerical_flux_noalloc(const double* rec, double* flux, int* codes, int n) {
    if (n < 0 || n > kMaxRecords) return VT_ERR_ARG;
    if (n == 0) return VT_OK;
    if (!rec || !flux || !codes) return VT_ERR_ARG;
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return VT_ERR_ARG;
    double* base = scratch_ptr(dev);
    if (!base) return VT_ERR_HIP;
    double* din = base + kRecOff;
    double* dout = base + kFluxOff;
    int* dcodes = reinterpret_cast<int*>(base + kCodeOff);
    // Every output entry of the n records is written by the kernel, so nothing
    // left from an earlier call is read back.
    HIP_TRY(hipMemcpy(din, rec, sizeof(double) * 10 * static_cast<size_t>(n), hipMemcpyHostToDevice));
    HIP_TRY(vt::launch_flux(din, dout, dcodes, n, nullptr));
    HIP_TRY(hipMemcpy(flux, dout, sizeof(double) * 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(codes, dcodes, sizeof(int) * static_cast<size_t>(n), hipMemcpyDeviceToHost));
    return VT_OK;
}
