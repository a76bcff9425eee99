// Entry point that allocates no device memory: msa_ke_packed_noalloc, the
// packed element stiffness of msa_ke_packed with its device staging (element
// coordinates, moduli, packed Ke, per-element status) in a file-scope device
// array instead of hipMalloc'ed buffers.
//
// Each GPU holds its own copy of `scratch` at its own address until the
// process ends.  In host code the array's name is only a host-side stand-in,
// so its device address is looked up once per device with
// hipGetSymbolAddress (noalloc_setup, called by msa_ctx_create with that
// device current) and only that address is used afterwards: as kernel
// arguments and as copy destinations and sources.
//
// Every call on one GPU shares that GPU's array.  The kernel writes every
// output entry it copies back (all 300 packed values and the status of each
// element), so nothing has to be cleared between calls; calls that overlap
// on one GPU (two threads, or two contexts on the same device) race on the
// array and must be serialised by the caller.
#include <mutex>

#include "msa_common.hpp"

namespace msa {

// Layout of one call's staging, in doubles: xyz8[kNoallocMaxElem][24] |
// E[kNoallocMaxElem] | ke300[kNoallocMaxElem][300] | status (ints).
constexpr size_t kNaXyz = 0;
constexpr size_t kNaE = kNaXyz + (size_t)kNoallocMaxElem * 24;
constexpr size_t kNaKe = kNaE + (size_t)kNoallocMaxElem;
constexpr size_t kNaStatus = kNaKe + (size_t)kNoallocMaxElem * 300;
constexpr size_t kNaDoubles = kNaStatus + ((size_t)kNoallocMaxElem + 1) / 2;

static __device__ __align__(256) double scratch[kNaDoubles];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
double *g_na_ptr[kNaMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

size_t noalloc_scratch_bytes() { return kNaDoubles * sizeof(double); }

// Look up this device's address of the array (the device must be current).
int noalloc_setup(int device)
{
    MSA_CHECK(device >= 0 && device < kNaMaxDevices, MSA_ERR_ARG,
              "device %d beyond the %d devices the no-allocation entry points track", device,
              kNaMaxDevices);
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (g_na_ptr[device]) return MSA_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                  hipGetErrorString(e));
        return MSA_ERR_HIP;
    }
    g_na_ptr[device] = (double *)p;
    return MSA_OK;
}

}  // namespace msa

using namespace msa;

extern "C" int msa_ke_packed_noalloc(msa_ctx *ctx, int n_elem, const double *xyz8, const double *E,
                                     double nu, double *ke300, int *elem_status)
{
    MSA_CHECK(ctx && xyz8 && E && ke300 && n_elem >= 0, MSA_ERR_ARG,
              "msa_ke_packed_noalloc: bad arguments");
    MSA_CHECK(n_elem <= kNoallocMaxElem, MSA_ERR_ARG,
              "msa_ke_packed_noalloc: %d elements exceed the limit of %d per call (the "
              "entry point stages into a fixed %zu-byte device array)",
              n_elem, kNoallocMaxElem, noalloc_scratch_bytes());
    if (n_elem == 0) return MSA_OK;
    MSA_CHECK(ctx->device >= 0 && ctx->device < kNaMaxDevices, MSA_ERR_ARG, "device %d",
              ctx->device);
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_na_mutex);
        base = g_na_ptr[ctx->device];
    }
    MSA_CHECK(base, MSA_ERR_STATE,
              "msa_ke_packed_noalloc: no device address of the staging array on device %d "
              "(the context was not created by msa_ctx_create)",
              ctx->device);
    MSA_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    double *dx = base + kNaXyz, *dE = base + kNaE, *dk = base + kNaKe;
    int *ds = (int *)(base + kNaStatus);
    const size_t n = (size_t)n_elem;
    MSA_HIP(hipMemcpyAsync(dx, xyz8, sizeof(double) * 24 * n, hipMemcpyHostToDevice, s));
    MSA_HIP(hipMemcpyAsync(dE, E, sizeof(double) * n, hipMemcpyHostToDevice, s));
    MSA_HIP(launch_ke_packed(s, n_elem, dx, dE, nu, dk, ds));
    MSA_HIP(hipMemcpyAsync(ke300, dk, sizeof(double) * 300 * n, hipMemcpyDeviceToHost, s));
    if (elem_status)
        MSA_HIP(hipMemcpyAsync(elem_status, ds, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    MSA_HIP(hipStreamSynchronize(s));
    return MSA_OK;
}
