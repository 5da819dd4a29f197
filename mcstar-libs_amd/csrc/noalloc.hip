// No-allocation form of the peak finder (mcs_findpeaks_noalloc): the kernels of mcs_findpeaks
// with their workspace in a file-scope device array instead of library-allocated memory, for
// callers that cannot allocate (code under HIP graph capture, tight serving loops).
#include "peaks.hpp"

#include <mutex>

// The mask, counts, offsets and the total of mcs_findpeaks_noalloc live in this file-scope device
// array: each GPU the library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: kernels get the per-device address
// hipGetSymbolAddress returns, looked up once per device.
#define MCS_PEAKS_SCRATCH_BYTES ((size_t)34 << 20)
static __device__ __align__(256) unsigned char scratch[MCS_PEAKS_SCRATCH_BYTES];

static constexpr int MCS_SCRATCH_MAX_DEVICES = 64;
static void* g_scratch_ptr[MCS_SCRATCH_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

static int scratch_address(void** out) {
    int dev = -1;
    MCS_HIP_TRY(hipGetDevice(&dev));
    MCS_CHECK(dev >= 0 && dev < MCS_SCRATCH_MAX_DEVICES, MCS_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MCS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MCS_CHECK(p, MCS_ERR_HIP, "no device address for the peak scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MCS_OK;
}

extern "C" int mcs_findpeaks_noalloc(const void* real, const int ext[3], double ths, int64_t* idx_dev,
                                     double* vals_dev, int64_t cap, int64_t* count, int prec) {
    int st = findpeaks_check(real, ext, count, cap, prec);
    if (st) return st;
    MCS_CHECK(cap == 0 || idx_dev, MCS_ERR_INVALID, "null argument");
    const size_t need = ((peaks_workspace_bytes(ext) + 255) & ~(size_t)255) + 256;
    MCS_CHECK(need <= MCS_PEAKS_SCRATCH_BYTES, MCS_ERR_UNSUPPORTED,
              "grid too large for the no-allocation peak finder (its mask must fit 34 MiB of scratch)");
    hipStream_t s = mcs_stream();
    MCS_CHECK(s, MCS_ERR_HIP, "no HIP device");
    void* ptr = nullptr;
    st = scratch_address(&ptr);
    if (st) return st;
    int64_t* d_total = reinterpret_cast<int64_t*>(reinterpret_cast<unsigned char*>(ptr) + need - 256);
    st = findpeaks_launch(real, ext, ths, idx_dev, vals_dev, cap, d_total, ptr, prec, s);
    if (st) return st;
    int64_t total = 0;
    MCS_HIP_TRY(hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipStreamSynchronize(s));
    *count = total;
    if (total > cap) {
        mcs_set_last_error(__FILE__, __LINE__, "peak capacity exceeded");
        return MCS_ERR_CAPACITY;
    }
    return MCS_OK;
}
