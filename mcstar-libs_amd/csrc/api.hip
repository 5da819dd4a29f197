// C ABI: status strings, device selection, memory and the library stream.
// Replaces the CUDA runtime calls of cuda/array_abstraction.hpp:74-131 and
// the CUERR convention of common/cu_error.hpp:5-11.
#include "mcs_common.hpp"

#include <cstring>
#include <mutex>

static thread_local char g_last_error[512] = "";
static hipStream_t g_stream = nullptr;
static int g_device = -1;
static std::mutex g_init_mu;

void mcs_set_last_error(const char* file, int line, const char* msg) {
    const char* base = strrchr(file, '/');
    snprintf(g_last_error, sizeof(g_last_error), "%s:%d %s", base ? base + 1 : file, line, msg ? msg : "");
}

extern "C" const char* mcs_last_error(void) { return g_last_error; }

extern "C" const char* mcs_error_string(int code) {
    switch (code) {
        case MCS_OK: return "MCS_OK";
        case MCS_ERR_INVALID: return "MCS_ERR_INVALID";
        case MCS_ERR_HIP: return "MCS_ERR_HIP";
        case MCS_ERR_UNSUPPORTED: return "MCS_ERR_UNSUPPORTED";
        case MCS_ERR_NOMEM: return "MCS_ERR_NOMEM";
        case MCS_ERR_CAPACITY: return "MCS_ERR_CAPACITY";
        default: return "<unknown>";
    }
}

static int ensure_init(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_stream && (device < 0 || device == g_device)) return MCS_OK;
    int n = 0;
    MCS_HIP_TRY(hipGetDeviceCount(&n));
    MCS_CHECK(n > 0, MCS_ERR_HIP, "no HIP device");
    if (device < 0) device = 0;
    MCS_CHECK(device < n, MCS_ERR_INVALID, "device index out of range");
    MCS_HIP_TRY(hipSetDevice(device));
    if (g_stream) (void)hipStreamDestroy(g_stream);
    // default priority, non-blocking with respect to the null stream
    MCS_HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = device;
    return MCS_OK;
}

hipStream_t mcs_stream() {
    if (!g_stream) (void)ensure_init(-1);
    return g_stream;
}

extern "C" int mcs_init(int device) { return ensure_init(device); }

extern "C" int mcs_sync(void) {
    int st = ensure_init(-1);
    if (st) return st;
    MCS_HIP_TRY(hipStreamSynchronize(g_stream));
    return MCS_OK;
}

extern "C" int mcs_malloc(void** ptr, size_t bytes) {
    MCS_CHECK(ptr, MCS_ERR_INVALID, "null argument");
    int st = ensure_init(-1);
    if (st) return st;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        *ptr = nullptr;
        mcs_set_last_error(__FILE__, __LINE__, "device allocation failed");
        return MCS_ERR_NOMEM;
    }
    return MCS_OK;
}

extern "C" int mcs_free(void* ptr) {
    if (!ptr) return MCS_OK;
    MCS_HIP_TRY(hipFree(ptr));
    return MCS_OK;
}

extern "C" int mcs_host_alloc(void** ptr, size_t bytes) {
    MCS_CHECK(ptr, MCS_ERR_INVALID, "null argument");
    int st = ensure_init(-1);
    if (st) return st;
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        mcs_set_last_error(__FILE__, __LINE__, "pinned host allocation failed");
        return MCS_ERR_NOMEM;
    }
    memset(*ptr, 0, bytes);
    return MCS_OK;
}

extern "C" int mcs_host_free(void* ptr) {
    if (!ptr) return MCS_OK;
    MCS_HIP_TRY(hipHostFree(ptr));
    return MCS_OK;
}

extern "C" int mcs_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    MCS_CHECK(dst && src, MCS_ERR_INVALID, "null argument");
    if (kind == MCS_H2H) { memcpy(dst, src, bytes); return MCS_OK; }
    int st = ensure_init(-1);
    if (st) return st;
    hipMemcpyKind k;
    switch (kind) {
        case MCS_H2D: k = hipMemcpyHostToDevice; break;
        case MCS_D2H: k = hipMemcpyDeviceToHost; break;
        case MCS_D2D: k = hipMemcpyDeviceToDevice; break;
        default: mcs_set_last_error(__FILE__, __LINE__, "bad copy kind"); return MCS_ERR_INVALID;
    }
    MCS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, g_stream));
    MCS_HIP_TRY(hipStreamSynchronize(g_stream));
    return MCS_OK;
}

extern "C" int mcs_memset(void* dst, int value, size_t bytes) {
    MCS_CHECK(dst, MCS_ERR_INVALID, "null argument");
    int st = ensure_init(-1);
    if (st) return st;
    MCS_HIP_TRY(hipMemsetAsync(dst, value, bytes, g_stream));
    MCS_HIP_TRY(hipStreamSynchronize(g_stream));
    return MCS_OK;
}

static void* g_ws_ptr[MCS_WS_SLOTS] = {};
static size_t g_ws_size[MCS_WS_SLOTS] = {};
static std::mutex g_ws_mu;

int mcs_workspace(int slot, size_t bytes, void** out) {
    MCS_CHECK(slot >= 0 && slot < MCS_WS_SLOTS && out, MCS_ERR_INVALID, "bad workspace slot");
    int st = ensure_init(-1);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (g_ws_size[slot] < bytes) {
        if (g_ws_ptr[slot]) {
            MCS_HIP_TRY(hipStreamSynchronize(g_stream));  // kernels of an earlier call may still read it
            (void)hipFree(g_ws_ptr[slot]);
            g_ws_ptr[slot] = nullptr;
            g_ws_size[slot] = 0;
        }
        // grow by at least a quarter so that slowly rising sizes do not reallocate every call
        const size_t want = bytes + bytes / 4;
        if (hipMalloc(&g_ws_ptr[slot], want) != hipSuccess) {
            g_ws_ptr[slot] = nullptr;
            mcs_set_last_error(__FILE__, __LINE__, "workspace allocation failed");
            return MCS_ERR_NOMEM;
        }
        g_ws_size[slot] = want;
    }
    *out = g_ws_ptr[slot];
    return MCS_OK;
}

static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

extern "C" int mcs_timer_start(void) {
    int st = ensure_init(-1);
    if (st) return st;
    if (!g_t0) {
        MCS_HIP_TRY(hipEventCreate(&g_t0));
        MCS_HIP_TRY(hipEventCreate(&g_t1));
    }
    MCS_HIP_TRY(hipEventRecord(g_t0, g_stream));
    return MCS_OK;
}

extern "C" int mcs_timer_stop(double* ms) {
    MCS_CHECK(ms && g_t0, MCS_ERR_INVALID, "timer not started");
    MCS_HIP_TRY(hipEventRecord(g_t1, g_stream));
    MCS_HIP_TRY(hipEventSynchronize(g_t1));
    float f = 0;
    MCS_HIP_TRY(hipEventElapsedTime(&f, g_t0, g_t1));
    *ms = f;
    return MCS_OK;
}
