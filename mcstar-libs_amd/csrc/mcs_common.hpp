// Shared device/host helpers for the MI355X (gfx950) PFC library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/mcstar_hip.h"

#define MCS_DEV __device__ __forceinline__
#define MCS_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- complex
template <typename F> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename F> using C2 = typename cplx_of<F>::type;

template <typename F> MCS_HD C2<F> mk(F x, F y) { C2<F> r; r.x = x; r.y = y; return r; }
// fused multiply-add written out: device code is built without implicit contraction, so
// these are the only fused operations and every kernel that shares an expression rounds it alike
MCS_HD float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
MCS_HD double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename T> MCS_HD T cadd(T a, T b) { T r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T> MCS_HD T csub(T a, T b) { T r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T> MCS_HD T cmul(T a, T b) {
    T r;
    r.x = fmad(a.x, b.x, -(a.y * b.y));
    r.y = fmad(a.x, b.y, a.y * b.x);
    return r;
}
// a * conj(b)
template <typename T> MCS_HD T cmulc(T a, T b) {
    T r;
    r.x = fmad(a.x, b.x, a.y * b.y);
    r.y = fmad(a.y, b.x, -(a.x * b.y));
    return r;
}
template <typename T> MCS_HD T cconj(T a) { T r; r.x = a.x; r.y = -a.y; return r; }
template <typename T, typename S> MCS_HD T cscale(T a, S s) { T r; r.x = a.x * s; r.y = a.y * s; return r; }
// multiply by +i (s = +1) or -i (s = -1)
template <int S, typename T> MCS_HD T cmuli(T a) {
    T r;
    if (S > 0) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
    return r;
}

// ---------------------------------------------------------------- errors
#define MCS_HIP_TRY(expr)                                                                  \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            mcs_set_last_error(__FILE__, __LINE__, hipGetErrorString(_e));                 \
            return MCS_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

#define MCS_CHECK(cond, code, msg)                                                         \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            mcs_set_last_error(__FILE__, __LINE__, msg);                                   \
            return (code);                                                                 \
        }                                                                                  \
    } while (0)

void mcs_set_last_error(const char* file, int line, const char* msg);
hipStream_t mcs_stream();  // library stream (default priority)

// Device workspaces the library keeps between calls instead of allocating per call: one
// grow-only buffer per slot on the library's device.  A buffer is reallocated only when a call
// needs more than it holds (after the stream has drained); its contents do not persist.  The
// entry points that use a slot are synchronous on the library stream, so calls from one host
// thread never overlap; calls from several host threads on the same slot must be serialised by
// the caller, as for the library stream itself.
enum { MCS_WS_PEAKS = 0, MCS_WS_KNN = 1, MCS_WS_BOND = 2, MCS_WS_SLOTS = 4 };
int mcs_workspace(int slot, size_t bytes, void** out);

static inline unsigned mcs_div_up(long a, long b) { return (unsigned)((a + b - 1) / b); }
