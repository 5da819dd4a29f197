// Local-maximum peak finder: call_kernel_findpeaks / call_kernel_findpeaks_2d
// (cuda/kernels/findpeaks.hpp:7-63, findpeaks_2d.hpp:7-56).
//
// A cell is a peak iff its value is strictly greater than each of its 26 (3D)
// or 8 (2D) periodic neighbours and not below the threshold.  The reference
// appends with an atomic counter into pinned host memory, so its order is
// arbitrary; here a count pass, a scan of per-block counts and an ordered
// write pass (wave ballots) give the list sorted by linear cell index with
// no atomics.  Reads are coalesced along x; neighbour reads hit L1/L2.
#include "peaks.hpp"

static constexpr int PBS = 256;
static constexpr int PITER = 16;  // cells per thread per block chunk
static constexpr long PCHUNK = (long)PBS * PITER;

template <typename F>
MCS_DEV bool is_peak(const F* __restrict__ v, long c, int nx, int ny, int nz, F ths) {
    const F lv = v[c];
    if (!(lv >= ths)) return false;
    const int i0 = (int)(c % nx);
    const long r = c / nx;
    const int i1 = (int)(r % ny), i2 = (int)(r / ny);
    const int lo2 = nz > 1 ? -1 : 0, hi2 = nz > 1 ? 1 : 0;
    for (int d2 = lo2; d2 <= hi2; ++d2) {
        int j2 = i2 + d2; j2 = j2 < 0 ? j2 + nz : (j2 >= nz ? j2 - nz : j2);
        for (int d1 = -1; d1 <= 1; ++d1) {
            int j1 = i1 + d1; j1 = j1 < 0 ? j1 + ny : (j1 >= ny ? j1 - ny : j1);
            const long rb = ((long)j2 * ny + j1) * nx;
            for (int d0 = -1; d0 <= 1; ++d0) {
                if (d0 == 0 && d1 == 0 && d2 == 0) continue;
                int j0 = i0 + d0; j0 = j0 < 0 ? j0 + nx : (j0 >= nx ? j0 - nx : j0);
                if (!(lv > v[rb + j0])) return false;
            }
        }
    }
    return true;
}

template <typename F>
__global__ void __launch_bounds__(PBS) k_peak_count(const F* v, int nx, int ny, int nz, F ths, long ncell, int* counts) {
    __shared__ int s[PBS];
    const long start = (long)blockIdx.x * PCHUNK;
    int cnt = 0;
    for (int j = 0; j < PITER; ++j) {
        const long c = start + (long)j * PBS + threadIdx.x;
        if (c < ncell && is_peak<F>(v, c, nx, ny, nz, ths)) ++cnt;
    }
    s[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = PBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = s[0];
}

// exclusive scan of n ints into off[0..n] (off[n] = total), one block of 1024 threads
__global__ void __launch_bounds__(1024) k_scan_counts(const int* counts, long n, int64_t* off) {
    __shared__ int64_t s[1024];
    const long per = (n + 1023) / 1024;
    const long b = (long)threadIdx.x * per;
    int64_t acc = 0;
    for (long i = b; i < b + per && i < n; ++i) acc += counts[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {  // Hillis-Steele inclusive scan
        const int64_t add = (int)threadIdx.x >= w ? s[threadIdx.x - w] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = s[threadIdx.x] - acc;
    for (long i = b; i < b + per && i < n; ++i) { off[i] = run; run += counts[i]; }
    if (threadIdx.x == 1023) off[n] = s[1023];
}

template <typename F>
__global__ void __launch_bounds__(PBS) k_peak_write(const F* v, int nx, int ny, int nz, F ths, long ncell,
                                                    const int64_t* off, int64_t* idx_out, double* val_out,
                                                    int64_t cap) {
    __shared__ int wave_cnt[PBS / 64];
    const long start = (long)blockIdx.x * PCHUNK;
    int64_t run = off[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < PITER; ++j) {
        const long c = start + (long)j * PBS + threadIdx.x;
        const bool f = c < ncell && is_peak<F>(v, c, nx, ny, nz, ths);
        const unsigned long long m = __ballot(f);
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        int64_t base = run;
        int tot = 0;
        for (int w = 0; w < PBS / 64; ++w) {
            if (w < wave) base += wave_cnt[w];
            tot += wave_cnt[w];
        }
        if (f) {
            const int64_t pos = base + __popcll(m & lt);
            if (pos < cap) {
                idx_out[pos] = c;
                if (val_out) val_out[pos] = (double)v[c];
            }
        }
        run += tot;
        __syncthreads();
    }
}

size_t peaks_workspace_bytes(const int ext[3]) {
    const long ncell = (long)ext[0] * ext[1] * ext[2];
    const long nb = (ncell + PCHUNK - 1) / PCHUNK;
    return (size_t)nb * sizeof(int) + (size_t)(nb + 1) * sizeof(int64_t) + 256;
}

int findpeaks_launch(const void* real, const int ext[3], double ths, int64_t* d_idx, double* d_vals, int64_t cap,
                     int64_t* d_total, void* workspace, int prec, hipStream_t s) {
    const long ncell = (long)ext[0] * ext[1] * ext[2];
    const long nb = (ncell + PCHUNK - 1) / PCHUNK;
    int* counts = reinterpret_cast<int*>(workspace);
    int64_t* off = reinterpret_cast<int64_t*>(reinterpret_cast<unsigned char*>(workspace) +
                                              (((size_t)nb * sizeof(int) + 255) & ~(size_t)255));
    if (prec == 64) {
        hipLaunchKernelGGL(k_peak_count<double>, dim3((unsigned)nb), dim3(PBS), 0, s, (const double*)real, ext[0],
                           ext[1], ext[2], ths, ncell, counts);
    } else {
        hipLaunchKernelGGL(k_peak_count<float>, dim3((unsigned)nb), dim3(PBS), 0, s, (const float*)real, ext[0], ext[1],
                           ext[2], (float)ths, ncell, counts);
    }
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, nb, off);
    if (prec == 64) {
        hipLaunchKernelGGL(k_peak_write<double>, dim3((unsigned)nb), dim3(PBS), 0, s, (const double*)real, ext[0],
                           ext[1], ext[2], ths, ncell, off, d_idx, d_vals, cap);
    } else {
        hipLaunchKernelGGL(k_peak_write<float>, dim3((unsigned)nb), dim3(PBS), 0, s, (const float*)real, ext[0], ext[1],
                           ext[2], (float)ths, ncell, off, d_idx, d_vals, cap);
    }
    MCS_HIP_TRY(hipGetLastError());
    if (d_total) MCS_HIP_TRY(hipMemcpyAsync(d_total, off + nb, sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    return MCS_OK;
}

extern "C" int mcs_findpeaks(const void* real, const int ext[3], double ths, int64_t* idx_out, double* vals_out,
                             int64_t cap, int64_t* count, int prec) {
    MCS_CHECK(real && ext && count && (cap == 0 || idx_out), MCS_ERR_INVALID, "null argument");
    MCS_CHECK(ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1, MCS_ERR_INVALID, "extent must be >=1 in all dimensions");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(cap >= 0, MCS_ERR_INVALID, "negative capacity");
    hipStream_t s = mcs_stream();
    const size_t wsb = peaks_workspace_bytes(ext);
    const long dcap = cap > 0 ? cap : 1;
    void* ws = nullptr;
    int64_t* d_idx = nullptr;
    double* d_vals = nullptr;
    int64_t* d_total = nullptr;
    int st = MCS_OK;
    if (hipMalloc(&ws, wsb) != hipSuccess || hipMalloc((void**)&d_idx, dcap * sizeof(int64_t)) != hipSuccess ||
        hipMalloc((void**)&d_total, sizeof(int64_t)) != hipSuccess ||
        (vals_out && hipMalloc((void**)&d_vals, dcap * sizeof(double)) != hipSuccess)) {
        mcs_set_last_error(__FILE__, __LINE__, "peak scratch allocation failed");
        st = MCS_ERR_NOMEM;
    }
    if (!st) st = findpeaks_launch(real, ext, ths, d_idx, d_vals, cap, d_total, ws, prec, s);
    int64_t total = 0;
    if (!st && hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) st = MCS_ERR_HIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
    const int64_t ncopy = total < cap ? total : cap;
    if (!st && ncopy > 0) {
        if (hipMemcpy(idx_out, d_idx, ncopy * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess) st = MCS_ERR_HIP;
        if (!st && vals_out && hipMemcpy(vals_out, d_vals, ncopy * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            st = MCS_ERR_HIP;
    }
    if (ws) (void)hipFree(ws);
    if (d_idx) (void)hipFree(d_idx);
    if (d_vals) (void)hipFree(d_vals);
    if (d_total) (void)hipFree(d_total);
    if (st) {
        if (st == MCS_ERR_HIP) mcs_set_last_error(__FILE__, __LINE__, "HIP failure in findpeaks");
        return st;
    }
    *count = total;
    if (total > cap) {
        mcs_set_last_error(__FILE__, __LINE__, "peak capacity exceeded");
        return MCS_ERR_CAPACITY;
    }
    return MCS_OK;
}
