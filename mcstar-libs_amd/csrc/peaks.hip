// Local-maximum peak finder: call_kernel_findpeaks / call_kernel_findpeaks_2d
// (cuda/kernels/findpeaks.hpp:7-63, findpeaks_2d.hpp:7-56).
//
// A cell is a peak iff its value is strictly greater than each of its 26 (3D)
// or 8 (2D) periodic neighbours and not below the threshold.  The reference
// appends with an atomic counter into pinned host memory, so its order is
// arbitrary; here the list comes out sorted by linear cell index, with no
// atomics, in one sweep over the field and three small passes over a bit mask:
//
//   k_peak_mark   workgroup = a tile of 64 x-columns by 16 rows (y), sweeping a
//                 chunk of z planes.  Each plane's (16+2) x (64+2) patch (with the
//                 periodic halo) is staged in LDS, double buffered, so a plane is
//                 read from HBM once by the tile that owns it.  The 26-neighbour
//                 maximum is separable: with A = max(v[x-1], v[x+1]) and
//                 B = max(v[x-1], v[x], v[x+1]) along a row,
//                   M8 = max(A, B(y-1), B(y+1))       (the 8 in-plane neighbours)
//                   P9 = max(M8, v)                   (the whole 3x3 of a plane)
//                   max26(z) = max(M8(z), P9(z-1), P9(z+1)),
//                 so a thread keeps M8 and P9 of two planes in registers while
//                 the sweep moves on.  One wave owns 64 consecutive x cells of a
//                 row and writes their peak flags as one 64-bit ballot word:
//                 bits[(z ny + y) * ceil(nx/64) + x/64].
//   k_pk_count, k_scan_counts, k_pk_write
//                 popcounts of 2048 words per workgroup, an exclusive scan of the
//                 counts, and the ordered write of the set bits' linear indices.
//                 The word order (row, x tile) is the linear index order.
//
// Every word of the mask is written on every call and the counts and offsets
// are overwritten, so the workspace needs no clearing.
#include "peaks.hpp"

#include <mutex>

static constexpr int PTX = 64;   // x columns per tile (one wave)
static constexpr int PTY = 16;   // rows per tile, 4 per thread
static constexpr int PBS = 256;  // threads per workgroup (4 waves)
static constexpr int PZC = 32;   // z planes per tile
static constexpr int PPW = PTX + 2, PPH = PTY + 2;  // plane patch with halo
static constexpr int PLOADS = (PPW * PPH + PBS - 1) / PBS;
static constexpr int KW = 8;                         // mask words per thread in the compaction passes
static constexpr long WPB = (long)PBS * KW;          // mask words per workgroup

MCS_DEV int wrapi(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

// maxima that propagate NaN (IEEE 754-2019 maximum), so that a NaN neighbour fails the strict
// comparison as in the reference's per-neighbour test (v > neighbour is false for a NaN
// neighbour); in FP32 a pair of them is one v_maximum3_f32 on gfx950
template <typename F> MCS_DEV F fmax2(F a, F b) { return __builtin_elementwise_maximum(a, b); }
template <typename F> MCS_DEV F fmax3(F a, F b, F c) { return fmax2(fmax2(a, b), c); }

// tile order: every XCD (workgroup b runs on XCD b % 8) takes a contiguous run of tiles, so the
// halo rows two neighbouring tiles share are fetched into one L2
MCS_DEV unsigned pk_tile(unsigned b, unsigned nb) {
    const unsigned q = nb / 8, r = nb % 8, k = b % 8;
    return k * q + (k < r ? k : r) + b / 8;
}

template <typename F>
__global__ void __launch_bounds__(PBS) k_peak_mark(const F* __restrict__ v, int nx, int ny, int nz, F ths, int tilesx,
                                                   int tilesy, int tilesz, unsigned long long* __restrict__ bits,
                                                   int wpr) {
    __shared__ F patch[2][PPH][PPW];
    const unsigned t = pk_tile(blockIdx.x, gridDim.x);
    const int tx = (int)(t % tilesx), ty = (int)((t / tilesx) % tilesy), tz = (int)(t / ((unsigned)tilesx * tilesy));
    const int x0 = tx * PTX, y0 = ty * PTY, z0 = tz * PZC;
    const int z1 = min(z0 + PZC, nz);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool is3d = nz > 1;
    const long plane = (long)nx * ny;

    // the patch slots this thread loads: fixed (row, column) offsets within a plane
    long off[PLOADS];
    int slot[PLOADS];
#pragma unroll
    for (int k = 0; k < PLOADS; ++k) {
        const int i = tid + k * PBS;
        slot[k] = i < PPW * PPH ? i : -1;
        const int r = i / PPW, c = i % PPW;
        off[k] = (long)wrapi(y0 - 1 + r, ny) * nx + wrapi(x0 - 1 + c, nx);
    }
    // the next plane in flight in registers while the current one is read from LDS (two planes in
    // flight were slower: 293 against 245 us at 512^3)
    F pre[PLOADS];
    auto load_plane = [&](F(&dst)[PLOADS], int z) {
        const long pb = (long)wrapi(z, nz) * plane;
#pragma unroll
        for (int k = 0; k < PLOADS; ++k) dst[k] = slot[k] >= 0 ? v[pb + off[k]] : F(0);
    };
    auto store_plane = [&](int buf) {
        F* p = &patch[buf][0][0];
#pragma unroll
        for (int k = 0; k < PLOADS; ++k)
            if (slot[k] >= 0) p[slot[k]] = pre[k];
    };

    // per row j (y = y0 + wave + 4j): value, in-plane 8-max and 9-max of planes z-1 and z
    F vc[4], m8c[4], p9p[4], p9c[4];
    const int c = lane + 1;
    const bool xok = x0 + lane < nx;
    // plane z0 - 1 (3D) first, then z0 .. z1 - 1, then z1 (3D): the first and last give P9 only
    const int zs = is3d ? z0 - 1 : z0, ze = is3d ? z1 : z1 - 1;
    load_plane(pre, zs);
    store_plane(0);
    __syncthreads();
    if (zs + 1 <= ze) load_plane(pre, zs + 1);
    for (int z = zs; z <= ze; ++z) {
        const int buf = (z - zs) & 1;
        F vn[4], m8n[4], p9n[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = wave + 4 * j + 1;
            const F* up = &patch[buf][r - 1][0];
            const F* mid = &patch[buf][r][0];
            const F* dn = &patch[buf][r + 1][0];
            const F bu = fmax3(up[c - 1], up[c], up[c + 1]);
            const F bd = fmax3(dn[c - 1], dn[c], dn[c + 1]);
            const F a = fmax2(mid[c - 1], mid[c + 1]);
            vn[j] = mid[c];
            m8n[j] = fmax3(a, bu, bd);
            p9n[j] = fmax2(m8n[j], vn[j]);
        }
        // the plane before z is now complete on both sides: finalise it
        if (is3d && z >= z0 + 1) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int y = y0 + wave + 4 * j;
                const F mx = fmax3(m8c[j], p9p[j], p9n[j]);
                const bool pk = xok && y < ny && vc[j] >= ths && vc[j] > mx;
                const unsigned long long m = __ballot(pk);
                if (lane == 0 && y < ny) bits[((long)(z - 1) * ny + y) * wpr + tx] = m;
            }
        }
        if (!is3d) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int y = y0 + wave + 4 * j;
                const bool pk = xok && y < ny && vn[j] >= ths && vn[j] > m8n[j];
                const unsigned long long m = __ballot(pk);
                if (lane == 0 && y < ny) bits[((long)z * ny + y) * wpr + tx] = m;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            p9p[j] = p9c[j];
            p9c[j] = p9n[j];
            vc[j] = vn[j];
            m8c[j] = m8n[j];
        }
        if (z + 1 <= ze) {
            store_plane(buf ^ 1);  // the buffer of plane z - 1, which every thread has finished reading
            __syncthreads();
            if (z + 2 <= ze) load_plane(pre, z + 2);
        }
    }
}

// popcounts of WPB words per workgroup
__global__ void __launch_bounds__(PBS) k_pk_count(const unsigned long long* __restrict__ bits, long nwords, int* counts) {
    __shared__ int s[PBS];
    const long w0 = (long)blockIdx.x * WPB + (long)threadIdx.x * KW;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < KW; ++j)
        if (w0 + j < nwords) cnt += __popcll(bits[w0 + j]);
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

// ordered write: word i of the mask is (row i / wpr, x tile i % wpr); bit b is x = 64 tile + b
template <typename F>
__global__ void __launch_bounds__(PBS) k_pk_write(const unsigned long long* __restrict__ bits, long nwords, int wpr, int nx,
                                                  const int64_t* __restrict__ off, const F* __restrict__ v,
                                                  int64_t* __restrict__ idx_out, double* __restrict__ val_out,
                                                  int64_t cap, int64_t* d_total) {
    __shared__ int s[PBS];
    const long w0 = (long)blockIdx.x * WPB + (long)threadIdx.x * KW;
    unsigned long long w[KW];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < KW; ++j) {
        w[j] = w0 + j < nwords ? bits[w0 + j] : 0ull;
        cnt += __popcll(w[j]);
    }
    s[threadIdx.x] = cnt;
    __syncthreads();
    for (int d = 1; d < PBS; d <<= 1) {  // inclusive scan of the per-thread counts
        const int add = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t pos = off[blockIdx.x] + s[threadIdx.x] - cnt;
#pragma unroll
    for (int j = 0; j < KW; ++j) {
        unsigned long long m = w[j];
        if (!m) continue;
        const long wi = w0 + j;
        const long row = wi / wpr;
        const long xb = (wi - row * wpr) * 64;
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            if (pos < cap) {
                const long c = row * nx + xb + b;
                idx_out[pos] = c;
                if (val_out) val_out[pos] = (double)v[c];
            }
            ++pos;
        }
    }
    if (d_total && blockIdx.x == 0 && threadIdx.x == 0) *d_total = off[gridDim.x];
}

// ------------------------------------------------------------------ host side
struct PeakGeom {
    int tilesx, tilesy, tilesz, wpr;
    long nwords, nblk;
    size_t bits_off, counts_off, off_off, bytes;
};

static PeakGeom peak_geom(const int ext[3]) {
    PeakGeom g;
    g.tilesx = (ext[0] + PTX - 1) / PTX;
    g.tilesy = (ext[1] + PTY - 1) / PTY;
    g.tilesz = (ext[2] + PZC - 1) / PZC;
    g.wpr = g.tilesx;
    g.nwords = (long)g.wpr * ext[1] * ext[2];
    g.nblk = (g.nwords + WPB - 1) / WPB;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    g.bits_off = 0;
    g.counts_off = al((size_t)g.nwords * sizeof(unsigned long long));
    g.off_off = g.counts_off + al((size_t)g.nblk * sizeof(int));
    g.bytes = g.off_off + al((size_t)(g.nblk + 1) * sizeof(int64_t));
    return g;
}

size_t peaks_workspace_bytes(const int ext[3]) { return peak_geom(ext).bytes; }

int findpeaks_launch(const void* real, const int ext[3], double ths, int64_t* d_idx, double* d_vals, int64_t cap,
                     int64_t* d_total, void* workspace, int prec, hipStream_t s) {
    const PeakGeom g = peak_geom(ext);
    const long tiles = (long)g.tilesx * g.tilesy * g.tilesz;
    MCS_CHECK(tiles <= 0x7fffffffL && g.nblk <= 0x7fffffffL, MCS_ERR_UNSUPPORTED, "grid too large for the peak finder");
    unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
    auto* bits = reinterpret_cast<unsigned long long*>(ws + g.bits_off);
    int* counts = reinterpret_cast<int*>(ws + g.counts_off);
    int64_t* off = reinterpret_cast<int64_t*>(ws + g.off_off);
    if (prec == 64)
        hipLaunchKernelGGL(k_peak_mark<double>, dim3((unsigned)tiles), dim3(PBS), 0, s, (const double*)real, ext[0],
                           ext[1], ext[2], ths, g.tilesx, g.tilesy, g.tilesz, bits, g.wpr);
    else
        hipLaunchKernelGGL(k_peak_mark<float>, dim3((unsigned)tiles), dim3(PBS), 0, s, (const float*)real, ext[0],
                           ext[1], ext[2], (float)ths, g.tilesx, g.tilesy, g.tilesz, bits, g.wpr);
    hipLaunchKernelGGL(k_pk_count, dim3((unsigned)g.nblk), dim3(PBS), 0, s, bits, g.nwords, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, g.nblk, off);
    if (prec == 64)
        hipLaunchKernelGGL(k_pk_write<double>, dim3((unsigned)g.nblk), dim3(PBS), 0, s, bits, g.nwords, g.wpr, ext[0],
                           off, (const double*)real, d_idx, d_vals, cap, d_total);
    else
        hipLaunchKernelGGL(k_pk_write<float>, dim3((unsigned)g.nblk), dim3(PBS), 0, s, bits, g.nwords, g.wpr, ext[0],
                           off, (const float*)real, d_idx, d_vals, cap, d_total);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

int findpeaks_check(const void* real, const int ext[3], int64_t* count, int64_t cap, int prec) {
    MCS_CHECK(real && ext && count, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1, MCS_ERR_INVALID, "extent must be >=1 in all dimensions");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(cap >= 0, MCS_ERR_INVALID, "negative capacity");
    return MCS_OK;
}

extern "C" int mcs_findpeaks(const void* real, const int ext[3], double ths, int64_t* idx_out, double* vals_out,
                             int64_t cap, int64_t* count, int prec) {
    int st = findpeaks_check(real, ext, count, cap, prec);
    if (st) return st;
    MCS_CHECK(cap == 0 || idx_out, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    // the mask workspace and the device copies of the outputs are kept by the library between calls
    const size_t wsb = peaks_workspace_bytes(ext);
    const size_t nidx = (size_t)(cap > 0 ? cap : 1);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_idx = al(wsb), o_vals = o_idx + al(nidx * sizeof(int64_t)), o_tot = o_vals + al(nidx * sizeof(double));
    void* buf = nullptr;
    st = mcs_workspace(MCS_WS_PEAKS, o_tot + 256, &buf);
    if (st) return st;
    unsigned char* b = reinterpret_cast<unsigned char*>(buf);
    int64_t* d_idx = reinterpret_cast<int64_t*>(b + o_idx);
    double* d_vals = vals_out ? reinterpret_cast<double*>(b + o_vals) : nullptr;
    int64_t* d_total = reinterpret_cast<int64_t*>(b + o_tot);
    st = findpeaks_launch(real, ext, ths, d_idx, d_vals, cap, d_total, buf, prec, s);
    if (st) return st;
    int64_t total = 0;
    MCS_HIP_TRY(hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipStreamSynchronize(s));
    const int64_t ncopy = total < cap ? total : cap;
    if (ncopy > 0) {
        MCS_HIP_TRY(hipMemcpyAsync(idx_out, d_idx, ncopy * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (vals_out) MCS_HIP_TRY(hipMemcpyAsync(vals_out, d_vals, ncopy * sizeof(double), hipMemcpyDeviceToHost, s));
        MCS_HIP_TRY(hipStreamSynchronize(s));
    }
    *count = total;
    if (total > cap) {
        mcs_set_last_error(__FILE__, __LINE__, "peak capacity exceeded");
        return MCS_ERR_CAPACITY;
    }
    return MCS_OK;
}
