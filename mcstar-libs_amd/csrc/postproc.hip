// Post-processing of the reference's analysis headers:
//
//   Gaussian filter table and spectral products    cuda/kernels/gaussian.hpp:22-68
//   squaring a real field (amplitude of a band)     cuda/kernels/lowpass.hpp:37-53
//   pair distances under the minimum image           particlepairs/main.cpp:8-53
//
// The filter table holds (e, 1 - e) per half-spectrum mode with
// e = exp(-(sx^2 kx^2 + sy^2 ky^2 + sz^2 kz^2) / 2), s_i = 1.1 * divs_per_sigma * h_i,
// so its real part is a low-pass and its imaginary part the complementary
// high-pass; kmult multiplies a spectrum by either part.
#include "pfc_math.hpp"

#include <vector>

namespace {

constexpr int BS = 256;

unsigned grid_for(long n) {
    const long g = (n + BS - 1) / BS;
    return (unsigned)(g < 1 ? 1 : (g > 65535L * 16 ? 65535L * 16 : g));
}

bool ext_ok(const int ext[3]) { return ext && ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1; }

MCS_DEV float exp_f(float x) { return expf(x); }
MCS_DEV double exp_f(double x) { return exp(x); }

template <typename F>
__global__ void __launch_bounds__(BS) k_gaussian_init(C2<F>* kg, int nx, int ny, int nz, F lx, F ly, F lz, F sx, F sy,
                                                      F sz) {
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const F a = sx * kwave<F>(i0, nx, lx), b = sy * kwave<F>(i1, ny, ly), c = sz * kwave<F>(i2, nz, lz);
        const F e = exp_f(-F(0.5) * (a * a + b * b + c * c));
        kg[idx] = mk<F>(e, F(1) - e);
    }
}

// arr1 <- arr1 * (part == 0 ? Re arr2 : Im arr2)
template <typename F>
__global__ void __launch_bounds__(BS) k_kmult(C2<F>* a1, const C2<F>* a2, long n, int part) {
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
        const C2<F> v = a1[i], w = a2[i];
        const F f = part == 0 ? w.x : w.y;
        a1[i] = mk<F>(v.x * f, v.y * f);
    }
}

template <typename F> __global__ void __launch_bounds__(BS) k_square(F* a, long n) {
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) a[i] = a[i] * a[i];
}

// minimum image of x2 - x1 in a periodic box of width w (particlepairs/main.cpp:8-15);
// contraction off so the distances equal the CPU's bit for bit
MCS_DEV double image_diff(double x2, double x1, double w) {
#pragma clang fp contract(off)
    const double dx = x2 - x1;
    if (fabs(dx) > w / 2) return dx > 0 ? dx - w : dx + w;
    return dx;
}

MCS_DEV double pair_dist(const double* xyz, int i, int j, double box) {
#pragma clang fp contract(off)
    const double dx = image_diff(xyz[3 * j], xyz[3 * i], box);
    const double dy = image_diff(xyz[3 * j + 1], xyz[3 * i + 1], box);
    const double dz = image_diff(xyz[3 * j + 2], xyz[3 * i + 2], box);
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// every ordered pair (i, j != i) in the reference's order: row i holds the P - 1 distances to j
__global__ void __launch_bounds__(BS) k_pair_dist(const double* xyz, int P, double box, double* out) {
    const long total = (long)P * (P - 1);
    for (long t = (long)blockIdx.x * BS + threadIdx.x; t < total; t += (long)gridDim.x * BS) {
        const int i = (int)(t / (P - 1));
        const int jj = (int)(t % (P - 1));
        const int j = jj < i ? jj : jj + 1;
        out[t] = pair_dist(xyz, i, j, box);
    }
}

// histogram of the distances of unordered pairs in [0, rmax) with nbins bins (per-block
// counts in LDS, then added to the global bins: integer adds, so the result is exact)
__global__ void __launch_bounds__(BS) k_pair_hist(const double* xyz, int P, double box, double rmax, int nbins,
                                                  unsigned long long* hist) {
    extern __shared__ unsigned int bins[];
    for (int b = threadIdx.x; b < nbins; b += BS) bins[b] = 0;
    __syncthreads();
    const long total = (long)P * (P - 1) / 2;
    for (long t = (long)blockIdx.x * BS + threadIdx.x; t < total; t += (long)gridDim.x * BS) {
        // t -> (i, j), i < j, rows of decreasing length
        long i = (long)((2.0 * P - 1 - sqrt((2.0 * P - 1) * (2.0 * P - 1) - 8.0 * t)) / 2);
        while (i > 0 && i * (2L * P - i - 1) / 2 > t) --i;
        while ((i + 1) * (2L * P - i - 2) / 2 <= t) ++i;
        const long j = t - i * (2L * P - i - 1) / 2 + i + 1;
        const double d = pair_dist(xyz, (int)i, (int)j, box);
        if (d < rmax) {
            const int b = (int)(d / rmax * nbins);
            atomicAdd(&bins[b < nbins ? b : nbins - 1], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += BS)
        if (bins[b]) atomicAdd(&hist[b], (unsigned long long)bins[b]);
}

int finish(hipStream_t s, const char* what) {
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, what);
        return MCS_ERR_HIP;
    }
    return MCS_OK;
}

}  // namespace

extern "C" int mcs_gaussian_init(void* kgauss, const int ext[3], const double lens[3], const double h[3],
                                 double divs_per_sigma, int prec) {
    MCS_CHECK(kgauss && lens && h && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    if (prec == 64) {
        const double sx = 1.1 * divs_per_sigma * h[0], sy = 1.1 * divs_per_sigma * h[1], sz = 1.1 * divs_per_sigma * h[2];
        hipLaunchKernelGGL(k_gaussian_init<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)kgauss, ext[0],
                           ext[1], ext[2], lens[0], lens[1], lens[2], sx, sy, sz);
    } else {
        // gaussian.hpp:39: Float(1.1) * divs_per_sigma * hh in the field's precision
        const float d = (float)divs_per_sigma;
        hipLaunchKernelGGL(k_gaussian_init<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)kgauss, ext[0], ext[1],
                           ext[2], (float)lens[0], (float)lens[1], (float)lens[2], 1.1f * d * (float)h[0],
                           1.1f * d * (float)h[1], 1.1f * d * (float)h[2]);
    }
    return finish(s, "HIP failure in mcs_gaussian_init");
}

extern "C" int mcs_kmult(void* karr1, const void* karr2, const int ext[3], int part, int prec) {
    MCS_CHECK(karr1 && karr2 && ext_ok(ext) && (part == 0 || part == 1), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    if (prec == 64)
        hipLaunchKernelGGL(k_kmult<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)karr1,
                           (const double2*)karr2, nc, part);
    else
        hipLaunchKernelGGL(k_kmult<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)karr1, (const float2*)karr2,
                           nc, part);
    return finish(s, "HIP failure in mcs_kmult");
}

extern "C" int mcs_square(void* real, const int ext[3], int prec) {
    MCS_CHECK(real && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1] * ext[2];
    if (prec == 64) hipLaunchKernelGGL(k_square<double>, dim3(grid_for(n)), dim3(BS), 0, s, (double*)real, n);
    else hipLaunchKernelGGL(k_square<float>, dim3(grid_for(n)), dim3(BS), 0, s, (float*)real, n);
    return finish(s, "HIP failure in mcs_square");
}

extern "C" int mcs_pair_distances(const double* xyz, int npts, double box, double* out, int64_t cap, int64_t* count) {
    MCS_CHECK(npts >= 0 && count && box > 0, MCS_ERR_INVALID, "bad argument");
    const int64_t total = (int64_t)npts * (npts > 0 ? npts - 1 : 0);
    *count = total;
    if (total == 0) return MCS_OK;
    MCS_CHECK(xyz && out, MCS_ERR_INVALID, "null argument");
    if (cap < total) {
        mcs_set_last_error(__FILE__, __LINE__, "output capacity below npts*(npts-1)");
        return MCS_ERR_CAPACITY;
    }
    hipStream_t s = mcs_stream();
    double *dx = nullptr, *dout = nullptr;
    int st = MCS_OK;
    if (hipMalloc((void**)&dx, 3 * (size_t)npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dout, (size_t)total * sizeof(double)) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, "pair-distance allocation failed");
        st = MCS_ERR_NOMEM;
    }
    if (!st && hipMemcpyAsync(dx, xyz, 3 * (size_t)npts * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        st = MCS_ERR_HIP;
    if (!st) {
        hipLaunchKernelGGL(k_pair_dist, dim3(grid_for(total)), dim3(BS), 0, s, dx, npts, box, dout);
        st = finish(s, "HIP failure in mcs_pair_distances");
    }
    if (!st && hipMemcpy(out, dout, (size_t)total * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        st = MCS_ERR_HIP;
    if (dx) (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    return st;
}

extern "C" int mcs_pair_histogram(const double* xyz, int npts, double box, double rmax, int nbins, int64_t* hist) {
    MCS_CHECK(npts >= 0 && hist && box > 0 && rmax > 0 && nbins > 0 && nbins <= 16384, MCS_ERR_INVALID,
              "bad argument");
    for (int b = 0; b < nbins; ++b) hist[b] = 0;
    if (npts < 2) return MCS_OK;
    MCS_CHECK(xyz, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    double* dx = nullptr;
    unsigned long long* dh = nullptr;
    int st = MCS_OK;
    if (hipMalloc((void**)&dx, 3 * (size_t)npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dh, (size_t)nbins * sizeof(unsigned long long)) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, "pair-histogram allocation failed");
        st = MCS_ERR_NOMEM;
    }
    if (!st && (hipMemcpyAsync(dx, xyz, 3 * (size_t)npts * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemsetAsync(dh, 0, (size_t)nbins * sizeof(unsigned long long), s) != hipSuccess))
        st = MCS_ERR_HIP;
    if (!st) {
        const long total = (long)npts * (npts - 1) / 2;
        hipLaunchKernelGGL(k_pair_hist, dim3(grid_for(total) < 4096 ? grid_for(total) : 4096), dim3(BS),
                           (size_t)nbins * sizeof(unsigned int), s, dx, npts, box, rmax, nbins, dh);
        st = finish(s, "HIP failure in mcs_pair_histogram");
    }
    std::vector<unsigned long long> h((size_t)nbins);
    if (!st && hipMemcpy(h.data(), dh, (size_t)nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
        st = MCS_ERR_HIP;
    if (!st)
        for (int b = 0; b < nbins; ++b) hist[b] = (int64_t)h[b];
    if (dx) (void)hipFree(dx);
    if (dh) (void)hipFree(dh);
    return st;
}
