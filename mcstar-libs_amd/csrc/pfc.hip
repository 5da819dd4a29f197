// Pointwise PFC kernels behind the reference's kernel wrappers:
//   call_kernel_psi3_split  (cuda/kernels/psi3_split.hpp:1-26)
//   call_kernel_norm        (cuda/kernels/norm.hpp:1-10)
//   call_kernel_update_eom  (cuda/kernels/eom.hpp:60-86)
//   NoiseHostApi            (cuda/noise.hpp:93-150), on Philox4x32-10
// plus the cubic-split update with in-register noise, the seeded initial
// field and FP64 reductions (mean density, free energy).
// All are streaming kernels: one read (and write) of each operand per call.
#include "fft_kernels.hpp"
#include "philox.hpp"
#include "pfc_math.hpp"

#include <cmath>
#include <vector>

static constexpr int BS = 256;
static constexpr int RED_BLOCKS_EL = 1024;  // blocks of the relaxation update (their maxima are reduced after)

static inline unsigned grid_for(long n, int per_thread = 1) {
    long g = (n + (long)BS * per_thread - 1) / ((long)BS * per_thread);
    if (g < 1) g = 1;
    if (g > 65535L * 64) g = 65535L * 64;
    return (unsigned)g;
}

static bool ext_ok(const int ext[3]) { return ext && ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1; }

// ------------------------------------------------------------------ psi3 / norm
template <typename F>
__global__ void __launch_bounds__(BS) k_psi3_split(F* a, const F* pot, F mul, F mu, long n) {
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
        const F t = a[i] * mul;
        F r = t * t * t - t - mu;
        if (pot) r += pot[i];
        a[i] = r;
    }
}

template <typename F>
__global__ void __launch_bounds__(BS) k_norm(F* a, long n, F nf) {
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) a[i] /= nf;
}

extern "C" int mcs_psi3_split(void* real, const void* pot, double mu, int norm, const int ext[3], int prec) {
    MCS_CHECK(real && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const long n = (long)ext[0] * ext[1] * ext[2];
    hipStream_t s = mcs_stream();
    if (prec == 64) {
        const double mul = norm ? 1.0 / (double)n : 1.0;
        hipLaunchKernelGGL(k_psi3_split<double>, dim3(grid_for(n)), dim3(BS), 0, s, (double*)real, (const double*)pot,
                           mul, mu, n);
    } else {
        // psi3_split.hpp:17-20: nf = 1 / elems in Float
        float mul = 1.0f;
        if (norm) mul /= (float)n;
        hipLaunchKernelGGL(k_psi3_split<float>, dim3(grid_for(n)), dim3(BS), 0, s, (float*)real, (const float*)pot, mul,
                           (float)mu, n);
    }
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

extern "C" int mcs_norm(void* real, const int ext[3], int prec) {
    MCS_CHECK(real && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const long n = (long)ext[0] * ext[1] * ext[2];
    hipStream_t s = mcs_stream();
    if (prec == 64)
        hipLaunchKernelGGL(k_norm<double>, dim3(grid_for(n)), dim3(BS), 0, s, (double*)real, n, (double)n);
    else
        hipLaunchKernelGGL(k_norm<float>, dim3(grid_for(n)), dim3(BS), 0, s, (float*)real, n, (float)n);
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ update_eom (reference form)
// eom.hpp:60-72 with the operation order kept:
//   namp *= sqrt(N); k = sqrt(-dop.x);
//   denom = 1 - dt((2-eps) dop.x + 2 dop.y + dop.x dop.y)
//   kpsi = (kpsi + dt dop.x knl + dt namp k knoise) / denom
template <typename F>
__global__ void __launch_bounds__(BS) k_update_eom(C2<F>* kpsi, const C2<F>* knl, const C2<F>* knoise, int nx, int ny,
                                                   int nz, F namp, F dt, F eps, F lx, F ly, F lz) {
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const F k0 = kwave<F>(i0, nx, lx), k1 = kwave<F>(i1, ny, ly), k2 = kwave<F>(i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F k = sqrt(-dx);
        const F denom = F(1) - dt * ((F(2) - eps) * dx + F(2) * dy + dx * dy);
        const C2<F> p = kpsi[idx], q = knl[idx];
        const C2<F> z = knoise ? knoise[idx] : mk<F>(0, 0);
        kpsi[idx] = mk<F>((p.x + dt * dx * q.x + dt * namp * k * z.x) / denom,
                          (p.y + dt * dx * q.y + dt * namp * k * z.y) / denom);
    }
}

extern "C" int mcs_update_eom(void* kpsi, const void* knonlin, const void* knoise, const int ext[3], double namp,
                              double dt, double eps, const double lens[3], int prec) {
    MCS_CHECK(kpsi && knonlin && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const long n = (long)ext[0] * ext[1] * ext[2];
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    hipStream_t s = mcs_stream();
    if (prec == 64) {
        // eom.hpp:62 scales by sqrtf(N) in both precisions
        const double na = namp * (double)sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)kpsi,
                           (const double2*)knonlin, (const double2*)knoise, ext[0], ext[1], ext[2], na, dt, eps,
                           lens[0], lens[1], lens[2]);
    } else {
        const float na = (float)namp * sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)kpsi,
                           (const float2*)knonlin, (const float2*)knoise, ext[0], ext[1], ext[2], na, (float)dt,
                           (float)eps, (float)lens[0], (float)lens[1], (float)lens[2]);
    }
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ update with Philox noise
template <typename F>
__global__ void __launch_bounds__(BS) k_update_pfc(C2<F>* kpsi, const C2<F>* knl, int nx, int ny, int nz, F lx, F ly,
                                                   F lz, F dt, F lin0, F amp, uint32_t key0, uint32_t key1,
                                                   uint32_t draw, int noise) {
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const F k0 = kwave<F>(i0, nx, lx), k1 = kwave<F>(i1, ny, ly), k2 = kwave<F>(i2, nz, lz);
        const F ksq = k0 * k0 + k1 * k1 + k2 * k2;
        C2<F> xi = mk<F>(0, 0);
        if (noise) xi = noise_xi<F>(i0, i1, i2, nx, ny, nz, key0, key1, draw);
        kpsi[idx] = pfc_mode_update<F>(kpsi[idx], knl[idx], xi, ksq, dt, lin0, amp);
    }
}

double pfc_noise_amp(const int ext[3], const double h[3], double dt, double temp) {
    if (temp <= 0) return 0.0;
    const double n = (double)ext[0] * ext[1] * ext[2];
    const double dV = h[0] * (ext[1] > 1 ? h[1] : 1.0) * (ext[2] > 1 ? h[2] : 1.0);
    return std::sqrt(2.0 * temp * dt / dV) * std::sqrt(n);
}

extern "C" int mcs_update_pfc(void* kpsi, const void* knl, const int ext[3], const double lens[3], double dt,
                              double eps, double temp, uint64_t seed, uint32_t draw, int split, int prec) {
    MCS_CHECK(kpsi && knl && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(split == MCS_SPLIT_CUBIC || split == MCS_SPLIT_REFERENCE, MCS_ERR_INVALID, "bad split");
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    const double h[3] = {lens[0] / ext[0], lens[1] / ext[1], lens[2] / ext[2]};
    const double amp = pfc_noise_amp(ext, h, dt, temp);
    const double lin0 = split == MCS_SPLIT_REFERENCE ? 1.0 - eps : -eps;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    hipStream_t s = mcs_stream();
    if (prec == 64)
        hipLaunchKernelGGL(k_update_pfc<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)kpsi,
                           (const double2*)knl, ext[0], ext[1], ext[2], lens[0], lens[1], lens[2], dt, lin0, amp, k0,
                           k1, draw, temp > 0);
    else
        hipLaunchKernelGGL(k_update_pfc<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)kpsi, (const float2*)knl,
                           ext[0], ext[1], ext[2], (float)lens[0], (float)lens[1], (float)lens[2], (float)dt,
                           (float)lin0, (float)amp, k0, k1, draw, temp > 0);
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ initial field
// psi = psi0 + amp * (2u - 1), u = (hi27 * 2^26 + lo26) * 2^-53 from Philox(seed, cell).
// Computed in FP64 with contraction off so the values match the FP64 oracle bit for bit.
template <typename F>
__global__ void __launch_bounds__(BS) k_init_field(F* psi, long n, double psi0, double amp, uint32_t key0,
                                                   uint32_t key1) {
#pragma clang fp contract(off)
    for (long c = (long)blockIdx.x * BS + threadIdx.x; c < n; c += (long)gridDim.x * BS) {
        const u32x4 o = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)((uint64_t)c >> 32), 0u, MCS_TAG_INIT}, key0, key1);
        const double u = ((double)(o.x >> 5) * 67108864.0 + (double)(o.y >> 6)) * (1.0 / 9007199254740992.0);
        const double t = 2.0 * u - 1.0;
        const double sc = amp * t;
        psi[c] = (F)(psi0 + sc);
    }
}

extern "C" int mcs_init_field(void* real, const int ext[3], double psi0, double amp, uint64_t seed, int prec) {
    MCS_CHECK(real && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const long n = (long)ext[0] * ext[1] * ext[2];
    hipStream_t s = mcs_stream();
    if (prec == 64)
        hipLaunchKernelGGL(k_init_field<double>, dim3(grid_for(n)), dim3(BS), 0, s, (double*)real, n, psi0, amp,
                           (uint32_t)seed, (uint32_t)(seed >> 32));
    else
        hipLaunchKernelGGL(k_init_field<float>, dim3(grid_for(n)), dim3(BS), 0, s, (float*)real, n, psi0, amp,
                           (uint32_t)seed, (uint32_t)(seed >> 32));
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ rng object
struct mcs_rng {
    uint64_t seed;
    uint32_t draw;
};

extern "C" int mcs_rng_create(mcs_rng** rng, uint64_t seed) {
    MCS_CHECK(rng, MCS_ERR_INVALID, "null argument");
    *rng = new mcs_rng{seed, 0u};
    return MCS_OK;
}
extern "C" int mcs_rng_destroy(mcs_rng* rng) { delete rng; return MCS_OK; }
extern "C" int mcs_rng_set_draw(mcs_rng* rng, uint32_t draw) {
    MCS_CHECK(rng, MCS_ERR_INVALID, "null argument");
    rng->draw = draw;
    return MCS_OK;
}
extern "C" int mcs_rng_get_draw(mcs_rng* rng, uint32_t* draw) {
    MCS_CHECK(rng && draw, MCS_ERR_INVALID, "null argument");
    *draw = rng->draw;
    return MCS_OK;
}

// four N(0, sigma^2) values per Philox block (two Box-Muller pairs)
template <typename F>
__global__ void __launch_bounds__(BS) k_fill_normal(F* out, long n, F sigma, uint32_t key0, uint32_t key1,
                                                    uint32_t draw) {
    const long nb = (n + 3) / 4;
    for (long b = (long)blockIdx.x * BS + threadIdx.x; b < nb; b += (long)gridDim.x * BS) {
        const u32x4 o = philox4x32_10(u32x4{(uint32_t)b, (uint32_t)((uint64_t)b >> 32), draw, MCS_TAG_REAL}, key0, key1);
        const uint32_t w[4] = {o.x, o.y, o.z, o.w};
        F z[4];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const F r = sqrt(F(-2) * log(u01_24<F>(w[2 * p])));
            F s, c;
            mcs_sincospi(F(2) * u01_24<F>(w[2 * p + 1]), &s, &c);
            z[2 * p] = r * c;
            z[2 * p + 1] = r * s;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (4 * b + j < n) out[4 * b + j] = sigma * z[j];
    }
}

template <typename F>
__global__ void __launch_bounds__(BS) k_fill_kspace(C2<F>* out, int nx, int ny, int nz, uint32_t key0, uint32_t key1,
                                                    uint32_t draw) {
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        out[idx] = noise_xi<F>(i0, (int)(r % ny), (int)(r / ny), nx, ny, nz, key0, key1, draw);
    }
}

extern "C" int mcs_rng_fill_normal(mcs_rng* rng, void* dst, size_t n, double sigma, int prec) {
    MCS_CHECK(rng && dst, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const uint32_t k0 = (uint32_t)rng->seed, k1 = (uint32_t)(rng->seed >> 32);
    const long nb = ((long)n + 3) / 4;
    if (prec == 64)
        hipLaunchKernelGGL(k_fill_normal<double>, dim3(grid_for(nb)), dim3(BS), 0, s, (double*)dst, (long)n, sigma, k0,
                           k1, rng->draw);
    else
        hipLaunchKernelGGL(k_fill_normal<float>, dim3(grid_for(nb)), dim3(BS), 0, s, (float*)dst, (long)n,
                           (float)sigma, k0, k1, rng->draw);
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    rng->draw++;
    return MCS_OK;
}

extern "C" int mcs_rng_fill_kspace(mcs_rng* rng, void* cmpl, const int ext[3], int prec) {
    MCS_CHECK(rng && cmpl && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const uint32_t k0 = (uint32_t)rng->seed, k1 = (uint32_t)(rng->seed >> 32);
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    if (prec == 64)
        hipLaunchKernelGGL(k_fill_kspace<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)cmpl, ext[0], ext[1],
                           ext[2], k0, k1, rng->draw);
    else
        hipLaunchKernelGGL(k_fill_kspace<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)cmpl, ext[0], ext[1],
                           ext[2], k0, k1, rng->draw);
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    rng->draw++;
    return MCS_OK;
}

__global__ void __launch_bounds__(BS) k_philox_blocks(uint32_t* out, long n, uint64_t c0, uint32_t c2, uint32_t c3,
                                                      uint32_t key0, uint32_t key1) {
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
        const uint64_t c = c0 + (uint64_t)i;
        const u32x4 o = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), c2, c3}, key0, key1);
        reinterpret_cast<uint4*>(out)[i] = make_uint4(o.x, o.y, o.z, o.w);
    }
}

extern "C" int mcs_philox_blocks(uint64_t seed, uint64_t c0, uint32_t c2, uint32_t c3, size_t n, uint32_t* out_dev) {
    MCS_CHECK(out_dev, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    hipLaunchKernelGGL(k_philox_blocks, dim3(grid_for((long)n)), dim3(BS), 0, s, out_dev, (long)n, c0, c2, c3,
                       (uint32_t)seed, (uint32_t)(seed >> 32));
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ update with the free-energy operator
// eom.hpp:88-109 (kernel_update_eom_with_fen) in its own operation order:
//   lk = (2 - eps) + 2 dop.x + dop.y, denom = 1 - dt dop.x lk,
//   kpsi <- (kpsi + dt dop.x knl + dt namp k knoise) / denom, and knoise <- lk * kpsi (the old
//   kpsi): the linear operator of the free energy applied to psi, for mcs_calc_fen after a C2R
template <typename F>
__global__ void __launch_bounds__(BS) k_update_eom_fen(C2<F>* kpsi, const C2<F>* knl, C2<F>* knoise, int nx, int ny,
                                                       int nz, F namp, F dt, F eps, F lx, F ly, F lz) {
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const F k0 = kwave<F>(i0, nx, lx), k1 = kwave<F>(i1, ny, ly), k2 = kwave<F>(i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F k = sqrt(-dx);
        const F lk = (F(2) - eps) + F(2) * dx + dy;
        const F denom = F(1) - dt * dx * lk;
        const C2<F> p = kpsi[idx], q = knl[idx], z = knoise[idx];
        kpsi[idx] = mk<F>((p.x + dt * dx * q.x + dt * namp * k * z.x) / denom,
                          (p.y + dt * dx * q.y + dt * namp * k * z.y) / denom);
        knoise[idx] = mk<F>(lk * p.x, lk * p.y);
    }
}

extern "C" int mcs_update_eom_with_fen(void* kpsi, const void* knonlin, void* knoise, const int ext[3], double namp,
                                       double dt, double eps, const double lens[3], int prec) {
    MCS_CHECK(kpsi && knonlin && knoise && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const long n = (long)ext[0] * ext[1] * ext[2];
    const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
    hipStream_t s = mcs_stream();
    if (prec == 64) {
        const double na = namp * (double)sqrtf((float)n);  // eom.hpp:90 scales by sqrtf(N)
        hipLaunchKernelGGL(k_update_eom_fen<double>, dim3(grid_for(nc)), dim3(BS), 0, s, (double2*)kpsi,
                           (const double2*)knonlin, (double2*)knoise, ext[0], ext[1], ext[2], na, dt, eps, lens[0],
                           lens[1], lens[2]);
    } else {
        const float na = (float)namp * sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom_fen<float>, dim3(grid_for(nc)), dim3(BS), 0, s, (float2*)kpsi,
                           (const float2*)knonlin, (float2*)knoise, ext[0], ext[1], ext[2], na, (float)dt, (float)eps,
                           (float)lens[0], (float)lens[1], (float)lens[2]);
    }
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

// ------------------------------------------------------------------ non-conserved relaxation
// el.hpp:3-45 (kernel_update_el, call_kernel_update_el): kpsi' = (kpsi - dt knl) / (1 + dt ((2 - eps) + 2 dop.x + dop.y)) and, per block,
// the largest |kpsi' - kpsi|; the block maxima are reduced by one more block (an exact max)
template <typename F>
__global__ void __launch_bounds__(BS) k_update_el(C2<F>* kpsi, const C2<F>* knl, int nx, int ny, int nz, F dt, F eps,
                                                  F lx, F ly, F lz, double* part) {
    __shared__ double m[BS];
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    double mx = 0;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const F k0 = kwave<F>(i0, nx, lx), k1 = kwave<F>(i1, ny, ly), k2 = kwave<F>(i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F denom = F(1) + dt * ((F(2) - eps) + F(2) * dx + dy);
        const C2<F> p = kpsi[idx], q = knl[idx];
        const C2<F> pn = mk<F>((p.x - dt * q.x) / denom, (p.y - dt * q.y) / denom);
        kpsi[idx] = pn;
        const F ex = p.x - pn.x, ey = p.y - pn.y;
        const F d = sqrt(ex * ex + ey * ey);
        mx = (double)d > mx ? (double)d : mx;
    }
    m[threadIdx.x] = mx;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) m[threadIdx.x] = m[threadIdx.x] > m[threadIdx.x + w] ? m[threadIdx.x] : m[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

__global__ void __launch_bounds__(BS) k_max_partials(const double* part, int n, double* out) {
    __shared__ double m[BS];
    double mx = 0;
    for (int i = threadIdx.x; i < n; i += BS) mx = part[i] > mx ? part[i] : mx;
    m[threadIdx.x] = mx;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) m[threadIdx.x] = m[threadIdx.x] > m[threadIdx.x + w] ? m[threadIdx.x] : m[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = m[0];
}

extern "C" int mcs_update_el(void* kpsi, const void* knonlin, const int ext[3], double dt, double eps,
                             const double lens[3], int prec, double* maxdiff) {
    MCS_CHECK(kpsi && knonlin && lens && maxdiff && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    double* part = nullptr;
    MCS_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * (RED_BLOCKS_EL + 1)));
    if (prec == 64)
        hipLaunchKernelGGL(k_update_el<double>, dim3(RED_BLOCKS_EL), dim3(BS), 0, s, (double2*)kpsi,
                           (const double2*)knonlin, ext[0], ext[1], ext[2], dt, eps, lens[0], lens[1], lens[2], part);
    else
        hipLaunchKernelGGL(k_update_el<float>, dim3(RED_BLOCKS_EL), dim3(BS), 0, s, (float2*)kpsi,
                           (const float2*)knonlin, ext[0], ext[1], ext[2], (float)dt, (float)eps, (float)lens[0],
                           (float)lens[1], (float)lens[2], part);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(BS), 0, s, part, RED_BLOCKS_EL, part + RED_BLOCKS_EL);
    int st = hipGetLastError() == hipSuccess ? MCS_OK : MCS_ERR_HIP;
    if (!st && hipMemcpyAsync(maxdiff, part + RED_BLOCKS_EL, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        st = MCS_ERR_HIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
    (void)hipFree(part);
    if (st) mcs_set_last_error(__FILE__, __LINE__, "HIP failure in mcs_update_el");
    return st;
}

// ------------------------------------------------------------------ FP64 reductions
// Deterministic: a fixed grid of per-block LDS tree sums, then one block sums the partials in order.
static constexpr int RED_BLOCKS = 1024;

template <typename F>
__global__ void __launch_bounds__(BS) k_sum_pow(const F* a, long n, double* part) {
    __shared__ double s1[BS], s4[BS];
    double v1 = 0, v4 = 0;
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
        const double v = (double)a[i];
        v1 += v;
        v4 += v * v * v * v;
    }
    s1[threadIdx.x] = v1;
    s4[threadIdx.x] = v4;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s4[threadIdx.x] += s4[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1[0]; part[2 * blockIdx.x + 1] = s4[0]; }
}

// sum over the half spectrum of w_k L(k) |psi_k|^2, w_k = 1 on the self-conjugate x planes, else 2
template <typename F>
__global__ void __launch_bounds__(BS) k_sum_lk(const C2<F>* kpsi, int nx, int ny, int nz, double lx, double ly,
                                               double lz, double eps, double* part) {
    __shared__ double s[BS];
    const int nxc = nx / 2 + 1;
    const long total = (long)nxc * ny * nz;
    double acc = 0;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const int i0 = (int)(idx % nxc);
        const long r = idx / nxc;
        const int i1 = (int)(r % ny), i2 = (int)(r / ny);
        const double k0 = kwave<double>(i0, nx, lx), k1 = kwave<double>(i1, ny, ly), k2 = kwave<double>(i2, nz, lz);
        const double ksq = k0 * k0 + k1 * k1 + k2 * k2;
        const double om = 1.0 - ksq;
        const double w = (i0 == 0 || ((nx & 1) == 0 && i0 == nx / 2)) ? 1.0 : 2.0;
        const C2<F> p = kpsi[idx];
        acc += w * (-eps + om * om) * ((double)p.x * p.x + (double)p.y * p.y);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(BS) k_sum_partials(const double* part, int n, int stride, double* out) {
    __shared__ double s[BS];
    for (int c = 0; c < stride; ++c) {
        double acc = 0;
        for (int i = threadIdx.x; i < n; i += BS) acc += part[(long)i * stride + c];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int w = BS / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = s[0];
        __syncthreads();
    }
}

extern "C" int mcs_field_stats(const void* real, const int ext[3], const double h[3], double eps, int prec,
                               double* mean, double* free_energy) {
    MCS_CHECK(real && h && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1] * ext[2];
    double* part = nullptr;
    double* res = nullptr;
    MCS_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * 2 * RED_BLOCKS));
    MCS_HIP_TRY(hipMalloc((void**)&res, sizeof(double) * 3));
    if (prec == 64)
        hipLaunchKernelGGL(k_sum_pow<double>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const double*)real, n, part);
    else
        hipLaunchKernelGGL(k_sum_pow<float>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const float*)real, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, part, RED_BLOCKS, 2, res);
    int st = MCS_OK;
    void* kbuf = nullptr;
    mcs_fft_plan* plan = nullptr;
    if (free_energy) {
        const size_t esz = prec == 64 ? 16 : 8;
        const long nc = (long)(ext[0] / 2 + 1) * ext[1] * ext[2];
        st = mcs_fft_plan_create(&plan, ext, prec, MCS_R2C);
        if (!st && hipMalloc(&kbuf, nc * esz) != hipSuccess) st = MCS_ERR_NOMEM;
        if (!st) st = fft_exec_async(plan, real, kbuf, s);
        if (!st) {
            const double lx = ext[0] * h[0], ly = ext[1] * h[1], lz = ext[2] * h[2];
            if (prec == 64)
                hipLaunchKernelGGL(k_sum_lk<double>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const double2*)kbuf, ext[0],
                                   ext[1], ext[2], lx, ly, lz, eps, part);
            else
                hipLaunchKernelGGL(k_sum_lk<float>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const float2*)kbuf, ext[0],
                                   ext[1], ext[2], lx, ly, lz, eps, part);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, part, RED_BLOCKS, 1, res + 2);
        }
    }
    double hres[3] = {0, 0, 0};
    if (!st && hipMemcpyAsync(hres, res, sizeof(hres), hipMemcpyDeviceToHost, s) != hipSuccess) st = MCS_ERR_HIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
    if (plan) mcs_fft_plan_destroy(plan);
    if (kbuf) (void)hipFree(kbuf);
    (void)hipFree(part);
    (void)hipFree(res);
    if (st) return st;
    if (mean) *mean = hres[0] / (double)n;
    if (free_energy) {
        const double dV = h[0] * (ext[1] > 1 ? h[1] : 1.0) * (ext[2] > 1 ? h[2] : 1.0);
        *free_energy = dV * (0.5 * hres[2] / (double)n + 0.25 * hres[1]);
    }
    return MCS_OK;
}

// eom.hpp:111-141 (kernel_calc_fen, call_kernel_calc_fen): per cell f = 0.5 psi (lpsi - psi + 0.5 psi^3), written into
// lpsi as the reference does, and F = dV sum f.  The sum is a deterministic FP64 reduction (the
// reference accumulated FP32 atomics); f itself is formed in double and stored in the field's type.
template <typename F>
__global__ void __launch_bounds__(BS) k_calc_fen(const F* psi, F* lpsi, long n, double* part) {
    __shared__ double s[BS];
    double acc = 0;
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
        const F p = psi[i];
        const F f = (F)(0.5 * p * (lpsi[i] - p + 0.5 * p * p * p));
        lpsi[i] = f;
        acc += (double)f;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

extern "C" int mcs_calc_fen(const void* psi, void* lpsi, const int ext[3], const double h[3], int prec, double* fen) {
    MCS_CHECK(psi && lpsi && h && fen && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1] * ext[2];
    double* part = nullptr;
    MCS_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * (RED_BLOCKS + 1)));
    if (prec == 64)
        hipLaunchKernelGGL(k_calc_fen<double>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const double*)psi, (double*)lpsi, n,
                           part);
    else
        hipLaunchKernelGGL(k_calc_fen<float>, dim3(RED_BLOCKS), dim3(BS), 0, s, (const float*)psi, (float*)lpsi, n,
                           part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, part, RED_BLOCKS, 1, part + RED_BLOCKS);
    double sum = 0;
    int st = hipGetLastError() == hipSuccess ? MCS_OK : MCS_ERR_HIP;
    if (!st && hipMemcpyAsync(&sum, part + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        st = MCS_ERR_HIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
    (void)hipFree(part);
    if (st) {
        mcs_set_last_error(__FILE__, __LINE__, "HIP failure in mcs_calc_fen");
        return st;
    }
    *fen = sum * h[0] * h[1] * h[2];  // eom.hpp:137 multiplies by hh.x hh.y hh.z
    return MCS_OK;
}
