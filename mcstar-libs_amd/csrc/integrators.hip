// Alternative k-space integrators of the reference's kernel headers, each a
// pointwise pass over the half spectrum (reference layout, one thread per mode):
//
//   exponential time differencing, conserved     cuda/kernels/eom-exp.hpp:3-77
//   exponential time differencing, non-conserved cuda/kernels/el-exp.hpp:3-61
//   conserved update with 2D "hpfc" noise        cuda/kernels/eom.hpp:16-58
//   conserved update with a polynomial operator  cuda/kernels/eom.hpp:143-180
//
// With L' = (2 - eps) + 2 dop.x + dop.y (dop.x = -k^2, dop.y = k^4) the
// conserved ETD step is psi' = E psi + S (k^2 N) + S' noise with
// E = exp(-k^2 L' dt) and S = (E - 1) / (k^2 L') (the limit -dt where the
// operator vanishes), the non-conserved one psi' = E psi + S N with
// E = exp(-L' dt) and S = (E - 1) / L'.  The (E, S) table is computed once per
// (dt, eps, box) by the init calls and read by the updates, as in the reference.
#include "pfc_math.hpp"

#include <vector>

namespace {

constexpr int BS = 256;

unsigned grid_for(long n) {
    const long g = (n + BS - 1) / BS;
    return (unsigned)(g < 1 ? 1 : (g > 65535L * 16 ? 65535L * 16 : g));
}

bool ext_ok(const int ext[3]) { return ext && ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1; }

struct Mode {
    long idx;
    int i0, i1, i2;
};

// mode coordinates of half-spectrum index idx
MCS_DEV Mode mode_of(long idx, int nx, int ny) {
    const int nxc = nx / 2 + 1;
    Mode m;
    m.idx = idx;
    m.i0 = (int)(idx % nxc);
    const long r = idx / nxc;
    m.i1 = (int)(r % ny);
    m.i2 = (int)(r / ny);
    return m;
}

// the reference takes expf in both builds (eom-exp.hpp:7, el-exp.hpp:6), so the FP64 tables
// carry single-precision exponentials too
MCS_DEV float exp_f(float x) { return expf(x); }
MCS_DEV double exp_f(double x) { return (double)expf((float)x); }

// ------------------------------------------------------------------ ETD operator tables
// eom-exp.hpp:3-10: linop = -dop.x ((2 - eps) + 2 dop.x + dop.y); (exp(-linop dt), (exp - 1) / linop),
// the second entry replaced by -dt where it is not a number (linop = 0)
template <typename F>
__global__ void __launch_bounds__(BS) k_init_eom_op(C2<F>* kop, int nx, int ny, int nz, F dt, F eps, F lx, F ly, F lz) {
    const long total = (long)(nx / 2 + 1) * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const Mode m = mode_of(idx, nx, ny);
        const F k0 = kwave<F>(m.i0, nx, lx), k1 = kwave<F>(m.i1, ny, ly), k2 = kwave<F>(m.i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F linop = -dx * ((F(2) - eps) + F(2) * dx + dy);
        const F e = exp_f(-linop * dt);
        const F second = (e - F(1)) / linop;
        kop[idx] = mk<F>(e, second == second ? second : -dt);
    }
}

// el-exp.hpp:3-9: linop = (2 - eps) + 2 dop.x + dop.y; (exp(-linop dt), (exp - 1) / linop)
template <typename F>
__global__ void __launch_bounds__(BS) k_init_el_op(C2<F>* kop, int nx, int ny, int nz, F dt, F eps, F lx, F ly, F lz) {
    const long total = (long)(nx / 2 + 1) * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const Mode m = mode_of(idx, nx, ny);
        const F k0 = kwave<F>(m.i0, nx, lx), k1 = kwave<F>(m.i1, ny, ly), k2 = kwave<F>(m.i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F linop = (F(2) - eps) + F(2) * dx + dy;
        const F e = exp_f(-linop * dt);
        kop[idx] = mk<F>(e, (e - F(1)) / linop);
    }
}

// ------------------------------------------------------------------ ETD updates
// eom-exp.hpp:18-34 (and 52-69 with fen): psi' = psi kop.x - dop.x knl kop.y - namp k knoise kop.y;
// with fen, knoise <- lk psi (the psi before the update), lk = (2 - eps) + 2 dop.x + dop.y
template <typename F, bool FEN>
__global__ void __launch_bounds__(BS) k_update_eom_etd(const C2<F>* kop, C2<F>* kpsi, const C2<F>* knl, C2<F>* knoise,
                                                       int nx, int ny, int nz, F namp, F eps, F lx, F ly, F lz) {
    const long total = (long)(nx / 2 + 1) * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const Mode m = mode_of(idx, nx, ny);
        const F k0 = kwave<F>(m.i0, nx, lx), k1 = kwave<F>(m.i1, ny, ly), k2 = kwave<F>(m.i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F dy = dx * dx;
        const F k = sqrt(-dx);
        const C2<F> p = kpsi[idx], q = knl[idx], o = kop[idx];
        const C2<F> z = knoise ? knoise[idx] : mk<F>(0, 0);
        kpsi[idx] = mk<F>(p.x * o.x - dx * q.x * o.y - namp * k * z.x * o.y,
                          p.y * o.x - dx * q.y * o.y - namp * k * z.y * o.y);
        if (FEN) {
            const F lk = (F(2) - eps) + F(2) * dx + dy;
            knoise[idx] = mk<F>(lk * p.x, lk * p.y);
        }
    }
}

// el-exp.hpp:17-34: psi' = psi kop.x + knl kop.y, and per block the largest |psi' - psi|
template <typename F>
__global__ void __launch_bounds__(BS) k_update_el_etd(const C2<F>* kop, C2<F>* kpsi, const C2<F>* knl, long total,
                                                      double* part) {
    __shared__ double mx[BS];
    double m = 0;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const C2<F> p = kpsi[idx], q = knl[idx], o = kop[idx];
        const C2<F> pn = mk<F>(p.x * o.x + q.x * o.y, p.y * o.x + q.y * o.y);
        kpsi[idx] = pn;
        const F ex = p.x - pn.x, ey = p.y - pn.y;
        const double d = (double)sqrt(ex * ex + ey * ey);
        m = d > m ? d : m;
    }
    mx[threadIdx.x] = m;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) mx[threadIdx.x] = mx[threadIdx.x] > mx[threadIdx.x + w] ? mx[threadIdx.x] : mx[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = mx[0];
}

__global__ void __launch_bounds__(BS) k_max_blocks(const double* part, int n, double* out) {
    __shared__ double mx[BS];
    double m = 0;
    for (int i = threadIdx.x; i < n; i += BS) m = part[i] > m ? part[i] : m;
    mx[threadIdx.x] = m;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) mx[threadIdx.x] = mx[threadIdx.x] > mx[threadIdx.x + w] ? mx[threadIdx.x] : mx[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = mx[0];
}

// ------------------------------------------------------------------ hpfc noise (2D)
// eom.hpp:16-46: three noise arrays combined with the in-plane wave vector,
//   noise = (c1 kx^2 + c2 ky^2) xi1 + (c2 kx^2 + c1 ky^2) xi2 + 2 kx ky xi3 for |k| <= 1/4, else 0,
//   kappa = 4/3, c2 = sqrt(1 + kappa/2 + sqrt(1 + kappa)), c1 = kappa / c2 / 2
template <typename F>
__global__ void __launch_bounds__(BS) k_update_eom_hpfc(C2<F>* kpsi, const C2<F>* knl, const C2<F>* n1,
                                                        const C2<F>* n2, const C2<F>* n3, int nx, int ny, F namp,
                                                        F dt, F eps, F lx, F ly) {
    const long total = (long)(nx / 2 + 1) * ny;
    const F kappa = F(4) / F(3);
    const F c2 = sqrt(F(1) + kappa / F(2) + sqrt(F(1) + kappa));
    const F c1 = kappa / c2 / F(2);
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const Mode m = mode_of(idx, nx, ny);
        const F k0 = kwave<F>(m.i0, nx, lx), k1 = kwave<F>(m.i1, ny, ly);
        const F dx = -k0 * k0 - k1 * k1;
        const F dy = dx * dx;
        const F k = sqrt(-dx);
        const F denom = F(1) - dt * ((F(2) - eps) * dx + F(2) * dy + dx * dy);
        const F a = c1 * k0 * k0 + c2 * k1 * k1, b = c2 * k0 * k0 + c1 * k1 * k1, c = F(2) * k0 * k1;
        const C2<F> x1 = n1[idx], x2 = n2[idx], x3 = n3[idx];
        C2<F> nz = mk<F>(a * x1.x + b * x2.x + c * x3.x, a * x1.y + b * x2.y + c * x3.y);
        if (!(k <= F(0.25))) nz = mk<F>(0, 0);
        const C2<F> p = kpsi[idx], q = knl[idx];
        kpsi[idx] = mk<F>((p.x + dt * dx * q.x + dt * namp * nz.x) / denom,
                          (p.y + dt * dx * q.y + dt * namp * nz.y) / denom);
    }
}

// ------------------------------------------------------------------ polynomial operator
// eom.hpp:143-171: the linear operator 1 - eps - P(k^2) with P a degree-6 polynomial in k^2
// (Horner, highest coefficient first), denom = 1 - dt dop.x (1 - eps - P)
template <typename F>
__global__ void __launch_bounds__(BS) k_update_eom_poly(C2<F>* kpsi, const C2<F>* knl, const C2<F>* knoise, int nx,
                                                        int ny, int nz, F namp, F dt, F eps, F lx, F ly, F lz) {
    // the fitted coefficients c0, c2, ..., c12 of eom.hpp:143-149 (data of the model, kept as given)
    const F c[7] = {F(-57.94636917114257), F(164.4074088310241), F(-181.1481995196213), F(100.6942650417575),
                    F(-30.20985042718309), F(4.678065276755401), F(-0.293423255328944)};
    const long total = (long)(nx / 2 + 1) * ny * nz;
    for (long idx = (long)blockIdx.x * BS + threadIdx.x; idx < total; idx += (long)gridDim.x * BS) {
        const Mode m = mode_of(idx, nx, ny);
        const F k0 = kwave<F>(m.i0, nx, lx), k1 = kwave<F>(m.i1, ny, ly), k2 = kwave<F>(m.i2, nz, lz);
        const F dx = -k0 * k0 - k1 * k1 - k2 * k2;
        const F ksq = -dx;
        const F k = sqrt(ksq);
        F poly = c[6];
        for (int j = 5; j >= 0; --j) poly = poly * ksq + c[j];
        const F denom = F(1) - dt * dx * (F(1) - eps - poly);
        const C2<F> p = kpsi[idx], q = knl[idx];
        const C2<F> z = knoise ? knoise[idx] : mk<F>(0, 0);
        kpsi[idx] = mk<F>((p.x + dt * dx * q.x + dt * namp * k * z.x) / denom,
                          (p.y + dt * dx * q.y + dt * namp * k * z.y) / denom);
    }
}

long half_modes(const int ext[3]) { return (long)(ext[0] / 2 + 1) * ext[1] * ext[2]; }

int finish(hipStream_t s, const char* what) {
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, what);
        return MCS_ERR_HIP;
    }
    return MCS_OK;
}

}  // namespace

extern "C" int mcs_init_eom_op(void* kop, const int ext[3], double dt, double eps, const double lens[3], int prec) {
    MCS_CHECK(kop && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const unsigned g = grid_for(half_modes(ext));
    if (prec == 64)
        hipLaunchKernelGGL(k_init_eom_op<double>, dim3(g), dim3(BS), 0, s, (double2*)kop, ext[0], ext[1], ext[2], dt,
                           eps, lens[0], lens[1], lens[2]);
    else
        hipLaunchKernelGGL(k_init_eom_op<float>, dim3(g), dim3(BS), 0, s, (float2*)kop, ext[0], ext[1], ext[2],
                           (float)dt, (float)eps, (float)lens[0], (float)lens[1], (float)lens[2]);
    return finish(s, "HIP failure in mcs_init_eom_op");
}

extern "C" int mcs_init_el_op(void* kop, const int ext[3], double dt, double eps, const double lens[3], int prec) {
    MCS_CHECK(kop && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const unsigned g = grid_for(half_modes(ext));
    if (prec == 64)
        hipLaunchKernelGGL(k_init_el_op<double>, dim3(g), dim3(BS), 0, s, (double2*)kop, ext[0], ext[1], ext[2], dt,
                           eps, lens[0], lens[1], lens[2]);
    else
        hipLaunchKernelGGL(k_init_el_op<float>, dim3(g), dim3(BS), 0, s, (float2*)kop, ext[0], ext[1], ext[2],
                           (float)dt, (float)eps, (float)lens[0], (float)lens[1], (float)lens[2]);
    return finish(s, "HIP failure in mcs_init_el_op");
}

extern "C" int mcs_update_eom_etd(const void* kop, void* kpsi, const void* knonlin, void* knoise, int with_fen,
                                  const int ext[3], double namp, double eps, const double lens[3], int prec) {
    MCS_CHECK(kop && kpsi && knonlin && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(!with_fen || knoise, MCS_ERR_INVALID, "the free-energy form needs the noise array");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1] * ext[2];
    const unsigned g = grid_for(half_modes(ext));
    if (prec == 64) {
        const double na = namp * (double)sqrtf((float)n);  // eom-exp.hpp:20 scales by sqrtf(N)
        auto fn = with_fen ? k_update_eom_etd<double, true> : k_update_eom_etd<double, false>;
        hipLaunchKernelGGL(fn, dim3(g), dim3(BS), 0, s, (const double2*)kop, (double2*)kpsi, (const double2*)knonlin,
                           (double2*)knoise, ext[0], ext[1], ext[2], na, eps, lens[0], lens[1], lens[2]);
    } else {
        const float na = (float)namp * sqrtf((float)n);
        auto fn = with_fen ? k_update_eom_etd<float, true> : k_update_eom_etd<float, false>;
        hipLaunchKernelGGL(fn, dim3(g), dim3(BS), 0, s, (const float2*)kop, (float2*)kpsi, (const float2*)knonlin,
                           (float2*)knoise, ext[0], ext[1], ext[2], na, (float)eps, (float)lens[0], (float)lens[1],
                           (float)lens[2]);
    }
    return finish(s, "HIP failure in mcs_update_eom_etd");
}

extern "C" int mcs_update_el_etd(const void* kop, void* kpsi, const void* knonlin, const int ext[3], int prec,
                                 double* maxdiff) {
    MCS_CHECK(kop && kpsi && knonlin && maxdiff && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    constexpr int NB = 1024;
    hipStream_t s = mcs_stream();
    double* part = nullptr;
    MCS_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * (NB + 1)));
    const long total = half_modes(ext);
    if (prec == 64)
        hipLaunchKernelGGL(k_update_el_etd<double>, dim3(NB), dim3(BS), 0, s, (const double2*)kop, (double2*)kpsi,
                           (const double2*)knonlin, total, part);
    else
        hipLaunchKernelGGL(k_update_el_etd<float>, dim3(NB), dim3(BS), 0, s, (const float2*)kop, (float2*)kpsi,
                           (const float2*)knonlin, total, part);
    hipLaunchKernelGGL(k_max_blocks, dim3(1), dim3(BS), 0, s, part, NB, part + NB);
    int st = finish(s, "HIP failure in mcs_update_el_etd");
    if (!st && hipMemcpy(maxdiff, part + NB, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) st = MCS_ERR_HIP;
    (void)hipFree(part);
    return st;
}

extern "C" int mcs_update_eom_hpfc(void* kpsi, const void* knonlin, const void* knoise1, const void* knoise2,
                                   const void* knoise3, const int ext[3], double namp, double dt, double eps,
                                   const double lens[3], int prec) {
    MCS_CHECK(kpsi && knonlin && knoise1 && knoise2 && knoise3 && lens && ext_ok(ext), MCS_ERR_INVALID,
              "bad argument");
    MCS_CHECK(ext[2] == 1, MCS_ERR_UNSUPPORTED, "the hpfc noise update is 2D only (eom.hpp:51-54)");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1];
    const unsigned g = grid_for(half_modes(ext));
    if (prec == 64) {
        const double na = namp * (double)sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom_hpfc<double>, dim3(g), dim3(BS), 0, s, (double2*)kpsi, (const double2*)knonlin,
                           (const double2*)knoise1, (const double2*)knoise2, (const double2*)knoise3, ext[0], ext[1],
                           na, dt, eps, lens[0], lens[1]);
    } else {
        const float na = (float)namp * sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom_hpfc<float>, dim3(g), dim3(BS), 0, s, (float2*)kpsi, (const float2*)knonlin,
                           (const float2*)knoise1, (const float2*)knoise2, (const float2*)knoise3, ext[0], ext[1], na,
                           (float)dt, (float)eps, (float)lens[0], (float)lens[1]);
    }
    return finish(s, "HIP failure in mcs_update_eom_hpfc");
}

extern "C" int mcs_update_eom_poly(void* kpsi, const void* knonlin, const void* knoise, const int ext[3], double namp,
                                   double dt, double eps, const double lens[3], int prec) {
    MCS_CHECK(kpsi && knonlin && lens && ext_ok(ext), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    hipStream_t s = mcs_stream();
    const long n = (long)ext[0] * ext[1] * ext[2];
    const unsigned g = grid_for(half_modes(ext));
    if (prec == 64) {
        const double na = namp * (double)sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom_poly<double>, dim3(g), dim3(BS), 0, s, (double2*)kpsi, (const double2*)knonlin,
                           (const double2*)knoise, ext[0], ext[1], ext[2], na, dt, eps, lens[0], lens[1], lens[2]);
    } else {
        const float na = (float)namp * sqrtf((float)n);
        hipLaunchKernelGGL(k_update_eom_poly<float>, dim3(g), dim3(BS), 0, s, (float2*)kpsi, (const float2*)knonlin,
                           (const float2*)knoise, ext[0], ext[1], ext[2], na, (float)dt, (float)eps, (float)lens[0],
                           (float)lens[1], (float)lens[2]);
    }
    return finish(s, "HIP failure in mcs_update_eom_poly");
}
