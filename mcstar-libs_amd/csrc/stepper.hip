// Fused PFC time step (the hot path).
//
// One step of the reference's loop (cuda/kernels/psi3_split.hpp:15-26,
// cuda/kernels/eom.hpp:60-86, cuda/noise.hpp:110-142, cufft_wrapper.hpp:81-85)
// is c2r -> psi^3 -> r2c -> noise -> update.  Here the real field psi(x,y,z)
// is packed as u(m,y,z) = psi(2m,y,z) + i psi(2m+1,y,z) and the state is
// U = FFT3(u) over the (nx/2) x ny x nz grid: N/2 complex values, rows of nx/2
// aligned elements, no Hermitian redundancy.  The full half spectrum follows
// from a pair of packed modes:
//     E = (U(p,y,z) + conj U(-p,-y,-z)) / 2,  O = -i (U(p,..) - conj U(-p,..)) / 2
//     Psi(p, y, z) = E + w^p O,  Psi(p + nx/2, y, z) = E - w^p O,  w = exp(-2 pi i / nx)
// so the semi-implicit update runs on row pairs (y,z) ~ (-y,-z) in the x pass.
//
// Step (3D), each a single sweep over HBM except the first:
//   S3  x rows:  FFT_x(W) -> unpack, update with Philox noise, repack -> U;  IFFT_x -> W
//   S2  y cols:  IFFT_y(W)
//   S1  z cols:  IFFT_z(W), nonlinearity (psi^3 or psi^3 - psi), FFT_z
//   S2  y cols:  FFT_y(W)
// 10 sweeps of N/2 complex values per step (S3 reads W and U, writes U and W).
// 2D: S3 on x rows, S1 on y columns (6 sweeps).
#include "fixed.hpp"
#include "pass_bodies.hpp"
#include "stepper.hpp"

#include <atomic>
#include <cmath>
#include <vector>

// ------------------------------------------------------------------ kernels (one work unit per workgroup;
// every launch uses blockDim.x == G * S, the invariant the bodies rely on)
template <typename F, int P, int DIR>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_row_reg(const C2<F>* src, C2<F>* dst, long nrows, int G, FftPlanDev pl,
                                                  const C2<F>* __restrict__ tw) {
    extern __shared__ unsigned char smem_raw[];
    row_body<F, P, DIR>(src, dst, nrows, G, pl, tw, reinterpret_cast<C2<F>*>(smem_raw), blockIdx.x, threadIdx.x);
}

// swz: the tile order of xcd_tile (col_tile_order)
template <typename F, int P, int SPLIT>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_col_nl(const C2<F>* src, C2<F>* W, ColGeom g, int G,
                                                 FftPlanDev pl, const C2<F>* __restrict__ tw, F scale, int64_t* d_step,
                                                 int swz, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int tilesx = (g.nxc + G - 1) / G;
    const unsigned blk = xcd_tile(blockIdx.x, gridDim.x, swz, tilesx, g.nouter);
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * pl.L);  // behind the exchange area
    col_nl_body<F, P, SPLIT>(src, W, g, G, pl, twl, scale, lds, blk % tilesx, blk / tilesx, threadIdx.x);
    if (d_step && blockIdx.x == 0 && threadIdx.x == 0) *d_step += 1;
}

template <typename F, int P, int SPLIT, bool WAVE, bool NOISE>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS) k_xupd(C2<F>* __restrict__ W, C2<F>* __restrict__ U, const int2* __restrict__ pairs,
                                              int npairs, int G, XArgs<F> ar, FftPlanDev pl,
                                              const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph,
                                              const int64_t* __restrict__ d_step, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const uint32_t draw = NOISE && ar.noise ? (uint32_t)*d_step : 0u;
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * 2 * pl.L);  // behind the exchange area
    xupd_body<F, P, SPLIT, WAVE, RunPlan, NOISE>(W, U, pairs, npairs, G, ar, pl, twl, wph, draw, lds, blockIdx.x,
                                                 threadIdx.x);
}

// ------------------------------------------------------------------ helpers
template <typename F>
__global__ void __launch_bounds__(256) k_scale_copy(const F* src, F* dst, long n, F s) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i] * s;
}

template <typename F>
__global__ void __launch_bounds__(256) k_sum_f64(const F* a, long n, double* part) {
    __shared__ double s[256];
    double acc = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) acc += (double)a[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

static int set_lds_attr(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

template <typename F, int P, int DIR>
static int launch_row_reg_t(const LinePlan& lp, const void* src, void* dst, long nrows, hipStream_t s) {
    const int S = lp.L / P;
    int G = S >= 256 ? 1 : 256 / S;
    const size_t esz = sizeof(C2<F>);
    while (G > 1 && (size_t)G * lp.L * esz > 65536) --G;
    const size_t lds = (size_t)G * lp.L * esz;
    auto fn = k_row_reg<F, P, DIR>;
    int st = set_lds_attr((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3((unsigned)((nrows + G - 1) / G)), dim3(G * S), lds, s, (const C2<F>*)src, (C2<F>*)dst,
                       nrows, G, lp.dev, (const C2<F>*)lp.tw);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int DIR>
static int launch_row_reg_d(const LinePlan& lp, const void* src, void* dst, long nrows, hipStream_t s) {
    switch (lp.P) {
        case 4: return launch_row_reg_t<F, 4, DIR>(lp, src, dst, nrows, s);
        case 8: return launch_row_reg_t<F, 8, DIR>(lp, src, dst, nrows, s);
        case 6: return launch_row_reg_t<F, 6, DIR>(lp, src, dst, nrows, s);
        case 10: return launch_row_reg_t<F, 10, DIR>(lp, src, dst, nrows, s);
        case 12: return launch_row_reg_t<F, 12, DIR>(lp, src, dst, nrows, s);
        case 20: return launch_row_reg_t<F, 20, DIR>(lp, src, dst, nrows, s);
        case 16: return launch_row_reg_t<F, 16, DIR>(lp, src, dst, nrows, s);
        case 32: return launch_row_reg_t<F, 32, DIR>(lp, src, dst, nrows, s);
        case 30: return launch_row_reg_t<F, 30, DIR>(lp, src, dst, nrows, s);
        default: mcs_set_last_error(__FILE__, __LINE__, "row plan not register-eligible"); return MCS_ERR_UNSUPPORTED;
    }
}

int launch_row_packed(const LinePlan& lp, const void* src, void* dst, long nrows, int dir, hipStream_t s) {
    if (lp.prec == 64)
        return dir < 0 ? launch_row_reg_d<double, -1>(lp, src, dst, nrows, s) : launch_row_reg_d<double, +1>(lp, src, dst, nrows, s);
    return dir < 0 ? launch_row_reg_d<float, -1>(lp, src, dst, nrows, s) : launch_row_reg_d<float, +1>(lp, src, dst, nrows, s);
}

template <typename F, int P, int SPLIT>
static int launch_col_nl_t(const LinePlan& lp, const void* src, void* W, const ColGeom& g, double scale,
                           int64_t* d_step, hipStream_t s) {
    const int S = lp.L / P;
    const size_t esz = sizeof(C2<F>);
    const int G = col_lines_per_wg(lp);
    const long tiles = (long)((g.nxc + G - 1) / G) * g.nouter;
    const int twn = twiddles_in_lds(lp, (size_t)G * lp.L * esz, tiles);
    const size_t lds = ((size_t)G * lp.L + twn) * esz;  // exchange area, twiddles
    auto fn = k_col_nl<F, P, SPLIT>;
    int st = set_lds_attr((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(G * S), lds, s, (const C2<F>*)src, (C2<F>*)W, g, G, lp.dev,
                       (const C2<F>*)lp.tw, (F)scale, d_step, col_tile_order(g, G, (long)esz), twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int SPLIT>
static int launch_col_nl_d(const LinePlan& lp, const void* src, void* W, const ColGeom& g, double scale,
                           int64_t* d_step, hipStream_t s) {
    switch (lp.P) {
        case 4: return launch_col_nl_t<F, 4, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 8: return launch_col_nl_t<F, 8, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 6: return launch_col_nl_t<F, 6, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 10: return launch_col_nl_t<F, 10, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 12: return launch_col_nl_t<F, 12, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 20: return launch_col_nl_t<F, 20, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 16: return launch_col_nl_t<F, 16, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 32: return launch_col_nl_t<F, 32, SPLIT>(lp, src, W, g, scale, d_step, s);
        case 30: return launch_col_nl_t<F, 30, SPLIT>(lp, src, W, g, scale, d_step, s);
        default: mcs_set_last_error(__FILE__, __LINE__, "column plan not register-eligible"); return MCS_ERR_UNSUPPORTED;
    }
}

// a column pass of the stepper on W in place: kind -1 forward, +1 inverse, 0 nonlinear
int launch_step_col(const LinePlan& lp, void* W, const ColGeom& g, int kind, double scale, int split,
                    int64_t* d_step, hipStream_t s) {
    return launch_step_col_to(lp, W, W, g, kind, scale, split, d_step, s);
}

int launch_step_col_to(const LinePlan& lp, const void* src, void* W, const ColGeom& g, int kind, double scale,
                       int split, int64_t* d_step, hipStream_t s) {
    if (kind == 0) return launch_col_nl(lp, src, W, g, scale, split, d_step, s);
    bool done = false;
    const int st = fixed_step_col(lp, src, W, g, kind, scale, split, d_step, s, &done);
    if (st || done) return st;
    return launch_col(lp, src, W, g, kind, s);
}

int launch_col_nl(const LinePlan& lp, const void* src, void* W, const ColGeom& g, double scale, int split,
                  int64_t* d_step, hipStream_t s) {
    bool done = false;
    const int st = fixed_step_col(lp, src, W, g, 0, scale, split, d_step, s, &done);
    if (st || done) return st;
    if (lp.prec == 64)
        return split == MCS_SPLIT_REFERENCE
                   ? launch_col_nl_d<double, MCS_SPLIT_REFERENCE>(lp, src, W, g, scale, d_step, s)
                   : launch_col_nl_d<double, MCS_SPLIT_CUBIC>(lp, src, W, g, scale, d_step, s);
    return split == MCS_SPLIT_REFERENCE ? launch_col_nl_d<float, MCS_SPLIT_REFERENCE>(lp, src, W, g, scale, d_step, s)
                                        : launch_col_nl_d<float, MCS_SPLIT_CUBIC>(lp, src, W, g, scale, d_step, s);
}

template <typename F, int P, int SPLIT>
static int launch_xupd_t(const StepGeom& sg, void* W, void* U, const XArgs<F>& ar, hipStream_t s) {
    const LinePlan& lp = sg.px;
    const int S = lp.L / P;
    const size_t esz = sizeof(C2<F>);
    const int G = xupd_pairs_per_wg(lp);
    const int twn = twiddles_in_lds(lp, (size_t)G * 2 * lp.L * esz, (sg.npairs + G - 1) / G);
    const size_t lds = ((size_t)G * 2 * lp.L + twn) * esz;  // exchange area, twiddles
    // the noise-free kernels carry no noise draw
    auto fn = ar.noise ? ((64 % S == 0) ? k_xupd<F, P, SPLIT, true, true> : k_xupd<F, P, SPLIT, false, true>)
                       : ((64 % S == 0) ? k_xupd<F, P, SPLIT, true, false> : k_xupd<F, P, SPLIT, false, false>);
    int st = set_lds_attr((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3((unsigned)((sg.npairs + G - 1) / G)), dim3(G * S), lds, s, (C2<F>*)W, (C2<F>*)U,
                       sg.pairs, sg.npairs, G, ar, lp.dev, (const C2<F>*)lp.tw, (const C2<F>*)sg.wph, sg.d_step,
                       twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int SPLIT>
static int launch_xupd_d(const StepGeom& sg, void* W, void* U, const XArgs<F>& ar, hipStream_t s) {
    switch (sg.px.P) {
        case 4: return launch_xupd_t<F, 4, SPLIT>(sg, W, U, ar, s);
        case 8: return launch_xupd_t<F, 8, SPLIT>(sg, W, U, ar, s);
        case 6: return launch_xupd_t<F, 6, SPLIT>(sg, W, U, ar, s);
        case 10: return launch_xupd_t<F, 10, SPLIT>(sg, W, U, ar, s);
        case 12: return launch_xupd_t<F, 12, SPLIT>(sg, W, U, ar, s);
        case 20: return launch_xupd_t<F, 20, SPLIT>(sg, W, U, ar, s);
        case 16: return launch_xupd_t<F, 16, SPLIT>(sg, W, U, ar, s);
        case 32: return launch_xupd_t<F, 32, SPLIT>(sg, W, U, ar, s);
        case 30: return launch_xupd_t<F, 30, SPLIT>(sg, W, U, ar, s);
        default: mcs_set_last_error(__FILE__, __LINE__, "x plan not register-eligible"); return MCS_ERR_UNSUPPORTED;
    }
}

int launch_xupd(const StepGeom& sg, const StepPhys& ph, void* W, void* U, hipStream_t s) {
    bool done = false;
    const int st = fixed_xupd(sg, ph, W, U, s, &done);
    if (st || done) return st;
    if (sg.prec == 64) {
        const XArgs<double> ar = make_xargs<double>(sg, ph);
        return ph.split == MCS_SPLIT_REFERENCE ? launch_xupd_d<double, MCS_SPLIT_REFERENCE>(sg, W, U, ar, s)
                                               : launch_xupd_d<double, MCS_SPLIT_CUBIC>(sg, W, U, ar, s);
    }
    const XArgs<float> ar = make_xargs<float>(sg, ph);
    return ph.split == MCS_SPLIT_REFERENCE ? launch_xupd_d<float, MCS_SPLIT_REFERENCE>(sg, W, U, ar, s)
                                           : launch_xupd_d<float, MCS_SPLIT_CUBIC>(sg, W, U, ar, s);
}

// ------------------------------------------------------------------ geometry
int step_geom_build(StepGeom* sg, const int ext[3], int prec) {
    MCS_CHECK(ext[0] >= 8 && ext[0] % 2 == 0, MCS_ERR_UNSUPPORTED, "stepper needs an even nx >= 8");
    MCS_CHECK(ext[1] >= 4, MCS_ERR_UNSUPPORTED, "stepper needs ny >= 4");
    MCS_CHECK(ext[2] == 1 || ext[2] >= 4, MCS_ERR_UNSUPPORTED, "stepper needs nz == 1 or nz >= 4");
    sg->nx = ext[0]; sg->ny = ext[1]; sg->nz = ext[2];
    sg->L = ext[0] / 2;
    sg->prec = prec;
    sg->is3d = ext[2] > 1;
    MCS_CHECK(line_plan_reg_P(sg->L, prec) && line_plan_reg_P(sg->ny, prec) &&
                  (!sg->is3d || line_plan_reg_P(sg->nz, prec)),
              MCS_ERR_UNSUPPORTED, "stepper axis lengths must be 2^a, 3*2^a or 5*2^a (nx/2, ny, nz >= 4)");
    MCS_CHECK(sg->L <= 2048, MCS_ERR_UNSUPPORTED, "nx/2 too long for the x pass");
    int st = line_plan_build(&sg->px, sg->L, prec, true, 8);
    if (!st) st = line_plan_build(&sg->py, sg->ny, prec);
    if (!st && sg->is3d) st = line_plan_build(&sg->pz, sg->nz, prec);
    if (!st) st = make_phase_table(&sg->wph, sg->nx, sg->L, prec);
    if (st) return st;
    // canonical row pairs (r, mirror r) with r <= mirror
    std::vector<int2> pr;
    for (int iz = 0; iz < sg->nz; ++iz)
        for (int iy = 0; iy < sg->ny; ++iy) {
            const int my = iy == 0 ? 0 : sg->ny - iy, mz = iz == 0 ? 0 : sg->nz - iz;
            const int r = iz * sg->ny + iy, m = mz * sg->ny + my;
            if (r <= m) pr.push_back(make_int2(r, m));
        }
    sg->npairs = (int)pr.size();
    MCS_HIP_TRY(hipMalloc((void**)&sg->pairs, pr.size() * sizeof(int2)));
    MCS_HIP_TRY(hipMemcpy(sg->pairs, pr.data(), pr.size() * sizeof(int2), hipMemcpyHostToDevice));
    MCS_HIP_TRY(hipMalloc((void**)&sg->d_step, sizeof(int64_t)));
    MCS_HIP_TRY(hipMemset(sg->d_step, 0, sizeof(int64_t)));
    return MCS_OK;
}

void step_geom_free(StepGeom* sg) {
    line_plan_free(&sg->px);
    line_plan_free(&sg->py);
    line_plan_free(&sg->pz);
    if (sg->wph) (void)hipFree(sg->wph);
    if (sg->pairs) (void)hipFree(sg->pairs);
    if (sg->d_step) (void)hipFree(sg->d_step);
    sg->wph = nullptr; sg->pairs = nullptr; sg->d_step = nullptr;
}

// column geometries on the packed [nz][ny][L] array
ColGeom step_geom_y(const StepGeom& sg) { return ColGeom{sg.L, (long)sg.L, sg.nz, (long)sg.L * sg.ny}; }
ColGeom step_geom_z(const StepGeom& sg) { return ColGeom{sg.L, (long)sg.L * sg.ny, sg.ny, (long)sg.L}; }
static ColGeom geom_y(const StepGeom& sg) { return step_geom_y(sg); }
static ColGeom geom_z(const StepGeom& sg) { return step_geom_z(sg); }

// src holds x-real rows (IFFT_x(U) in W in the spectral form, the state S in the mixed form);
// leaves FFT_yz(N(psi)) in W
int step_nonlinear_half(const StepGeom& sg, const StepPhys& ph, const void* src, void* W, int64_t* d_step,
                        hipStream_t s) {
    const double scale = step_nl_scale(sg);
    int st;
    if (sg.is3d) {
        st = launch_step_col_to(sg.py, src, W, geom_y(sg), +1, scale, ph.split, nullptr, s);
        if (!st) st = launch_step_col(sg.pz, W, geom_z(sg), 0, scale, ph.split, d_step, s);
        if (!st) st = launch_step_col(sg.py, W, geom_y(sg), -1, scale, ph.split, nullptr, s);
    } else {
        st = launch_step_col_to(sg.py, src, W, geom_y(sg), 0, scale, ph.split, d_step, s);
    }
    return st;
}

// the rows the first column pass of a step reads: W (spectral form) or the state (mixed form)
static const void* step_rows(const StepGeom& sg, const void* W, const void* U) {
    return sg.form == MCS_FORM_MIXED ? U : W;
}

// the same passes as step_run with an event after each; returns per-pass milliseconds summed over nsteps
int step_run_timed(const StepGeom& sg, const StepPhys& ph, void* W, void* U, int nsteps, hipStream_t s,
                   double* acc, int* npass) {
    const int np = sg.is3d ? 4 : 2;
    *npass = np;
    hipEvent_t ev[5];
    for (int i = 0; i <= np; ++i) MCS_HIP_TRY(hipEventCreate(&ev[i]));
    for (int i = 0; i < np; ++i) acc[i] = 0;
    const double scale = step_nl_scale(sg);
    const void* rows = step_rows(sg, W, U);
    int st = MCS_OK;
    for (int it = 0; it < nsteps && !st; ++it) {
        (void)hipEventRecord(ev[0], s);
        st = launch_xpass(sg, ph, W, U, s);
        (void)hipEventRecord(ev[1], s);
        if (sg.is3d) {
            if (!st) st = launch_step_col_to(sg.py, rows, W, geom_y(sg), +1, scale, ph.split, nullptr, s);
            (void)hipEventRecord(ev[2], s);
            if (!st) st = launch_step_col(sg.pz, W, geom_z(sg), 0, scale, ph.split, sg.d_step, s);
            (void)hipEventRecord(ev[3], s);
            if (!st) st = launch_step_col(sg.py, W, geom_y(sg), -1, scale, ph.split, nullptr, s);
            (void)hipEventRecord(ev[4], s);
        } else {
            if (!st) st = launch_step_col_to(sg.py, rows, W, geom_y(sg), 0, scale, ph.split, sg.d_step, s);
            (void)hipEventRecord(ev[2], s);
        }
        if (!st && hipEventSynchronize(ev[np]) != hipSuccess) st = MCS_ERR_HIP;
        for (int i = 0; i < np && !st; ++i) {
            float f = 0;
            if (hipEventElapsedTime(&f, ev[i], ev[i + 1]) != hipSuccess) st = MCS_ERR_HIP;
            acc[i] += f;
        }
    }
    for (int i = 0; i <= np; ++i) (void)hipEventDestroy(ev[i]);
    return st;
}

int step_run(const StepGeom& sg, const StepPhys& ph, void* W, void* U, int nsteps, hipStream_t s) {
    const void* rows = step_rows(sg, W, U);
    for (int i = 0; i < nsteps; ++i) {
        int st = launch_xpass(sg, ph, W, U, s);
        if (!st) st = step_nonlinear_half(sg, ph, rows, W, sg.d_step, s);
        if (st) return st;
    }
    return MCS_OK;
}

// psi (real, reference layout, N values) -> the state U; W primed with the nonlinear term;
// returns sum(psi) in *mass.  Spectral form: U = FFT_yz(FFT_x(u)); mixed form: U = FFT_yz(u)
// (u is psi read as N/2 complex values)
int step_import(const StepGeom& sg, const StepPhys& ph, const void* psi, void* W, void* U, double* mass,
                hipStream_t s) {
    const long rows = (long)sg.ny * sg.nz;
    const long n = (long)sg.nx * sg.ny * sg.nz;
    int st = MCS_OK;
    const void* src = psi;
    if (sg.form != MCS_FORM_MIXED) {
        st = launch_row_packed(sg.px, psi, U, rows, -1, s);
        src = U;
    }
    if (!st && sg.is3d) {
        st = launch_col(sg.pz, src, U, geom_z(sg), -1, s);
        src = U;
    }
    if (!st) st = launch_col(sg.py, src, U, geom_y(sg), -1, s);
    if (st) return st;
    double* part = nullptr;
    MCS_HIP_TRY(hipMalloc((void**)&part, 1024 * sizeof(double)));
    if (sg.prec == 64) hipLaunchKernelGGL(k_sum_f64<double>, dim3(1024), dim3(256), 0, s, (const double*)psi, n, part);
    else hipLaunchKernelGGL(k_sum_f64<float>, dim3(1024), dim3(256), 0, s, (const float*)psi, n, part);
    std::vector<double> hp(1024);
    st = hipMemcpyAsync(hp.data(), part, 1024 * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess
             ? MCS_OK
             : MCS_ERR_HIP;
    (void)hipFree(part);
    if (st) { mcs_set_last_error(__FILE__, __LINE__, "mass reduction failed"); return st; }
    double m = 0;
    for (double v : hp) m += v;
    *mass = m;
    return step_prime(sg, ph, W, U, s);
}

// W as the last pass of a step leaves it (FFT_yz(N(psi)) in x-real layout), from the state U
int step_prime(const StepGeom& sg, const StepPhys& ph, void* W, const void* U, hipStream_t s) {
    if (sg.form == MCS_FORM_MIXED) return step_nonlinear_half(sg, ph, U, W, nullptr, s);
    int st = launch_row_packed(sg.px, U, W, (long)sg.ny * sg.nz, +1, s);
    if (!st) st = step_nonlinear_half(sg, ph, W, W, nullptr, s);
    return st;
}

// the state U -> psi (real, reference layout); tmp must hold N/2 complex values
int step_export(const StepGeom& sg, const void* U, void* psi, void* tmp, hipStream_t s) {
    const long rows = (long)sg.ny * sg.nz;
    const long n = (long)sg.nx * sg.ny * sg.nz;
    int st = MCS_OK;
    const void* src = U;
    if (sg.form != MCS_FORM_MIXED) {
        st = launch_row_packed(sg.px, U, tmp, rows, +1, s);
        src = tmp;
    }
    if (!st) st = launch_col(sg.py, src, tmp, geom_y(sg), +1, s);
    if (!st && sg.is3d) st = launch_col(sg.pz, tmp, tmp, geom_z(sg), +1, s);
    if (st) return st;
    // the packed inverse over N/2 points returns (N/2) u; over y and z of the mixed state, ny nz u
    const double sc = step_nl_scale(sg);
    if (sg.prec == 64)
        hipLaunchKernelGGL(k_scale_copy<double>, dim3(2048), dim3(256), 0, s, (const double*)tmp, (double*)psi, n, sc);
    else
        hipLaunchKernelGGL(k_scale_copy<float>, dim3(2048), dim3(256), 0, s, (const float*)tmp, (float*)psi, n,
                           (float)sc);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

// converts the state U in place to the other form: spectral -> mixed IFFT_x and 1/L,
// mixed -> spectral FFT_x (W, the nonlinear term, is the same in both forms)
int step_convert_form(StepGeom* sg, void* U, int form, hipStream_t s) {
    if (form == sg->form) return MCS_OK;
    const long rows = (long)sg->ny * sg->nz;
    const long n = (long)sg->nx * sg->ny * sg->nz;
    int st;
    if (form == MCS_FORM_MIXED) {
        st = launch_row_packed(sg->px, U, U, rows, +1, s);
        if (st) return st;
        const double sc = 1.0 / sg->L;
        if (sg->prec == 64)
            hipLaunchKernelGGL(k_scale_copy<double>, dim3(2048), dim3(256), 0, s, (const double*)U, (double*)U, n, sc);
        else
            hipLaunchKernelGGL(k_scale_copy<float>, dim3(2048), dim3(256), 0, s, (const float*)U, (float*)U, n,
                               (float)sc);
        MCS_HIP_TRY(hipGetLastError());
    } else {
        st = launch_row_packed(sg->px, U, U, rows, -1, s);
        if (st) return st;
    }
    sg->form = form;
    return MCS_OK;
}

// ------------------------------------------------------------------ C ABI
struct mcs_stepper {
    mcs_pfc_params p;
    StepGeom sg;
    StepPhys ph;
    void* U = nullptr;  // the state (U in the spectral form, S in the mixed form)
    void* W = nullptr;
    bool loaded = false;
    bool planned = false;  // explicit plans (not tuned)
    int64_t step = 0;
};

static std::atomic<int> g_form{MCS_FORM_SPECTRAL};

extern "C" int mcs_set_step_form(int form) {
    MCS_CHECK(form == MCS_FORM_SPECTRAL || form == MCS_FORM_MIXED, MCS_ERR_INVALID, "unknown step form");
    g_form.store(form);
    return MCS_OK;
}

extern "C" int mcs_get_step_form(int* form) {
    MCS_CHECK(form, MCS_ERR_INVALID, "null argument");
    *form = g_form.load();
    return MCS_OK;
}

static int stepper_create(mcs_stepper** out, const mcs_pfc_params* p, const int* plan) {
    MCS_CHECK(out && p, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(p->ext[0] >= 1 && p->ext[1] >= 1 && p->ext[2] >= 1, MCS_ERR_INVALID, "extent must be >=1");
    MCS_CHECK(p->prec == 32 || p->prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(p->split == MCS_SPLIT_CUBIC || p->split == MCS_SPLIT_REFERENCE, MCS_ERR_INVALID, "bad split");
    MCS_CHECK(p->dt > 0 && p->h[0] > 0 && p->h[1] > 0 && p->h[2] > 0, MCS_ERR_INVALID, "dt and h must be positive");
    if (plan) {
        const int rc = step_plan_validate(p->ext, p->prec, plan);
        if (rc) return rc;
    }
    (void)mcs_stream();
    auto* st = new mcs_stepper();
    st->p = *p;
    int rc = step_geom_build(&st->sg, p->ext, p->prec);
    const size_t bytes = (size_t)p->ext[0] * p->ext[1] * p->ext[2] * (p->prec == 64 ? 8 : 4);
    if (!rc && (hipMalloc(&st->U, bytes) != hipSuccess || hipMalloc(&st->W, bytes) != hipSuccess)) {
        mcs_set_last_error(__FILE__, __LINE__, "stepper state allocation failed");
        rc = MCS_ERR_NOMEM;
    }
    if (rc) { mcs_stepper_destroy(st); return rc; }
    st->ph.h[0] = p->h[0]; st->ph.h[1] = p->h[1]; st->ph.h[2] = p->h[2];
    st->ph.eps = p->eps; st->ph.dt = p->dt; st->ph.temp = p->temp; st->ph.seed = p->seed; st->ph.split = p->split;
    st->ph.mass = 0;
    st->sg.form = g_form.load();
    st->planned = plan != nullptr;
    rc = plan ? step_geom_apply_plan(&st->sg, plan) : step_geom_tune(&st->sg, st->ph, st->W, st->U, mcs_stream());
    if (rc) { mcs_stepper_destroy(st); return rc; }
    *out = st;
    return MCS_OK;
}

extern "C" int mcs_stepper_create(mcs_stepper** out, const mcs_pfc_params* p) { return stepper_create(out, p, nullptr); }

extern "C" int mcs_stepper_create_planned(mcs_stepper** out, const mcs_pfc_params* p, const int plan[6]) {
    MCS_CHECK(plan, MCS_ERR_INVALID, "null plan");
    return stepper_create(out, p, plan);
}

extern "C" int mcs_stepper_get_state(mcs_stepper* st, void* spec, double* mass, int64_t* step) {
    MCS_CHECK(st && spec && mass && step, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(st->loaded, MCS_ERR_INVALID, "stepper has no field (call mcs_stepper_set_psi)");
    const size_t bytes = (size_t)st->sg.nx * st->sg.ny * st->sg.nz * (st->sg.prec == 64 ? 8 : 4);
    MCS_HIP_TRY(hipMemcpyAsync(spec, st->U, bytes, hipMemcpyDefault, mcs_stream()));
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    *mass = st->ph.mass;
    *step = st->step;
    return MCS_OK;
}

extern "C" int mcs_stepper_set_state(mcs_stepper* st, const void* spec, double mass, int64_t step) {
    MCS_CHECK(st && spec && step >= 0, MCS_ERR_INVALID, "bad argument");
    hipStream_t s = mcs_stream();
    const size_t bytes = (size_t)st->sg.nx * st->sg.ny * st->sg.nz * (st->sg.prec == 64 ? 8 : 4);
    MCS_HIP_TRY(hipMemcpyAsync(st->U, spec, bytes, hipMemcpyDefault, s));
    st->ph.mass = mass;
    st->step = step;
    MCS_HIP_TRY(hipMemcpyAsync(st->sg.d_step, &st->step, sizeof(int64_t), hipMemcpyHostToDevice, s));
    // W as the last pass of a step leaves it: FFT_yz(N(psi)) in x-real layout
    int rc = step_prime(st->sg, st->ph, st->W, st->U, s);
    if (rc) return rc;
    MCS_HIP_TRY(hipStreamSynchronize(s));
    st->loaded = true;
    return MCS_OK;
}

extern "C" int mcs_stepper_plan_info(mcs_stepper* st, int info[6]) {
    MCS_CHECK(st && info, MCS_ERR_INVALID, "null argument");
    step_geom_plan_info(st->sg, info);
    return MCS_OK;
}

extern "C" int mcs_stepper_set_form(mcs_stepper* st, int form) {
    MCS_CHECK(st, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(form == MCS_FORM_SPECTRAL || form == MCS_FORM_MIXED, MCS_ERR_INVALID, "unknown step form");
    if (form == st->sg.form) return MCS_OK;
    hipStream_t s = mcs_stream();
    if (st->loaded) {
        int rc = step_convert_form(&st->sg, st->U, form, s);
        if (rc) return rc;
        MCS_HIP_TRY(hipStreamSynchronize(s));
        return MCS_OK;
    }
    st->sg.form = form;
    // no field yet: the buffers are free, so tuned plans are tuned again for this form's x pass
    if (!st->planned) {
        int rc = step_geom_tune(&st->sg, st->ph, st->W, st->U, s);
        if (rc) return rc;
    }
    return MCS_OK;
}

extern "C" int mcs_stepper_get_form(mcs_stepper* st, int* form) {
    MCS_CHECK(st && form, MCS_ERR_INVALID, "null argument");
    *form = st->sg.form;
    return MCS_OK;
}

extern "C" int mcs_stepper_destroy(mcs_stepper* st) {
    if (!st) return MCS_OK;
    step_geom_free(&st->sg);
    if (st->U) (void)hipFree(st->U);
    if (st->W) (void)hipFree(st->W);
    delete st;
    return MCS_OK;
}

extern "C" int mcs_stepper_set_psi(mcs_stepper* st, const void* psi) {
    MCS_CHECK(st && psi, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    // the draw number read by the x pass is the step counter
    MCS_HIP_TRY(hipMemcpyAsync(st->sg.d_step, &st->step, sizeof(int64_t), hipMemcpyHostToDevice, s));
    int rc = step_import(st->sg, st->ph, psi, st->W, st->U, &st->ph.mass, s);
    if (rc) return rc;
    MCS_HIP_TRY(hipStreamSynchronize(s));
    st->loaded = true;
    return MCS_OK;
}

extern "C" int mcs_stepper_get_psi(mcs_stepper* st, void* psi) {
    MCS_CHECK(st && psi, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(st->loaded, MCS_ERR_INVALID, "stepper has no field (call mcs_stepper_set_psi)");
    hipStream_t s = mcs_stream();
    const size_t bytes = (size_t)st->sg.nx * st->sg.ny * st->sg.nz * (st->p.prec == 64 ? 8 : 4);
    void* tmp = nullptr;
    MCS_HIP_TRY(hipMalloc(&tmp, bytes));
    int rc = step_export(st->sg, st->U, psi, tmp, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MCS_ERR_HIP;
    (void)hipFree(tmp);
    return rc;
}

extern "C" int mcs_stepper_step_async(mcs_stepper* st, int nsteps) {
    MCS_CHECK(st && nsteps >= 0, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(st->loaded, MCS_ERR_INVALID, "stepper has no field (call mcs_stepper_set_psi)");
    int rc = step_run(st->sg, st->ph, st->W, st->U, nsteps, mcs_stream());
    if (rc) return rc;
    st->step += nsteps;
    return MCS_OK;
}

extern "C" int mcs_stepper_step(mcs_stepper* st, int nsteps) {
    int rc = mcs_stepper_step_async(st, nsteps);
    if (rc) return rc;
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    return MCS_OK;
}

extern "C" int mcs_stepper_pass_times(mcs_stepper* st, int nsteps, double* ms, int cap, int* npass) {
    MCS_CHECK(st && ms && npass && nsteps > 0 && cap >= 4, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(st->loaded, MCS_ERR_INVALID, "stepper has no field (call mcs_stepper_set_psi)");
    double acc[4];
    int rc = step_run_timed(st->sg, st->ph, st->W, st->U, nsteps, mcs_stream(), acc, npass);
    if (rc) return rc;
    for (int i = 0; i < *npass; ++i) ms[i] = acc[i] / nsteps;
    st->step += nsteps;
    return MCS_OK;
}

extern "C" int mcs_stepper_get_step(mcs_stepper* st, int64_t* step) {
    MCS_CHECK(st && step, MCS_ERR_INVALID, "null argument");
    *step = st->step;
    return MCS_OK;
}

extern "C" int mcs_stepper_set_step(mcs_stepper* st, int64_t step) {
    MCS_CHECK(st && step >= 0, MCS_ERR_INVALID, "bad argument");
    st->step = step;
    MCS_HIP_TRY(hipMemcpyAsync(st->sg.d_step, &st->step, sizeof(int64_t), hipMemcpyHostToDevice, mcs_stream()));
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    return MCS_OK;
}

extern "C" int mcs_stepper_mass(mcs_stepper* st, double* mean) {
    MCS_CHECK(st && mean, MCS_ERR_INVALID, "null argument");
    *mean = st->ph.mass / ((double)st->sg.nx * st->sg.ny * st->sg.nz);
    return MCS_OK;
}
