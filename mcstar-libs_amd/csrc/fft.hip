// Boundary FFT: FFT::CufftPlan<F, R2C|C2R> (cuda/cufft_wrapper.hpp:61-86)
// rebuilt as the library's own Stockham passes.
//
//   R2C: x rows (real -> half spectrum) | y columns | z columns
//   C2R: z columns | y columns | x rows (half spectrum -> real)
// Column passes keep T = 16 (f32) / 8 (f64) adjacent x columns per
// workgroup, so every global access is a 128-byte row segment.  The x pass of
// an even nx packs the real row as nx/2 complex values and splits the
// spectrum with one extra twiddle, so it costs one FFT of half length.
#include "fft_kernels.hpp"
#include "fixed.hpp"
#include "row_kernels.hpp"

#include <vector>

// ------------------------------------------------------------------ kernels
// Column pass, register engine.  src/dst may alias (each workgroup reads its
// whole tile before it writes).
template <typename F, int P, int DIR>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_col_reg(const C2<F>* src, C2<F>* dst, ColGeom g, int G, FftPlanDev pl,
                                                  const C2<F>* __restrict__ tw, int swz, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int S = pl.L / P;
    const int tilesx = (g.nxc + G - 1) / G;
    const int blk = (int)xcd_tile(blockIdx.x, gridDim.x, swz, tilesx, g.nouter);
    const int tile = blk % tilesx, outer = blk / tilesx;
    const int c = threadIdx.x % G, t = threadIdx.x / G;
    const int x = tile * G + c;
    const bool valid = x < g.nxc;
    const long base = (long)outer * g.os + x;
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? src[base + (long)(t + m * S) * g.ls] : mk<F>(0, 0);
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * pl.L);  // behind the exchange area
    reg_fft<F, P, DIR>(a, pl, twl, t, lds, c, G);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) dst[base + (long)(t + m * S) * g.ls] = a[m];
    }
}

// Column pass, generic engine.
template <typename F, int DIR>
__global__ void __launch_bounds__(256) k_col_gen(const C2<F>* src, C2<F>* dst, ColGeom g, int G, int Lp,
                                                 FftPlanDev pl, const C2<F>* __restrict__ tw, int swz) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* b0 = reinterpret_cast<C2<F>*>(smem_raw);
    C2<F>* b1 = b0 + (long)G * Lp;
    const int L = pl.L;
    const int tilesx = (g.nxc + G - 1) / G;
    const int blk = (int)xcd_tile(blockIdx.x, gridDim.x, swz, tilesx, g.nouter);
    const int tile = blk % tilesx, outer = blk / tilesx;
    const long base = (long)outer * g.os + (long)tile * G;
    for (int idx = threadIdx.x; idx < G * L; idx += blockDim.x) {
        const int c = idx % G, pos = idx / G;
        const bool valid = tile * G + c < g.nxc;
        b0[c * Lp + pos] = valid ? src[base + c + (long)pos * g.ls] : mk<F>(0, 0);
    }
    __syncthreads();
    C2<F>* res = lds_fft<F, DIR>(b0, b1, pl, tw, Lp, G, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < G * L; idx += blockDim.x) {
        const int c = idx % G, pos = idx / G;
        if (tile * G + c < g.nxc) dst[base + c + (long)pos * g.ls] = res[c * Lp + pos];
    }
}

// x pass of the boundary transform (generic engine, G rows per workgroup).
//   MODE 0: R2C, even nx: packed FFT of nx/2, then X_k = E_k + w^k O_k, k in [0, nx/2]
//   MODE 1: R2C, odd nx: FFT of (x, 0), keep nx/2+1
//   MODE 2: C2R, even nx: Z_k = (X_k + X*_{L-k}) + i (X_k - X*_{L-k}) w^-k, inverse FFT of nx/2
//   MODE 3: C2R, odd nx: rebuild the Hermitian row, inverse FFT, real part
template <typename F, int MODE>
__global__ void __launch_bounds__(256) k_row_x(const void* srcv, void* dstv, long nrows, int nx, int G, int Lp,
                                               FftPlanDev pl, const C2<F>* __restrict__ tw,
                                               const C2<F>* __restrict__ wph) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* b0 = reinterpret_cast<C2<F>*>(smem_raw);
    C2<F>* b1 = b0 + (long)G * Lp;
    const int L = pl.L;
    const int nxc = nx / 2 + 1;
    const long r0 = (long)blockIdx.x * G;
    const int nr = (int)((nrows - r0) < G ? (nrows - r0) : G);
    const int tid = threadIdx.x, nthr = blockDim.x;
    constexpr int DIR = (MODE < 2) ? -1 : +1;

    if (MODE == 0) {
        const C2<F>* src = reinterpret_cast<const C2<F>*>(srcv);
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, pos = idx - c * L;
            b0[c * Lp + pos] = c < nr ? src[(r0 + c) * L + pos] : mk<F>(0, 0);
        }
    } else if (MODE == 1) {
        const F* src = reinterpret_cast<const F*>(srcv);
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, pos = idx - c * L;
            b0[c * Lp + pos] = mk<F>(c < nr ? src[(r0 + c) * nx + pos] : F(0), F(0));
        }
    } else if (MODE == 2) {
        const C2<F>* src = reinterpret_cast<const C2<F>*>(srcv);
        for (int idx = tid; idx < G * (L + 1); idx += nthr) {
            const int c = idx / (L + 1), pos = idx - c * (L + 1);
            b1[c * Lp + pos] = c < nr ? src[(r0 + c) * (L + 1) + pos] : mk<F>(0, 0);
        }
        __syncthreads();
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, k = idx - c * L;
            C2<F> xk = b1[c * Lp + k];
            C2<F> xp = b1[c * Lp + (L - k)];
            if (k == 0) { xk.y = F(0); xp.y = F(0); }
            const C2<F> s = cadd(xk, cconj(xp));
            const C2<F> d = cmulc(csub(xk, cconj(xp)), wph[k]);  // (X_k - X*_{L-k}) * conj(w^k)
            b0[c * Lp + k] = cadd(s, cmuli<+1>(d));
        }
    } else {
        const C2<F>* src = reinterpret_cast<const C2<F>*>(srcv);
        for (int idx = tid; idx < G * nxc; idx += nthr) {
            const int c = idx / nxc, pos = idx - c * nxc;
            b0[c * Lp + pos] = c < nr ? src[(r0 + c) * nxc + pos] : mk<F>(0, 0);
        }
        __syncthreads();
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, pos = idx - c * L;
            if (pos >= nxc) b0[c * Lp + pos] = cconj(b0[c * Lp + (L - pos)]);
            if (pos == 0) b0[c * Lp].y = F(0);
        }
    }
    __syncthreads();
    C2<F>* res = lds_fft<F, DIR>(b0, b1, pl, tw, Lp, G, tid, nthr);

    if (MODE == 0) {
        C2<F>* dst = reinterpret_cast<C2<F>*>(dstv);
        for (int idx = tid; idx < G * (L + 1); idx += nthr) {
            const int c = idx / (L + 1), k = idx - c * (L + 1);
            if (c >= nr) continue;
            const int kk = k == L ? 0 : k;
            const int kp = kk == 0 ? 0 : L - kk;
            const C2<F> zk = res[c * Lp + kk];
            const C2<F> zp = cconj(res[c * Lp + kp]);
            const C2<F> e = cscale(cadd(zk, zp), F(0.5));
            const C2<F> o = cscale(cmuli<-1>(csub(zk, zp)), F(0.5));
            dst[(r0 + c) * (L + 1) + k] = cadd(e, cmul(wph[k], o));
        }
    } else if (MODE == 1) {
        C2<F>* dst = reinterpret_cast<C2<F>*>(dstv);
        for (int idx = tid; idx < G * nxc; idx += nthr) {
            const int c = idx / nxc, k = idx - c * nxc;
            if (c < nr) dst[(r0 + c) * nxc + k] = res[c * Lp + k];
        }
    } else if (MODE == 2) {
        C2<F>* dst = reinterpret_cast<C2<F>*>(dstv);
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, pos = idx - c * L;
            if (c < nr) dst[(r0 + c) * L + pos] = res[c * Lp + pos];
        }
    } else {
        F* dst = reinterpret_cast<F*>(dstv);
        for (int idx = tid; idx < G * L; idx += nthr) {
            const int c = idx / L, pos = idx - c * L;
            if (c < nr) dst[(r0 + c) * nx + pos] = res[c * Lp + pos].x;
        }
    }
}

// ------------------------------------------------------------------ launchers
// Measured on 512^3 FP32: the step's y passes take 207 us in order 0 and 191 us in
// order 1; the boundary transform's z passes over 257-wide rows cost 0.79 ms per R2C in
// order 1 and 0.72 ms in order 2; the step's z pass (256-wide rows, no shared lines) is
// fastest in order 0.
int col_tile_order(const ColGeom& g, int G, long esz) {
    if (g.ls * esz <= 65536) return 1;
    return (g.nxc % G != 0 || (g.nxc * esz) % 128 != 0) ? 2 : 0;
}

static int set_lds(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

template <typename F, int P, int DIR>
static int launch_col_reg_t(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, hipStream_t s) {
    const int S = lp.L / P;
    const int esz = (int)sizeof(C2<F>);
    const int G = col_lines_per_wg(lp);
    const long tiles = (long)((g.nxc + G - 1) / G) * g.nouter;
    const int twn = twiddles_in_lds(lp, (size_t)G * lp.L * esz, tiles);
    const size_t lds = ((size_t)G * lp.L + twn) * esz;  // exchange area, twiddles
    auto fn = k_col_reg<F, P, DIR>;
    int st = set_lds((const void*)fn, lds);
    if (st) return st;
    const int swz = col_tile_order(g, G, esz);
    hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(G * S), lds, s, (const C2<F>*)src, (C2<F>*)dst, g, G, lp.dev,
                       (const C2<F>*)lp.tw, swz, twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int DIR>
static int launch_col_gen_t(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, hipStream_t s) {
    const int Lp = line_plan_gen_Lp(lp);
    const int esz = (int)sizeof(C2<F>);
    int G = (int)(128 / esz);
    while (G > 1 && line_plan_gen_lds(lp, G) > 65536) --G;
    const size_t lds = line_plan_gen_lds(lp, G);
    auto fn = k_col_gen<F, DIR>;
    int st = set_lds((const void*)fn, lds);
    if (st) return st;
    const long tiles = (long)((g.nxc + G - 1) / G) * g.nouter;
    const int swz = col_tile_order(g, G, esz);
    hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(256), lds, s, (const C2<F>*)src, (C2<F>*)dst, g, G, Lp, lp.dev,
                       (const C2<F>*)lp.tw, swz);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int DIR>
static int launch_col_t(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, hipStream_t s) {
    if (lp.engine == ENGINE_REG) {
        switch (lp.P) {
            case 4: return launch_col_reg_t<F, 4, DIR>(lp, src, dst, g, s);
            case 8: return launch_col_reg_t<F, 8, DIR>(lp, src, dst, g, s);
            case 6: return launch_col_reg_t<F, 6, DIR>(lp, src, dst, g, s);
            case 10: return launch_col_reg_t<F, 10, DIR>(lp, src, dst, g, s);
            case 12: return launch_col_reg_t<F, 12, DIR>(lp, src, dst, g, s);
            case 20: return launch_col_reg_t<F, 20, DIR>(lp, src, dst, g, s);
            case 16: return launch_col_reg_t<F, 16, DIR>(lp, src, dst, g, s);
            case 32: return launch_col_reg_t<F, 32, DIR>(lp, src, dst, g, s);
            case 30: return launch_col_reg_t<F, 30, DIR>(lp, src, dst, g, s);
            default: break;
        }
    }
    return launch_col_gen_t<F, DIR>(lp, src, dst, g, s);
}

int launch_col(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, int dir, hipStream_t s) {
    bool done = false;
    const int st = fixed_col(lp, src, dst, g, dir, s, &done);
    if (st || done) return st;
    if (lp.prec == 64) return dir < 0 ? launch_col_t<double, -1>(lp, src, dst, g, s) : launch_col_t<double, +1>(lp, src, dst, g, s);
    return dir < 0 ? launch_col_t<float, -1>(lp, src, dst, g, s) : launch_col_t<float, +1>(lp, src, dst, g, s);
}

template <typename F, int MODE>
static int launch_row_t(const LinePlan& lp, const void* wph, const void* src, void* dst, long nrows, int nx,
                        hipStream_t s) {
    const int Lp = line_plan_gen_Lp(lp);
    int G = 16;
    while (G > 1 && line_plan_gen_lds(lp, G) > 65536) --G;
    const size_t lds = line_plan_gen_lds(lp, G);
    auto fn = k_row_x<F, MODE>;
    int st = set_lds((const void*)fn, lds);
    if (st) return st;
    const long blocks = (nrows + G - 1) / G;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(256), lds, s, src, dst, nrows, nx, G, Lp, lp.dev,
                       (const C2<F>*)lp.tw, (const C2<F>*)wph);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int P>
static int launch_row_reg_t(int mode, const LinePlan& lp, const void* wph, const void* src, void* dst, long nrows,
                            hipStream_t s) {
    const int S = lp.L / P;
    int G = S >= 256 ? 1 : 256 / S;
    const size_t esz = sizeof(C2<F>);
    while (G > 1 && (size_t)G * lp.L * esz > 65536) --G;
    const size_t lds = (size_t)G * lp.L * esz;
    bool done = false;
    const int fst = fixed_row(mode, lp, wph, src, dst, nrows, G, lds, s, &done);
    if (fst || done) return fst;
    const bool wave = 64 % S == 0;
    auto fn = mode == 0 ? (wave ? k_row_r2c_reg<F, P, true> : k_row_r2c_reg<F, P, false>)
                        : (wave ? k_row_c2r_reg<F, P, true> : k_row_c2r_reg<F, P, false>);
    int st = set_lds((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3((unsigned)((nrows + G - 1) / G)), dim3(G * S), lds, s, (const C2<F>*)src, (C2<F>*)dst,
                       nrows, G, lp.dev, (const C2<F>*)lp.tw, (const C2<F>*)wph);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F>
static int launch_row_f(int mode, const LinePlan& lp, const void* wph, const void* src, void* dst, long nrows, int nx,
                        hipStream_t s) {
    if (lp.engine == ENGINE_REG && (mode == 0 || mode == 2)) {
        switch (lp.P) {
            case 4: return launch_row_reg_t<F, 4>(mode, lp, wph, src, dst, nrows, s);
            case 8: return launch_row_reg_t<F, 8>(mode, lp, wph, src, dst, nrows, s);
            case 12: return launch_row_reg_t<F, 12>(mode, lp, wph, src, dst, nrows, s);
            case 16: return launch_row_reg_t<F, 16>(mode, lp, wph, src, dst, nrows, s);
            case 20: return launch_row_reg_t<F, 20>(mode, lp, wph, src, dst, nrows, s);
            default: break;
        }
    }
    switch (mode) {
        case 0: return launch_row_t<F, 0>(lp, wph, src, dst, nrows, nx, s);
        case 1: return launch_row_t<F, 1>(lp, wph, src, dst, nrows, nx, s);
        case 2: return launch_row_t<F, 2>(lp, wph, src, dst, nrows, nx, s);
        default: return launch_row_t<F, 3>(lp, wph, src, dst, nrows, nx, s);
    }
}

// ------------------------------------------------------------------ plan object
struct mcs_fft_plan {
    int ext[3];
    int prec;
    int dir;
    LinePlan px, py, pz;  // px: nx/2 (even) or nx (odd)
    void* wph = nullptr;  // exp(-2 pi i k / nx), k in [0, nx/2]
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
};

extern "C" int mcs_fft_plan_create(mcs_fft_plan** plan, const int ext[3], int prec, int dir) {
    MCS_CHECK(plan && ext, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1, MCS_ERR_INVALID, "extent must be >=1 in all dimensions");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(dir == MCS_R2C || dir == MCS_C2R, MCS_ERR_INVALID, "direction must be MCS_R2C or MCS_C2R");
    auto* p = new mcs_fft_plan();
    p->ext[0] = ext[0]; p->ext[1] = ext[1]; p->ext[2] = ext[2];
    p->prec = prec;
    p->dir = dir;
    const int nx = ext[0];
    const bool even = nx % 2 == 0;
    // even nx: the packed row of nx/2 points on the register engine when its length allows
    // (for powers of two 8 points per thread up to 512 and 16 beyond, else the default register
    // plan up to 20, at most 256 threads per row), otherwise the generic engine.  At 512^3 FP32,
    // with the fixed plans' twiddles in LDS, 8 points took R2C from 0.74-0.76 to 0.70-0.71 ms.
    int st = MCS_OK;
    bool reg = false;
    const int rowP = nx / 2 <= 512 ? 8 : 16;
    if (even && line_plan_build_reg(&p->px, nx / 2, prec, rowP) == MCS_OK) reg = true;
    if (even && !reg && rowP != 16 && line_plan_build_reg(&p->px, nx / 2, prec, 16) == MCS_OK) reg = true;
    if (even && !reg && line_plan_build(&p->px, nx / 2, prec, true, 8) == MCS_OK) {
        reg = p->px.engine == ENGINE_REG && p->px.P <= 20;
        if (!reg) line_plan_free(&p->px);
    }
    if (reg && nx / 2 / p->px.P > 256) {
        line_plan_free(&p->px);
        reg = false;
    }
    if (!reg) st = line_plan_build(&p->px, even ? nx / 2 : nx, prec, false);
    if (!st && ext[1] > 1) st = line_plan_build(&p->py, ext[1], prec);
    if (!st && ext[2] > 1) st = line_plan_build(&p->pz, ext[2], prec);
    if (!st && even) st = make_phase_table(&p->wph, nx, nx / 2 + 1, prec);
    if (!st && dir == MCS_C2R && (ext[1] > 1 || ext[2] > 1)) {
        const size_t esz = prec == 64 ? 16 : 8;
        p->scratch_bytes = (size_t)(nx / 2 + 1) * ext[1] * ext[2] * esz;
        if (hipMalloc(&p->scratch, p->scratch_bytes) != hipSuccess) {
            mcs_set_last_error(__FILE__, __LINE__, "scratch allocation failed");
            st = MCS_ERR_NOMEM;
        }
    }
    if (st) { mcs_fft_plan_destroy(p); return st; }
    *plan = p;
    return MCS_OK;
}

extern "C" int mcs_fft_plan_destroy(mcs_fft_plan* p) {
    if (!p) return MCS_OK;
    line_plan_free(&p->px);
    line_plan_free(&p->py);
    line_plan_free(&p->pz);
    if (p->wph) (void)hipFree(p->wph);
    if (p->scratch) (void)hipFree(p->scratch);
    delete p;
    return MCS_OK;
}

int fft_exec_async(mcs_fft_plan* p, const void* in, void* out, hipStream_t s) {
    const int nx = p->ext[0], ny = p->ext[1], nz = p->ext[2];
    const int nxc = nx / 2 + 1;
    const long rows = (long)ny * nz;
    const bool even = nx % 2 == 0;
    int st = MCS_OK;
    auto row = [&](int mode, const void* src, void* dst) {
        return p->prec == 64 ? launch_row_f<double>(mode, p->px, p->wph, src, dst, rows, nx, s)
                             : launch_row_f<float>(mode, p->px, p->wph, src, dst, rows, nx, s);
    };
    const ColGeom gy{nxc, (long)nxc, nz, (long)nxc * ny};
    const ColGeom gz{nxc, (long)nxc * ny, ny, (long)nxc};
    if (p->dir == MCS_R2C) {
        st = row(even ? 0 : 1, in, out);
        if (!st && ny > 1) st = launch_col(p->py, out, out, gy, -1, s);
        if (!st && nz > 1) st = launch_col(p->pz, out, out, gz, -1, s);
    } else {
        const void* src = in;
        if (nz > 1) { st = launch_col(p->pz, src, p->scratch, gz, +1, s); src = p->scratch; }
        if (!st && ny > 1) { st = launch_col(p->py, src, p->scratch, gy, +1, s); src = p->scratch; }
        if (!st) st = row(even ? 2 : 3, src, out);
    }
    return st;
}

extern "C" int mcs_fft_exec(mcs_fft_plan* p, const void* in, void* out) {
    MCS_CHECK(p && in && out, MCS_ERR_INVALID, "null argument");
    // the row pass of R2C reads real rows of nx values and writes complex rows of nx/2 + 1, so
    // in place other workgroups would overwrite rows before they are read; C2R stages its column
    // passes through the plan's scratch, except in 1D where the row pass alone runs
    MCS_CHECK(in != out || (p->dir == MCS_C2R && (p->ext[1] > 1 || p->ext[2] > 1)), MCS_ERR_INVALID,
              "in-place transform not supported (R2C, and C2R of one row, need separate input and output)");
    int st = fft_exec_async(p, in, out, mcs_stream());
    if (st) return st;
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    return MCS_OK;
}
