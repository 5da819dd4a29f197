// Fixed-plan kernels of the single-grid fused step and of the boundary FFT for
// the shapes that carry the headline workloads: axis lengths 128, 256 and 512
// in FP32 and FP64, and the lengths of the 3D ensemble's sizes (32 to 192, with
// 12- and 20-point plans for 3*2^a and 5*2^a) in FP32, with the ensembles'
// work-table column kernels for the same plans.  The ensembles' x pass keeps the
// runtime plans: with fixed bodies (up to 428 VGPRs in the 20-point x pass) the
// 3D ensemble step was slower (0.69 against 0.61 ms).
//
// They compute exactly what the runtime-plan kernels compute (same butterflies,
// same twiddle table, same operation order) with the whole line plan known at
// compile time.  Against a plan read from kernel arguments this removes the
// register-array copies between passes and the per-butterfly index arithmetic:
// the 512-point z pass (32 points per thread, two transforms) drops from about
// 3600 to about 2200 vector instructions per thread.  Column tiles address
// HBM through a uniform tile base plus a 32-bit per-thread byte offset, which
// the launcher checks.
//
// A plan the table below does not hold runs the runtime-plan kernel; so does
// every pass whose bit mcs_set_fixed_kernels has cleared, and every plan for
// which the tuner timed the runtime-plan kernel faster (LinePlan::nofx).
#include "fixed.hpp"
#include "ens_kernels.hpp"
#include "row_kernels.hpp"

#include <atomic>

static std::atomic<int> g_fixed{MCS_FIXED_COLUMNS | MCS_FIXED_XPASS};

extern "C" int mcs_set_fixed_kernels(int mask) {
    MCS_CHECK(mask >= 0 && mask <= (MCS_FIXED_COLUMNS | MCS_FIXED_XPASS), MCS_ERR_INVALID, "unknown fixed-kernel bits");
    g_fixed.store(mask);
    return MCS_OK;
}

int fixed_kernel_mask() { return g_fixed.load(); }

// lane shuffles in the fixed-plan x passes: 0 never, 1 where the tuner timed them faster
// (default), 2 wherever a plan has them
static std::atomic<int> g_shuffle{1};

extern "C" int mcs_set_lane_shuffle(int mode) {
    MCS_CHECK(mode >= 0 && mode <= 2, MCS_ERR_INVALID, "lane shuffle mode is 0, 1 or 2");
    g_shuffle.store(mode);
    return MCS_OK;
}

extern "C" int mcs_get_lane_shuffle(int* mode) {
    MCS_CHECK(mode, MCS_ERR_INVALID, "null argument");
    *mode = g_shuffle.load();
    return MCS_OK;
}

int lane_shuffle_mode() { return g_shuffle.load(); }

extern "C" int mcs_get_fixed_kernels(int* mask) {
    MCS_CHECK(mask, MCS_ERR_INVALID, "null argument");
    *mask = g_fixed.load();
    return MCS_OK;
}

// ------------------------------------------------------------------ kernels
// a tile of G adjacent columns of W (in place), S = L/P threads per column; blockDim.x == G * S
template <typename F, int P, typename FP>
MCS_DEV void fx_col_load(C2<F> (&a)[P], const char* wb, long mst, unsigned off, bool valid) {
#pragma unroll
    for (int m = 0; m < P; ++m)
        a[m] = valid ? *reinterpret_cast<const C2<F>*>(wb + m * mst + off) : mk<F>(0, 0);
}

template <typename F, int P, typename FP>
MCS_DEV void fx_col_store(const C2<F> (&a)[P], char* wb, long mst, unsigned off, bool valid) {
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) *reinterpret_cast<C2<F>*>(wb + m * mst + off) = a[m];
    }
}

// a column pass (stepper y passes, state import/export, boundary FFT); src and dst may alias
// (a workgroup reads its whole tile before it writes).  swz is the tile order of xcd_tile.
template <typename F, int P, int DIR, typename FP>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_col_fx(const C2<F>* src, C2<F>* dst, ColGeom g,
                                                                     int G, const C2<F>* __restrict__ tw, int swz) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    constexpr int S = FP::L / P;
    const int tilesx = (g.nxc + G - 1) / G;
    const int blk = (int)xcd_tile(blockIdx.x, gridDim.x, swz, tilesx, g.nouter);
    const int tile = blk % tilesx, outer = blk / tilesx;
    const int c = threadIdx.x % G, t = threadIdx.x / G;
    const bool valid = tile * G + c < g.nxc;
    const long tb = (long)outer * g.os + (long)tile * G;
    const char* sb = reinterpret_cast<const char*>(src + tb);
    char* db = reinterpret_cast<char*>(dst + tb);
    const long mst = (long)S * g.ls * (long)sizeof(C2<F>);
    const unsigned off = (unsigned)((c + (long)t * g.ls) * (long)sizeof(C2<F>));
    // the twiddle table in LDS behind the exchange area: a pass reads its twiddles in the
    // middle of the transform, where an L2 round trip stalls the wave
    constexpr int NTW = fix_tw_count(FP{});
    C2<F>* twl = lds + G * FP::L;
    for (int i = threadIdx.x; i < NTW; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    C2<F> a[P];
    fx_col_load<F, P, FP>(a, sb, mst, off, valid);
    fix_fft<F, P, DIR, false>(a, twl, t, lds, c, G, FP{});
    fx_col_store<F, P, FP>(a, db, mst, off, valid);
}

// the z pass of the step (the y pass in 2D): inverse transform, N = psi^3 (or psi^3 - psi),
// forward transform, src -> W (src may be W)
template <typename F, int P, int SPLIT, typename FP>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_col_nl_fx(const C2<F>* src, C2<F>* W, ColGeom g,
                                                                        int G, const C2<F>* __restrict__ tw, F scale,
                                                                        int64_t* d_step, int swz) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    constexpr int S = FP::L / P;
    const int tilesx = (g.nxc + G - 1) / G;
    const int blk = (int)xcd_tile(blockIdx.x, gridDim.x, swz, tilesx, g.nouter);
    const int tile = blk % tilesx, outer = blk / tilesx;
    const int c = threadIdx.x % G, t = threadIdx.x / G;
    const bool valid = tile * G + c < g.nxc;
    const char* sb = reinterpret_cast<const char*>(src + (long)outer * g.os + (long)tile * G);
    char* wb = reinterpret_cast<char*>(W + (long)outer * g.os + (long)tile * G);
    const long mst = (long)S * g.ls * (long)sizeof(C2<F>);
    const unsigned off = (unsigned)((c + (long)t * g.ls) * (long)sizeof(C2<F>));
    // the twiddle table in LDS behind the exchange area, read by both transforms
    constexpr int NTW = fix_tw_count(FP{});
    C2<F>* twl = lds + G * FP::L;
    for (int i = threadIdx.x; i < NTW; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    C2<F> a[P];
    fx_col_load<F, P, FP>(a, sb, mst, off, valid);
    fix_fft<F, P, +1, false>(a, twl, t, lds, c, G, FP{});
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const F u = a[m].x * scale, v = a[m].y * scale;
        if (SPLIT == MCS_SPLIT_REFERENCE) a[m] = mk<F>(fmad(u * u, u, -u), fmad(v * v, v, -v));
        else a[m] = mk<F>(u * u * u, v * v * v);
    }
    fix_fft<F, P, -1, false>(a, twl, t, lds, c, G, FP{});
    fx_col_store<F, P, FP>(a, wb, mst, off, valid);
    if (d_step && blockIdx.x == 0 && threadIdx.x == 0) *d_step += 1;
}

template <typename F, int P, int SPLIT, bool WAVE, typename FP, bool NOISE, bool SHUF>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS) k_xupd_fx(C2<F>* __restrict__ W, C2<F>* __restrict__ U,
                                                 const int2* __restrict__ pairs, int npairs, int G, XArgs<F> ar,
                                                 const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph,
                                                 const int64_t* __restrict__ d_step) {
    extern __shared__ unsigned char smem_raw[];
    const uint32_t draw = NOISE && ar.noise ? (uint32_t)*d_step : 0u;
    const FftPlanDev unused{};
    // the twiddle table in LDS behind the two rows per pair, read by both transforms of every
    // row (with the 4-point plan's three twiddled passes per transform, the 512^3 FP32 pass
    // went from 380 to 374 us)
    constexpr int NTW = fix_tw_count(FP{});
    C2<F>* twl = reinterpret_cast<C2<F>*>(smem_raw) + G * 2 * FP::L;
    for (int i = threadIdx.x; i < NTW; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    xupd_body<F, P, SPLIT, WAVE, FP, NOISE, SHUF>(W, U, pairs, npairs, G, ar, unused, twl, wph, draw,
                                            reinterpret_cast<C2<F>*>(smem_raw), blockIdx.x, threadIdx.x);
}

// Signatures of the kernel families the tables hold (every instance of a family has its
// family's parameters): launches go through these types, so that the arguments are checked
// against the parameters when the library compiles.
template <typename F> using ColFxFn = void (*)(const C2<F>*, C2<F>*, ColGeom, int, const C2<F>*, int);
template <typename F> using ColNlFxFn = void (*)(const C2<F>*, C2<F>*, ColGeom, int, const C2<F>*, F, int64_t*, int);
template <typename F>
using XupdFxFn = void (*)(C2<F>*, C2<F>*, const int2*, int, int, XArgs<F>, const C2<F>*, const C2<F>*,
                          const int64_t*);
template <typename F>
using RowFxFn = void (*)(const C2<F>*, C2<F>*, long, int, FftPlanDev, const C2<F>*, const C2<F>*);
template <typename T> static T typed(const void* fn) { return reinterpret_cast<T>(const_cast<void*>(fn)); }

// ------------------------------------------------------------------ table
// Each entry: precision, points per thread, length, radices in pass order (the
// runtime plan builder's order for that length and P), and the kernels.
struct ColEntry {
    int prec, P, L, np;
    int rad[4];
    const void* fwd;
    const void* inv;
    const void* nl_ns;   // psi^3
    const void* nl_ref;  // psi^3 - psi
};
struct XEntry {
    int prec, P, L, np;
    int rad[4];
    int ntw;  // twiddle table entries, staged in LDS behind the rows
    // [shuffle][split][noise]: split 0 psi^3, 1 psi^3 - psi; noise 0 without the draw (noise off);
    // the lane-shuffle forms are null where the plan has none
    const void* fn[2][2][2];
};

template <typename F, int P, int L, int... Rs> static ColEntry col_entry() {
    using FP = FixPlan<L, Rs...>;
    ColEntry e{};
    e.prec = sizeof(F) == 8 ? 64 : 32;
    e.P = P;
    e.L = L;
    const int r[] = {Rs...};
    e.np = (int)sizeof...(Rs);
    for (int i = 0; i < e.np; ++i) e.rad[i] = r[i];
    e.fwd = (const void*)k_col_fx<F, P, -1, FP>;
    e.inv = (const void*)k_col_fx<F, P, +1, FP>;
    e.nl_ns = (const void*)k_col_nl_fx<F, P, MCS_SPLIT_CUBIC, FP>;
    e.nl_ref = (const void*)k_col_nl_fx<F, P, MCS_SPLIT_REFERENCE, FP>;
    return e;
}

template <typename F, int P, int L, int... Rs> static XEntry x_entry() {
    using FP = FixPlan<L, Rs...>;
    constexpr bool wave = 64 % (L / P) == 0;
    constexpr bool shuf = wave && fix_shuffle_ok<P, L, Rs...>();
    XEntry e{};
    e.prec = sizeof(F) == 8 ? 64 : 32;
    e.P = P;
    e.L = L;
    e.ntw = fix_tw_count(FP{});
    const int r[] = {Rs...};
    e.np = (int)sizeof...(Rs);
    for (int i = 0; i < e.np; ++i) e.rad[i] = r[i];
    e.fn[0][0][1] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, true, false>;
    e.fn[0][1][1] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, true, false>;
    e.fn[0][0][0] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, false, false>;
    e.fn[0][1][0] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, false, false>;
    if constexpr (shuf) {
        e.fn[1][0][1] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, true, true>;
        e.fn[1][1][1] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, true, true>;
        e.fn[1][0][0] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, false, true>;
        e.fn[1][1][0] = (const void*)k_xupd_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, false, true>;
    }
    return e;
}

static const ColEntry* col_table(int* n) {
    static const ColEntry t[] = {
        // single grids of 256^3 and 512^3
        col_entry<float, 8, 512, 8, 8, 8>(),   col_entry<float, 16, 512, 2, 16, 16>(),
        col_entry<float, 32, 512, 16, 32>(),   col_entry<float, 8, 256, 4, 8, 8>(),
        col_entry<float, 16, 256, 16, 16>(),   col_entry<double, 8, 512, 8, 8, 8>(),
        col_entry<double, 16, 512, 2, 16, 16>(), col_entry<double, 8, 256, 4, 8, 8>(),
        col_entry<double, 16, 256, 16, 16>(),
        // the 3D ensemble's sizes {64, 96, 128, 160, 192}^3 (3*2^a and 5*2^a: 12 and 20 points)
        col_entry<float, 8, 64, 8, 8>(),       col_entry<float, 16, 64, 4, 16>(),
        col_entry<float, 12, 96, 3, 2, 4, 4>(), col_entry<float, 8, 128, 2, 8, 8>(),
        col_entry<float, 16, 128, 8, 16>(),    col_entry<float, 20, 160, 5, 2, 4, 4>(),
        col_entry<float, 12, 192, 3, 4, 4, 4>(),
    };
    *n = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}

static const XEntry* x_table(int* n) {
    static const XEntry t[] = {
        x_entry<float, 8, 256, 4, 8, 8>(),  x_entry<float, 16, 256, 16, 16>(), x_entry<float, 8, 128, 2, 8, 8>(),
        x_entry<float, 16, 128, 8, 16>(),   x_entry<double, 8, 256, 4, 8, 8>(), x_entry<double, 8, 128, 2, 8, 8>(),
        // 4 points per thread (a row pair per wavefront, more waves per SIMD): the 512^3 FP32
        // x pass 385 -> 378 us against the 8-point plan
        x_entry<float, 4, 256, 4, 4, 4, 4>(), x_entry<float, 4, 128, 2, 4, 4, 4>(),
        x_entry<double, 4, 256, 4, 4, 4, 4>(), x_entry<double, 4, 128, 2, 4, 4, 4>(),
        // x lengths nx/2 of the 3D ensemble
        x_entry<float, 8, 32, 4, 8>(),      x_entry<float, 12, 48, 3, 4, 4>(), x_entry<float, 8, 64, 8, 8>(),
        x_entry<float, 20, 80, 5, 4, 4>(),  x_entry<float, 12, 96, 3, 2, 4, 4>(),
    };
    *n = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}

template <typename E> static bool plan_matches(const E& e, const LinePlan& lp) {
    if (lp.engine != ENGINE_REG || lp.prec != e.prec || lp.P != e.P || lp.L != e.L || lp.dev.np != e.np) return false;
    for (int i = 0; i < e.np; ++i)
        if (lp.dev.radix[i] != e.rad[i]) return false;
    return true;
}

static int set_lds_fx(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

// ------------------------------------------------------------------ launchers
static const ColEntry* col_match(const LinePlan& lp, const ColGeom& g) {
    if (!(g_fixed.load() & MCS_FIXED_COLUMNS) || lp.nofx) return nullptr;
    int n = 0;
    const ColEntry* t = col_table(&n);
    const ColEntry* e = nullptr;
    for (int i = 0; i < n && !e; ++i)
        if (plan_matches(t[i], lp)) e = &t[i];
    if (!e) return nullptr;
    const int S = lp.L / lp.P;
    const long esz = lp.prec == 64 ? 16 : 8;
    // per-thread byte offsets (column + t * ls) stay below 2^32
    if (((long)S * g.ls + col_lines_per_wg(lp)) * esz >= (1L << 32)) return nullptr;
    return e;
}

int fixed_col(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, int dir, hipStream_t s, bool* done) {
    *done = false;
    const ColEntry* e = col_match(lp, g);
    if (!e) return MCS_OK;
    const int S = lp.L / lp.P;
    const int G = col_lines_per_wg(lp);
    const long esz = lp.prec == 64 ? 16 : 8;
    const int swz = col_tile_order(g, G, esz);
    const void* fn = dir < 0 ? e->fwd : e->inv;
    const size_t lds = ((size_t)G * lp.L + lp.twcount) * esz;  // the exchange area and the twiddle table
    int st = set_lds_fx(fn, lds);
    if (st) return st;
    const long tiles = (long)((g.nxc + G - 1) / G) * g.nouter;
    if (lp.prec == 64)
        hipLaunchKernelGGL(typed<ColFxFn<double>>(fn), dim3((unsigned)tiles), dim3(G * S), lds, s,
                           (const C2<double>*)src, (C2<double>*)dst, g, G, (const C2<double>*)lp.tw, swz);
    else
        hipLaunchKernelGGL(typed<ColFxFn<float>>(fn), dim3((unsigned)tiles), dim3(G * S), lds, s,
                           (const C2<float>*)src, (C2<float>*)dst, g, G, (const C2<float>*)lp.tw, swz);
    MCS_HIP_TRY(hipGetLastError());
    *done = true;
    return MCS_OK;
}

int fixed_step_col(const LinePlan& lp, const void* src, void* W, const ColGeom& g, int kind, double scale, int split,
                   int64_t* d_step, hipStream_t s, bool* done) {
    if (kind != 0) return fixed_col(lp, src, W, g, kind, s, done);
    *done = false;
    const ColEntry* e = col_match(lp, g);
    if (!e) return MCS_OK;
    const int S = lp.L / lp.P;
    const int G = col_lines_per_wg(lp);
    const long esz = lp.prec == 64 ? 16 : 8;
    const void* fn = split == MCS_SPLIT_REFERENCE ? e->nl_ref : e->nl_ns;
    const size_t lds = ((size_t)G * lp.L + lp.twcount) * esz;  // the exchange area and the twiddle table
    int st = set_lds_fx(fn, lds);
    if (st) return st;
    const long tiles = (long)((g.nxc + G - 1) / G) * g.nouter;
    const int swz = col_tile_order(g, G, esz);
    if (lp.prec == 64)
        hipLaunchKernelGGL(typed<ColNlFxFn<double>>(fn), dim3((unsigned)tiles), dim3(G * S), lds, s,
                           (const C2<double>*)src, (C2<double>*)W, g, G, (const C2<double>*)lp.tw, scale, d_step, swz);
    else
        hipLaunchKernelGGL(typed<ColNlFxFn<float>>(fn), dim3((unsigned)tiles), dim3(G * S), lds, s,
                           (const C2<float>*)src, (C2<float>*)W, g, G, (const C2<float>*)lp.tw, (float)scale, d_step,
                           swz);
    MCS_HIP_TRY(hipGetLastError());
    *done = true;
    return MCS_OK;
}

template <typename F>
static int fixed_xupd_t(const XEntry& e, const StepGeom& sg, const StepPhys& ph, void* W, void* U, hipStream_t s) {
    const LinePlan& lp = sg.px;
    XArgs<F> ar = make_xargs<F>(sg, ph);
    const int S = lp.L / lp.P;
    const int G = xupd_pairs_per_wg(lp);
    const size_t lds = ((size_t)G * 2 * lp.L + e.ntw) * sizeof(C2<F>);  // two rows per pair, twiddles
    const int mode = g_shuffle.load();
    const bool sh = e.fn[1][0][0] && (mode == 2 || (mode == 1 && lp.shuf));
    const void* fn = e.fn[sh ? 1 : 0][ph.split == MCS_SPLIT_REFERENCE ? 1 : 0][ar.noise ? 1 : 0];
    int st = set_lds_fx(fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(typed<XupdFxFn<F>>(fn), dim3((unsigned)((sg.npairs + G - 1) / G)), dim3(G * S), lds, s,
                       (C2<F>*)W, (C2<F>*)U, sg.pairs, sg.npairs, G, ar, (const C2<F>*)lp.tw, (const C2<F>*)sg.wph,
                       (const int64_t*)sg.d_step);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

int fixed_xupd(const StepGeom& sg, const StepPhys& ph, void* W, void* U, hipStream_t s, bool* done) {
    *done = false;
    if (!(g_fixed.load() & MCS_FIXED_XPASS) || sg.px.nofx) return MCS_OK;
    int n = 0;
    const XEntry* t = x_table(&n);
    for (int i = 0; i < n; ++i)
        if (plan_matches(t[i], sg.px)) {
            *done = true;
            return sg.prec == 64 ? fixed_xupd_t<double>(t[i], sg, ph, W, U, s) : fixed_xupd_t<float>(t[i], sg, ph, W, U, s);
        }
    return MCS_OK;
}

// ------------------------------------------------------------------ ensemble column passes
// The work-table column kernels of the batched ensembles with the plans of the column table
// above.  Their x pass keeps the runtime plan: its fixed bodies held up to 428 VGPRs (the
// 20-point pass) and ran slower.
struct EnsColEntry {
    int prec, P, L, np;
    int rad[4];
    const void* fwd;
    const void* inv;
    const void* nl_ns;
    const void* nl_ref;
};

template <typename F, int P, int L, int... Rs> static EnsColEntry ens_col_entry() {
    using FP = FixPlan<L, Rs...>;
    EnsColEntry e{};
    e.prec = sizeof(F) == 8 ? 64 : 32;
    e.P = P;
    e.L = L;
    const int r[] = {Rs...};
    e.np = (int)sizeof...(Rs);
    for (int i = 0; i < e.np; ++i) e.rad[i] = r[i];
    e.fwd = (const void*)k_ens_col<F, P, -1, FP>;
    e.inv = (const void*)k_ens_col<F, P, +1, FP>;
    e.nl_ns = (const void*)k_ens_col_nl<F, P, MCS_SPLIT_CUBIC, FP>;
    e.nl_ref = (const void*)k_ens_col_nl<F, P, MCS_SPLIT_REFERENCE, FP>;
    return e;
}

static const EnsColEntry* ens_col_table(int* n) {
    static const EnsColEntry t[] = {
        ens_col_entry<float, 8, 512, 8, 8, 8>(),   ens_col_entry<float, 16, 512, 2, 16, 16>(),
        ens_col_entry<float, 32, 512, 16, 32>(),   ens_col_entry<float, 8, 256, 4, 8, 8>(),
        ens_col_entry<float, 16, 256, 16, 16>(),   ens_col_entry<double, 8, 512, 8, 8, 8>(),
        ens_col_entry<double, 16, 512, 2, 16, 16>(), ens_col_entry<double, 8, 256, 4, 8, 8>(),
        ens_col_entry<double, 16, 256, 16, 16>(),
        ens_col_entry<float, 8, 64, 8, 8>(),       ens_col_entry<float, 16, 64, 4, 16>(),
        ens_col_entry<float, 12, 96, 3, 2, 4, 4>(), ens_col_entry<float, 8, 128, 2, 8, 8>(),
        ens_col_entry<float, 16, 128, 8, 16>(),    ens_col_entry<float, 20, 160, 5, 2, 4, 4>(),
        ens_col_entry<float, 12, 192, 3, 4, 4, 4>(),
    };
    *n = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}

const void* fixed_ens_col_kernel(const LinePlan& lp, int kind, int split) {
    if (!(g_fixed.load() & MCS_FIXED_COLUMNS) || lp.nofx) return nullptr;
    int n = 0;
    const EnsColEntry* t = ens_col_table(&n);
    for (int i = 0; i < n; ++i)
        if (plan_matches(t[i], lp)) {
            if (kind < 0) return t[i].fwd;
            if (kind > 0) return t[i].inv;
            return split == MCS_SPLIT_REFERENCE ? t[i].nl_ref : t[i].nl_ns;
        }
    return nullptr;
}

// ------------------------------------------------------------------ boundary row passes
// The x (row) passes of the boundary transform, R2C and C2R, for the register plans the
// boundary FFT builds for nx/2 = 128, 256 and 512 (the MCS_FIXED_XPASS bit).
struct RowEntry {
    int prec, P, L, np;
    int rad[4];
    int ntw;  // twiddle table entries, staged in LDS behind the rows
    const void* r2c;
    const void* c2r;
};

template <typename F, int P, int L, int... Rs> static RowEntry row_entry() {
    using FP = FixPlan<L, Rs...>;
    constexpr bool wave = 64 % (L / P) == 0;
    RowEntry e{};
    e.prec = sizeof(F) == 8 ? 64 : 32;
    e.P = P;
    e.L = L;
    e.ntw = fix_tw_count(FP{});
    const int r[] = {Rs...};
    e.np = (int)sizeof...(Rs);
    for (int i = 0; i < e.np; ++i) e.rad[i] = r[i];
    e.r2c = (const void*)k_row_r2c_reg<F, P, wave, FP>;
    e.c2r = (const void*)k_row_c2r_reg<F, P, wave, FP>;
    return e;
}

static const RowEntry* row_table(int* n) {
    static const RowEntry t[] = {
        row_entry<float, 16, 128, 8, 16>(),   row_entry<float, 16, 256, 16, 16>(),
        row_entry<float, 8, 256, 4, 8, 8>(),  row_entry<float, 8, 128, 2, 8, 8>(),
        row_entry<float, 8, 512, 8, 8, 8>(),
        row_entry<float, 32, 512, 16, 32>(),  row_entry<double, 8, 128, 2, 8, 8>(),
        row_entry<double, 8, 256, 4, 8, 8>(), row_entry<double, 8, 512, 8, 8, 8>(),
    };
    *n = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}

int fixed_row(int mode, const LinePlan& lp, const void* wph, const void* src, void* dst, long nrows, int G,
              size_t lds, hipStream_t s, bool* done) {
    *done = false;
    if (!(g_fixed.load() & MCS_FIXED_XPASS) || lp.nofx) return MCS_OK;
    int n = 0;
    const RowEntry* t = row_table(&n);
    const RowEntry* e = nullptr;
    for (int i = 0; i < n && !e; ++i)
        if (plan_matches(t[i], lp)) e = &t[i];
    if (!e) return MCS_OK;
    const void* fn = mode == 0 ? e->r2c : e->c2r;
    lds += (size_t)e->ntw * (lp.prec == 64 ? 16 : 8);  // the twiddle table behind the rows
    int st = set_lds_fx(fn, lds);
    if (st) return st;
    const dim3 grid((unsigned)((nrows + G - 1) / G)), block(G * (lp.L / lp.P));
    if (lp.prec == 64)
        hipLaunchKernelGGL(typed<RowFxFn<double>>(fn), grid, block, lds, s, (const C2<double>*)src, (C2<double>*)dst,
                           nrows, G, lp.dev, (const C2<double>*)lp.tw, (const C2<double>*)wph);
    else
        hipLaunchKernelGGL(typed<RowFxFn<float>>(fn), grid, block, lds, s, (const C2<float>*)src, (C2<float>*)dst,
                           nrows, G, lp.dev, (const C2<float>*)lp.tw, (const C2<float>*)wph);
    MCS_HIP_TRY(hipGetLastError());
    *done = true;
    return MCS_OK;
}

bool fixed_shuffle_available(const LinePlan& lp) {
    if (!(g_fixed.load() & MCS_FIXED_XPASS) || g_shuffle.load() == 0) return false;
    int n = 0;
    const XEntry* t = x_table(&n);
    for (int i = 0; i < n; ++i)
        if (plan_matches(t[i], lp)) return t[i].fn[1][0][0] != nullptr;
    return false;
}

bool fixed_available(const LinePlan& lp, bool xpass) {
    if (!(g_fixed.load() & (xpass ? MCS_FIXED_XPASS : MCS_FIXED_COLUMNS))) return false;
    int n = 0;
    if (xpass) {
        const XEntry* t = x_table(&n);
        for (int i = 0; i < n; ++i)
            if (plan_matches(t[i], lp)) return true;
    } else {
        const ColEntry* t = col_table(&n);
        for (int i = 0; i < n; ++i)
            if (plan_matches(t[i], lp)) return true;
    }
    return false;
}
