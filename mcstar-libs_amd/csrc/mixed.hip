// Kernels of the stepper's mixed form (the x pass that reads the nonlinear term W
// and the state S = IFFT_x(U) / L and writes S only): runtime-plan kernels for
// every register plan, and fixed-plan kernels for the x lengths of the headline
// single grids (nx/2 = 128 and 256, FP32 and FP64) with the radix sequences of
// the spectral form's fixed x kernels.  Both compute the same operations in the
// same order, so the results are the same bits.
#include "fixed.hpp"
#include "pass_mixed.hpp"
#include "stepper.hpp"

template <typename F, int P, int SPLIT, bool WAVE, bool NOISE>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS)
    k_xupd_mx(const C2<F>* __restrict__ W, C2<F>* __restrict__ Sx, const int2* __restrict__ pairs, int npairs, int G,
              XArgs<F> ar, FftPlanDev pl, const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph,
              const int64_t* __restrict__ d_step, F invL, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const uint32_t draw = NOISE && ar.noise ? (uint32_t)*d_step : 0u;
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * 3 * pl.L);  // behind the exchange area
    xupd_mixed_body<F, P, SPLIT, WAVE, RunPlan, NOISE>(W, Sx, pairs, npairs, G, ar, pl, twl, wph, draw, invL, lds,
                                                       blockIdx.x, threadIdx.x);
}

// SHUF: the last exchange of every line transform through lane shuffles (plans with fix_shuffle_ok)
template <typename F, int P, int SPLIT, bool WAVE, typename FP, bool NOISE, bool SHUF>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS)
    k_xupd_mx_fx(const C2<F>* __restrict__ W, C2<F>* __restrict__ Sx, const int2* __restrict__ pairs, int npairs,
                 int G, XArgs<F> ar, const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph,
                 const int64_t* __restrict__ d_step, F invL) {
    extern __shared__ unsigned char smem_raw[];
    const uint32_t draw = NOISE && ar.noise ? (uint32_t)*d_step : 0u;
    const FftPlanDev unused{};
    // the twiddle table in LDS behind the three rows per pair, read by all three transforms
    constexpr int NTW = fix_tw_count(FP{});
    C2<F>* twl = reinterpret_cast<C2<F>*>(smem_raw) + G * 3 * FP::L;
    for (int i = threadIdx.x; i < NTW; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    xupd_mixed_body<F, P, SPLIT, WAVE, FP, NOISE, SHUF>(W, Sx, pairs, npairs, G, ar, unused, twl, wph, draw, invL,
                                                        reinterpret_cast<C2<F>*>(smem_raw), blockIdx.x, threadIdx.x);
}

// row pairs per workgroup of the mixed x pass: the plan's tile, reduced until three rows of LDS
// per pair fit 64 KB (a single pair may use up to MCS_LDS_MAX); the tile does not change results
int xupd_mixed_pairs_per_wg(const LinePlan& lp) {
    const size_t esz = lp.prec == 64 ? 16 : 8;
    int G = xupd_pairs_per_wg(lp);
    while (G > 1 && (size_t)G * 3 * lp.L * esz > 65536) --G;
    return G;
}

static size_t mixed_lds(const LinePlan& lp, int G) { return (size_t)G * 3 * lp.L * (lp.prec == 64 ? 16 : 8); }

static int set_lds_mx(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

// ------------------------------------------------------------------ fixed-plan table
struct MxEntry {
    int prec, P, L, np;
    int rad[4];
    int ntw;  // twiddle table entries, staged in LDS behind the rows
    // [shuffle][split][noise]: split 0 psi^3, 1 psi^3 - psi; noise 0 without the draw (noise off);
    // the shuffle forms are null where the plan has none
    const void* fn[2][2][2];
};

template <typename F, int P, int L, int... Rs> static MxEntry mx_entry() {
    using FP = FixPlan<L, Rs...>;
    constexpr bool wave = 64 % (L / P) == 0;
    constexpr bool shuf = wave && fix_shuffle_ok<P, L, Rs...>();
    MxEntry e{};
    e.prec = sizeof(F) == 8 ? 64 : 32;
    e.P = P;
    e.L = L;
    e.ntw = fix_tw_count(FP{});
    const int r[] = {Rs...};
    e.np = (int)sizeof...(Rs);
    for (int i = 0; i < e.np; ++i) e.rad[i] = r[i];
    e.fn[0][0][1] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, true, false>;
    e.fn[0][1][1] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, true, false>;
    e.fn[0][0][0] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, false, false>;
    e.fn[0][1][0] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, false, false>;
    if constexpr (shuf) {
        e.fn[1][0][1] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, true, true>;
        e.fn[1][1][1] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, true, true>;
        e.fn[1][0][0] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_CUBIC, wave, FP, false, true>;
        e.fn[1][1][0] = (const void*)k_xupd_mx_fx<F, P, MCS_SPLIT_REFERENCE, wave, FP, false, true>;
    }
    return e;
}

static const MxEntry* mx_table(int* n) {
    static const MxEntry t[] = {
        mx_entry<float, 8, 256, 4, 8, 8>(),  mx_entry<float, 16, 256, 16, 16>(), mx_entry<float, 8, 128, 2, 8, 8>(),
        mx_entry<float, 16, 128, 8, 16>(),   mx_entry<double, 8, 256, 4, 8, 8>(), mx_entry<double, 8, 128, 2, 8, 8>(),
        // 4 points per thread, as in the spectral form's table
        mx_entry<float, 4, 256, 4, 4, 4, 4>(), mx_entry<float, 4, 128, 2, 4, 4, 4>(),
        mx_entry<double, 4, 256, 4, 4, 4, 4>(), mx_entry<double, 4, 128, 2, 4, 4, 4>(),
    };
    *n = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}

static const MxEntry* mx_match(const LinePlan& lp) {
    if (!(fixed_kernel_mask() & MCS_FIXED_XPASS) || lp.nofx) return nullptr;
    if (lp.engine != ENGINE_REG) return nullptr;
    int n = 0;
    const MxEntry* t = mx_table(&n);
    for (int i = 0; i < n; ++i) {
        const MxEntry& e = t[i];
        if (lp.prec != e.prec || lp.P != e.P || lp.L != e.L || lp.dev.np != e.np) continue;
        bool same = true;
        for (int j = 0; j < e.np; ++j) same = same && lp.dev.radix[j] == e.rad[j];
        if (same) return &e;
    }
    return nullptr;
}

bool fixed_mixed_available(const LinePlan& lp) { return mx_match(lp) != nullptr; }

bool fixed_mixed_shuffle_available(const LinePlan& lp) {
    const MxEntry* e = mx_match(lp);
    return e && e->fn[1][0][0] && lane_shuffle_mode() != 0;
}

// whether the launch takes the shuffle form of e for lp
static bool use_shuffle(const MxEntry& e, const LinePlan& lp) {
    const int mode = lane_shuffle_mode();
    return e.fn[1][0][0] && (mode == 2 || (mode == 1 && lp.shuf));
}

// ------------------------------------------------------------------ launchers
template <typename F>
static int launch_mx_fx(const MxEntry& e, const StepGeom& sg, const XArgs<F>& ar, int split, const void* W, void* Sx,
                        hipStream_t s) {
    const LinePlan& lp = sg.px;
    const int S = lp.L / lp.P;
    const int G = xupd_mixed_pairs_per_wg(lp);
    const size_t lds = mixed_lds(lp, G) + (size_t)e.ntw * sizeof(C2<F>);  // rows, twiddles
    const void* fn = e.fn[use_shuffle(e, lp) ? 1 : 0][split == MCS_SPLIT_REFERENCE ? 1 : 0][ar.noise ? 1 : 0];
    int st = set_lds_mx(fn, lds);
    if (st) return st;
    // every fixed-plan kernel of the table has this signature: launched through it, so that the
    // arguments are checked against the parameters when the library compiles
    using MxFn = void (*)(const C2<F>*, C2<F>*, const int2*, int, int, XArgs<F>, const C2<F>*, const C2<F>*,
                          const int64_t*, F);
    hipLaunchKernelGGL(reinterpret_cast<MxFn>(const_cast<void*>(fn)), dim3((unsigned)((sg.npairs + G - 1) / G)),
                       dim3(G * S), lds, s, (const C2<F>*)W, (C2<F>*)Sx, sg.pairs, sg.npairs, G, ar,
                       (const C2<F>*)lp.tw, (const C2<F>*)sg.wph, (const int64_t*)sg.d_step, (F)(1.0 / lp.L));
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int P, int SPLIT>
static int launch_mx_t(const StepGeom& sg, const XArgs<F>& ar, const void* W, void* Sx, hipStream_t s) {
    const LinePlan& lp = sg.px;
    const int S = lp.L / P;
    const int G = xupd_mixed_pairs_per_wg(lp);
    const int twn = twiddles_in_lds(lp, mixed_lds(lp, G), (sg.npairs + G - 1) / G);
    const size_t lds = mixed_lds(lp, G) + (size_t)twn * sizeof(C2<F>);  // exchange area, twiddles
    // the noise-free kernels carry no noise draw
    auto fn = ar.noise ? ((64 % S == 0) ? k_xupd_mx<F, P, SPLIT, true, true> : k_xupd_mx<F, P, SPLIT, false, true>)
                       : ((64 % S == 0) ? k_xupd_mx<F, P, SPLIT, true, false> : k_xupd_mx<F, P, SPLIT, false, false>);
    int st = set_lds_mx((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3((unsigned)((sg.npairs + G - 1) / G)), dim3(G * S), lds, s, (const C2<F>*)W,
                       (C2<F>*)Sx, sg.pairs, sg.npairs, G, ar, lp.dev, (const C2<F>*)lp.tw, (const C2<F>*)sg.wph,
                       sg.d_step, (F)(1.0 / lp.L), twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int SPLIT>
static int launch_mx_d(const StepGeom& sg, const XArgs<F>& ar, const void* W, void* Sx, hipStream_t s) {
    switch (sg.px.P) {
        case 4: return launch_mx_t<F, 4, SPLIT>(sg, ar, W, Sx, s);
        case 8: return launch_mx_t<F, 8, SPLIT>(sg, ar, W, Sx, s);
        case 6: return launch_mx_t<F, 6, SPLIT>(sg, ar, W, Sx, s);
        case 10: return launch_mx_t<F, 10, SPLIT>(sg, ar, W, Sx, s);
        case 12: return launch_mx_t<F, 12, SPLIT>(sg, ar, W, Sx, s);
        case 20: return launch_mx_t<F, 20, SPLIT>(sg, ar, W, Sx, s);
        case 16: return launch_mx_t<F, 16, SPLIT>(sg, ar, W, Sx, s);
        case 32: return launch_mx_t<F, 32, SPLIT>(sg, ar, W, Sx, s);
        case 30: return launch_mx_t<F, 30, SPLIT>(sg, ar, W, Sx, s);
        default: mcs_set_last_error(__FILE__, __LINE__, "x plan not register-eligible"); return MCS_ERR_UNSUPPORTED;
    }
}

template <typename F>
static int launch_mx(const StepGeom& sg, const StepPhys& ph, const void* W, void* Sx, hipStream_t s) {
    const XArgs<F> ar = make_xargs<F>(sg, ph);
    if (const MxEntry* e = mx_match(sg.px)) return launch_mx_fx<F>(*e, sg, ar, ph.split, W, Sx, s);
    return ph.split == MCS_SPLIT_REFERENCE ? launch_mx_d<F, MCS_SPLIT_REFERENCE>(sg, ar, W, Sx, s)
                                           : launch_mx_d<F, MCS_SPLIT_CUBIC>(sg, ar, W, Sx, s);
}

int launch_xupd_mixed(const StepGeom& sg, const StepPhys& ph, const void* W, void* Sx, hipStream_t s) {
    MCS_CHECK(sg.px.engine == ENGINE_REG, MCS_ERR_UNSUPPORTED, "x plan not register-eligible");
    return sg.prec == 64 ? launch_mx<double>(sg, ph, W, Sx, s) : launch_mx<float>(sg, ph, W, Sx, s);
}
