// Row (x) passes of the boundary transform on the register engine, for the
// runtime line plan (RunPlan) or a fixed plan (pl unused, FixPlan<L, ...>).
#pragma once

#include "fft_kernels.hpp"

// a fixed plan's twiddle table copied into LDS behind the G rows (the launcher sizes the LDS for
// it); runtime plans read the global table
template <typename F, typename FP> MCS_DEV const C2<F>* row_twiddles(const C2<F>* tw, C2<F>* behind) {
    if constexpr (FP::fixed) {
        constexpr int NTW = fix_tw_count(FP{});
        for (int i = threadIdx.x; i < NTW; i += blockDim.x) behind[i] = tw[i];
        __syncthreads();
        return behind;
    } else {
        (void)behind;
        return tw;
    }
}

// x pass of the boundary transform for even nx on the register engine: G rows of
// L = nx/2 packed points per workgroup, S = L/P threads per row.
//   R2C: packed FFT in registers, then X_k = E_k + w^k O_k with the mirror Z_{L-k}
//        read back through LDS; k = L (Nyquist) comes from k = 0 with w^L = -1
//   C2R: Z_k = (X_k + X*_{L-k}) + i (X_k - X*_{L-k}) w^-k with the mirror element
//        read back through LDS, inverse FFT in registers
template <typename F, int P, bool WAVE, typename FP = RunPlan>
__global__ void __launch_bounds__(256) k_row_r2c_reg(const C2<F>* __restrict__ src, C2<F>* __restrict__ dst,
                                                     long nrows, int G, FftPlanDev pl,
                                                     const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int L = FP::fixed ? FP::L : pl.L, S = L / P;
    const int t = threadIdx.x % S, g = threadIdx.x / S;
    const long row = (long)blockIdx.x * G + g;
    const bool valid = row < nrows && g < G;
    const int base = (g < G ? g : 0) * L;
    const C2<F>* twr = row_twiddles<F, FP>(tw, lds + G * L);
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? src[row * L + t + m * S] : mk<F>(0, 0);
    line_fft<F, P, -1, WAVE, FP>(a, pl, twr, t, lds, base, 1);
    lds_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) lds[base + t + m * S] = a[m];
    lds_sync<WAVE>();
    if (!valid) return;
    C2<F>* out = dst + row * (L + 1);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int k = t + m * S;
        const C2<F> zk = a[m];
        const C2<F> zp = cconj(lds[base + (k == 0 ? 0 : L - k)]);
        const C2<F> e = cscale(cadd(zk, zp), F(0.5));
        const C2<F> o = cscale(cmuli<-1>(csub(zk, zp)), F(0.5));
        out[k] = cadd(e, cmul(wph[k], o));
        if (k == 0) out[L] = csub(e, o);
    }
}

template <typename F, int P, bool WAVE, typename FP = RunPlan>
__global__ void __launch_bounds__(256) k_row_c2r_reg(const C2<F>* __restrict__ src, C2<F>* __restrict__ dst,
                                                     long nrows, int G, FftPlanDev pl,
                                                     const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int L = FP::fixed ? FP::L : pl.L, S = L / P;
    const int t = threadIdx.x % S, g = threadIdx.x / S;
    const long row = (long)blockIdx.x * G + g;
    const bool valid = row < nrows && g < G;
    const int base = (g < G ? g : 0) * L;
    const C2<F>* twr = row_twiddles<F, FP>(tw, lds + G * L);
    const C2<F>* in = src + (valid ? row : 0) * (L + 1);
    // the row is read once; the mirror X_{L-k} comes back through LDS (X_L, the
    // mirror of k = 0, is read by the row's first thread)
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? in[t + m * S] : mk<F>(0, 0);
    const C2<F> xnyq = valid && t == 0 ? in[L] : mk<F>(0, 0);
#pragma unroll
    for (int m = 0; m < P; ++m) lds[base + t + m * S] = a[m];
    lds_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int k = t + m * S;
        C2<F> xk = a[m];
        C2<F> xp = k == 0 ? xnyq : lds[base + L - k];
        if (k == 0) { xk.y = F(0); xp.y = F(0); }
        const C2<F> sm = cadd(xk, cconj(xp));
        const C2<F> d = cmulc(csub(xk, cconj(xp)), wph[k]);
        a[m] = cadd(sm, cmuli<+1>(d));
    }
    lds_sync<WAVE>();  // every mirror is read before the transform's exchanges reuse the LDS
    line_fft<F, P, +1, WAVE, FP>(a, pl, twr, t, lds, base, 1);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) dst[row * L + t + m * S] = a[m];
    }
}
