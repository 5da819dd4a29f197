// The x pass of the stepper's mixed form.
//
// In the spectral form the stepper keeps U = FFT3(u) and its x pass reads and
// writes both U and the work array W (four sweeps).  In the mixed form the
// state is S = IFFT_x(U) / L: real along x, spectral along y and z (the array
// the spectral x pass leaves in W, scaled by 1/L).  The x pass then reads W
// and S and writes S only (three sweeps):
//     U   = FFT_x(S)               the state's spectrum, in registers
//     N_k = FFT_x(W)               the nonlinear term's spectrum
//     U'  = update(U, N_k, noise)  on mode pairs, exactly as in the spectral form
//     S'  = IFFT_x(U') / L
// and the y inverse pass reads S and writes W.  A 3D step moves 9 sweeps of
// N/2 complex values instead of 10 (2D: 5 instead of 6); the x pass does three
// pairs of line transforms instead of two.
#pragma once

#include "pass_bodies.hpp"

// G row pairs per unit, S = L/P threads per pair; LDS 3 rows per pair (the two spectra of the
// nonlinear term, then the mirror row's state spectrum).  invL = 1/L (exact for L a power of two).
// SHUF: a fixed plan's last exchange of every line transform through lane shuffles (fix_fft_shuf).
template <typename F, int P, int SPLIT, bool WAVE = false, typename FP = RunPlan, bool NOISE = true,
          bool SHUF = false>
MCS_DEV void xupd_mixed_body(const C2<F>* __restrict__ W, C2<F>* __restrict__ Sx, const int2* __restrict__ pairs,
                             int npairs, int G, const XArgs<F>& ar, const FftPlanDev& pl,
                             const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph, uint32_t draw, F invL,
                             C2<F>* lds, int blk, int tid) {
    const int L = FP::fixed ? FP::L : pl.L, S = L / P;
    constexpr bool XI_INLINE = P <= 12 || (P == 16 && sizeof(F) == 4);
    const int t = tid % S, g = tid / S;
    const int q = blk * G + g;
    const bool valid = q < npairs && g < G;
    const int2 pr = pairs[valid ? q : 0];
    const long rc = pr.x, rm = pr.y;
    const bool self = rc == rm;
    const int base = (g < G ? g : 0) * 3 * L;
    C2<F>* Nc = lds + base;
    C2<F>* Nm = Nc + L;
    C2<F>* Xm = Nm + L;

    C2<F> a[P], b[P];
    // the nonlinear term's spectra of both rows, kept in LDS
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        a[m] = valid ? W[rc * L + p] : mk<F>(0, 0);
        b[m] = valid ? W[rm * L + p] : mk<F>(0, 0);
    }
    line_fft2<F, P, -1, WAVE, FP, SHUF>(a, b, pl, tw, t, lds, base, base + L, 1);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        Nc[p] = a[m];
        Nm[p] = b[m];
    }
    // the state's spectra: row c stays in registers, row m goes through LDS to the owners of -p
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        a[m] = valid ? Sx[rc * L + p] : mk<F>(0, 0);
        b[m] = valid ? Sx[rm * L + p] : mk<F>(0, 0);
    }
    line_fft<F, P, -1, WAVE, FP, SHUF>(a, pl, tw, t, lds, base + 2 * L, 1);
    line_fft<F, P, -1, WAVE, FP, SHUF>(b, pl, tw, t, lds, base + 2 * L, 1);
#pragma unroll
    for (int m = 0; m < P; ++m) Xm[t + m * S] = b[m];
    lds_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        b[m] = Xm[p == 0 ? 0 : L - p];
    }
    const int iyc = (int)(rc % ar.ny), izc = (int)(rc / ar.ny);
    const int iym = (int)(rm % ar.ny), izm = (int)(rm / ar.ny);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        const int pp = p == 0 ? 0 : L - p;
        C2<F> uc, um;
        xupd_pair<F, NOISE, XI_INLINE>(Nc[p], Nm[pp], a[m], b[m], p, iyc, izc, iym, izm, wph[p], ar, draw, &uc, &um);
        a[m] = uc;
        b[m] = um;
    }
    // the mirror row's new modes back to their owners
    lds_sync<WAVE>();
    if (valid && !self) {
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int p = t + m * S;
            Nm[p == 0 ? 0 : L - p] = b[m];
        }
    }
    lds_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) b[m] = self ? a[m] : Nm[t + m * S];
    lds_sync<WAVE>();
    line_fft2<F, P, +1, WAVE, FP, SHUF>(a, b, pl, tw, t, lds, base, base + L, 1);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) {
            Sx[rc * L + t + m * S] = cscale(a[m], invL);
            if (!self) Sx[rm * L + t + m * S] = cscale(b[m], invL);
        }
    }
}
