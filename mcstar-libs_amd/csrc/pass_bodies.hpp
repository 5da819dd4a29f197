// Device bodies of the stepper's passes, shared by the single-grid kernels and
// the ensemble's work-table kernels.  Each body processes one work unit (a
// tile of columns, a block of rows or a block of row pairs) with the LDS the
// caller provides; a persistent caller loops over units and must
// __syncthreads() between them.
#pragma once

#include "fft_kernels.hpp"
#include "pfc_math.hpp"
#include "philox.hpp"

// The bodies take the thread index as an argument.  Persistent callers pass
// opaque_tid() on every iteration: with threadIdx.x itself, the compiler hoists
// all address arithmetic derived from it (P offsets per access, LDS and twiddle
// indices) out of the unit loop and keeps it live across the whole body, which
// doubled the register use of the loop kernels and spilled the P = 12 ones.
MCS_DEV int opaque_tid() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// ------------------------------------------------------------------ column pass
// tile of G adjacent x columns (thread c = tid % G, t = tid / G), L = pl.L points each
template <typename F, int P, int DIR, typename FP = RunPlan>
MCS_DEV void col_body(const C2<F>* src, C2<F>* dst, const ColGeom& g, int G, const FftPlanDev& pl,
                      const C2<F>* __restrict__ tw, C2<F>* lds, int tile, int outer, int tid) {
    const int S = (FP::fixed ? FP::L : pl.L) / P;
    const int c = tid % G, t = tid / G;
    const int x = tile * G + c;
    const bool valid = x < g.nxc && t < S;
    const long base = (long)outer * g.os + x;
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? src[base + (long)(t + m * S) * g.ls] : mk<F>(0, 0);
    line_fft<F, P, DIR, false, FP>(a, pl, tw, t < S ? t : 0, lds, c, G);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) dst[base + (long)(t + m * S) * g.ls] = a[m];
    }
}

// column pass with the nonlinearity between the inverse and forward transforms, src -> W
// (src may be W; FP: RunPlan, the line plan pl, or a FixPlan):
// u = scale * value holds psi at (2m, 2m+1); N = psi^3 (or psi^3 - psi)
template <typename F, int P, int SPLIT, typename FP = RunPlan>
MCS_DEV void col_nl_body(const C2<F>* src, C2<F>* W, const ColGeom& g, int G, const FftPlanDev& pl,
                         const C2<F>* __restrict__ tw, F scale, C2<F>* lds, int tile, int outer, int tid) {
    const int S = (FP::fixed ? FP::L : pl.L) / P;
    const int c = tid % G, t = tid / G;
    const int x = tile * G + c;
    const bool valid = x < g.nxc && t < S;
    const int tt = t < S ? t : 0;
    const long base = (long)outer * g.os + x;
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? src[base + (long)(t + m * S) * g.ls] : mk<F>(0, 0);
    line_fft<F, P, +1, false, FP>(a, pl, tw, tt, lds, c, G);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const F u = a[m].x * scale, v = a[m].y * scale;
        if (SPLIT == MCS_SPLIT_REFERENCE) a[m] = mk<F>(fmad(u * u, u, -u), fmad(v * v, v, -v));
        else a[m] = mk<F>(u * u * u, v * v * v);
    }
    line_fft<F, P, -1, false, FP>(a, pl, tw, tt, lds, c, G);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) W[base + (long)(t + m * S) * g.ls] = a[m];
    }
}

// ------------------------------------------------------------------ row pass on packed rows
template <typename F, int P, int DIR, bool WAVE = false>
MCS_DEV void row_body(const C2<F>* src, C2<F>* dst, long nrows, int G, const FftPlanDev& pl,
                      const C2<F>* __restrict__ tw, C2<F>* lds, long blk, int tid) {
    const int L = pl.L, S = L / P;
    const int t = tid % S, g = tid / S;
    const long row = blk * G + g;
    const bool valid = row < nrows && g < G;
    C2<F> a[P];
#pragma unroll
    for (int m = 0; m < P; ++m) a[m] = valid ? src[row * L + t + m * S] : mk<F>(0, 0);
    reg_fft<F, P, DIR, WAVE>(a, pl, tw, t, lds, (g < G ? g : 0) * L, 1);
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) dst[row * L + t + m * S] = a[m];
    }
}

// ------------------------------------------------------------------ x pass with the update
template <typename F>
struct XArgs {
    int nx, ny, nz, L;
    F dkx, dky, dkz;  // 2 pi / (n h) per axis
    F dt, lin0, amp;
    uint32_t key0, key1;
    int noise;
    double mass;  // FP64 k=0 mode (sum of psi), re-imposed exactly
};

MCS_DEV int ksigned(int i, int n) { return i < n / 2 + 1 ? i : i - n; }

// One packed mode p of row c = (iy, iz) together with mode -p of the mirror row m = (my, mz):
//   E = (A + conj B)/2, O = -i (A - conj B)/2 with A = U(c, p), B = U(m, -p)
//   Psi1 = E + w^p O (full mode p), Psi2 = E - w^p O (full mode p + nx/2)
// both modes are updated, then E' = (Psi1' + Psi2')/2, O' = (Psi1' - Psi2') conj(w^p) / 2 and
//   U'(c, p) = E' + i O',  U'(m, -p) = conj(E') + i conj(O').
// NOISE = false compiles the noise draw out (ar.noise must then be 0): the draw's call keeps
// many registers live across the unrolled update (the 512-point FP32 x pass: 125 VGPRs against 191)
template <typename F, bool NOISE = true, bool INL = false>
MCS_DEV void xupd_pair(C2<F> nA, C2<F> nBr, C2<F> pA, C2<F> pBr, int p, int iy, int iz, int my, int mz, C2<F> w,
                       const XArgs<F>& ar, uint32_t draw, C2<F>* uc, C2<F>* um) {
    const C2<F> nB = cconj(nBr), pB = cconj(pBr);
    const F h = F(0.5);
    const C2<F> eN = cscale(cadd(nA, nB), h), oN = cscale(cmuli<-1>(csub(nA, nB)), h);
    const C2<F> eP = cscale(cadd(pA, pB), h), oP = cscale(cmuli<-1>(csub(pA, pB)), h);
    const C2<F> woN = cmul(w, oN), woP = cmul(w, oP);
    const C2<F> n1 = cadd(eN, woN), n2 = csub(eN, woN);
    const C2<F> p1 = cadd(eP, woP), p2 = csub(eP, woP);
    const F ky = F(ksigned(iy, ar.ny)) * ar.dky, kz = F(ksigned(iz, ar.nz)) * ar.dkz;
    const F kyz = fmad(ky, ky, kz * kz);
    const F kx1 = F(p) * ar.dkx;
    const F kx2 = F(ar.L - p) * ar.dkx;  // |k_x| of index p + nx/2 (the Nyquist index for p = 0)
    C2<F> xi1 = mk<F>(0, 0), xi2 = mk<F>(0, 0);
    if (NOISE && ar.noise) {
        // one Philox block gives both modes; a self-mirrored row takes mode p + L as the
        // conjugate of its partner L - p (a second block) and real modes at p = 0
        const long rc = (long)iz * ar.ny + iy;
        const uint64_t ctr = (uint64_t)rc * ar.L + p;
        const XiPair<F> x = INL ? xi_block_inl<F>(ctr, ar.key0, ar.key1, draw) : xi_block<F>(ctr, ar.key0, ar.key1, draw);
        xi1 = x.a;
        xi2 = x.b;
        if (iy == my && iz == mz) {
            const F s2 = F(1.41421356237309504880);
            if (p == 0) {
                xi1 = mk<F>(s2 * x.a.x, F(0));
                xi2 = mk<F>(s2 * x.b.x, F(0));
            } else {
                const uint64_t ctr2 = (uint64_t)rc * ar.L + (ar.L - p);
                xi2 = cconj((INL ? xi_block_inl<F>(ctr2, ar.key0, ar.key1, draw) : xi_block<F>(ctr2, ar.key0, ar.key1, draw)).a);
            }
        }
    }
    const C2<F> q1 = pfc_mode_update<F>(p1, n1, xi1, fmad(kx1, kx1, kyz), ar.dt, ar.lin0, ar.amp);
    const C2<F> q2 = pfc_mode_update<F>(p2, n2, xi2, fmad(kx2, kx2, kyz), ar.dt, ar.lin0, ar.amp);
    if (p == 0 && iy == 0 && iz == 0) {
        // k = 0: the mean density is carried in FP64 and re-imposed exactly
        const double m = ar.mass;
        const double er = 0.5 * (m + (double)q2.x), ei = 0.5 * (double)q2.y;
        const double or_ = 0.5 * (m - (double)q2.x), oi = -0.5 * (double)q2.y;
        *uc = mk<F>((F)(er - oi), (F)(ei + or_));
        *um = *uc;
        return;
    }
    const C2<F> e = cscale(cadd(q1, q2), h);
    const C2<F> o = cscale(cmulc(csub(q1, q2), w), h);
    *uc = cadd(e, cmuli<+1>(o));
    *um = cadd(cconj(e), cmuli<+1>(cconj(o)));
}

// G row pairs (row, mirror row) per unit, S = L/P threads per pair; LDS 2 rows per pair.
// Each thread updates the mode pairs {(c, p), (m, -p)} of its own positions p once; the
// mirror row's results move to their owners through LDS.  A self-mirrored row (c == m)
// is processed as row c alone.
// FP: RunPlan (the line plan pl) or a FixPlan of length L = nx/2; NOISE as for xupd_pair;
// SHUF: a fixed plan's last exchange of every line transform through lane shuffles (fix_fft_shuf).
template <typename F, int P, int SPLIT, bool WAVE = false, typename FP = RunPlan, bool NOISE = true,
          bool SHUF = false>
MCS_DEV void xupd_body(C2<F>* __restrict__ W, C2<F>* __restrict__ U, const int2* __restrict__ pairs, int npairs,
                       int G, const XArgs<F>& ar, const FftPlanDev& pl, const C2<F>* __restrict__ tw,
                       const C2<F>* __restrict__ wph, uint32_t draw, C2<F>* lds, int blk, int tid) {
    const int L = FP::fixed ? FP::L : pl.L, S = L / P;
    // the noise draw inline where the unrolled body stays in registers (with it, the 20-point
    // FP32 and the 16- and 20-point FP64 bodies went to scratch); a call elsewhere
    constexpr bool XI_INLINE = P <= 12 || (P == 16 && sizeof(F) == 4);
    const int t = tid % S, g = tid / S;
    const int q = blk * G + g;
    const bool valid = q < npairs && g < G;
    const int2 pr = pairs[valid ? q : 0];
    const long rc = pr.x, rm = pr.y;
    const bool self = rc == rm;
    const int base = (g < G ? g : 0) * 2 * L;
    C2<F>* Nc = lds + base;
    C2<F>* Nm = Nc + L;

    // two P-element arrays live at a time: the rows' spectra, then the state pairs updated in place
    C2<F> a[P], b[P];
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
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int p = t + m * S;
        const int pp = p == 0 ? 0 : L - p;
        a[m] = valid ? U[rc * L + p] : mk<F>(0, 0);
        b[m] = valid ? U[rm * L + pp] : mk<F>(0, 0);
    }
    lds_sync<WAVE>();
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
    lds_sync<WAVE>();
    if (valid) {
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int p = t + m * S;
            const int pp = p == 0 ? 0 : L - p;
            U[rc * L + p] = a[m];
            if (!self) {
                U[rm * L + pp] = b[m];
                Nm[pp] = b[m];
            }
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
            W[rc * L + t + m * S] = a[m];
            if (!self) W[rm * L + t + m * S] = b[m];
        }
    }
}
