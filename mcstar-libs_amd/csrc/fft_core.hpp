// Stockham autosort FFT engines for CDNA4 (gfx950), written for this library.
//
// Pass p of a length-L transform with radix R and sub-length Ns (product of
// the earlier radices) maps butterfly j in [0, L/R):
//     k   = j mod Ns
//     v_q = x[j + q*L/R] * w_{Ns*R}^{q*k}             q in [0, R)
//     y[(j/Ns)*Ns*R + k + s*Ns] = DFT_R(v)_s           s in [0, R)
// which takes natural-order input to natural-order output.
//
// Register engine: a line of L points is owned by S = L/P threads, thread t
// holding the P points t + m*S (m in [0, P)).  For every radix R dividing P
// the inputs of butterflies j = t + b*S (b in [0, P/R)) are exactly those
// points, so each pass reads registers, writes its outputs to LDS, and the
// exchange reads the same pattern back.  The last pass lands its outputs in
// the thread's own registers at natural positions t + m*S, so an inverse
// transform, a pointwise op and a forward transform chain without LDS.
//
// Generic engine: ping-pong LDS buffers and a work-item loop, for the
// lengths the register engine does not take (factors 2, 3, 5, 7, 11, 13).
#pragma once

#include "mcs_common.hpp"

#define MCS_MAX_PASSES 16
// workgroup size bound of register-engine kernels: with 16+ complex floats (8+ doubles) of data
// per thread a 1024-thread workgroup would cap the kernel at 128 VGPRs
#define MCS_REG_MAX_THREADS(P, F) (((P) * sizeof(F) > 64 || (sizeof(F) == 4 && (P) >= 16)) ? 256 : 1024)
inline int mcs_reg_max_threads(int P, int prec) {
    return (P * (prec == 64 ? 8 : 4) > 64 || (prec != 64 && P >= 16)) ? 256 : 1024;
}

struct FftPlanDev {
    int L;
    int np;
    int radix[MCS_MAX_PASSES];
    int ns[MCS_MAX_PASSES];
    unsigned nsm[MCS_MAX_PASSES];  // ceil(2^32 / ns): j % ns = j - ns * umulhi(j, nsm) for the j of a line
    int twoff[MCS_MAX_PASSES];
};

// ------------------------------------------------------------ butterflies
// In-place DFT of size R over v[0..R-1]: X_s = sum_q v_q exp(DIR*2*pi*i*q*s/R).
template <typename F, int DIR> MCS_DEV void dft2(C2<F>* v) {
    const C2<F> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename F, int DIR> MCS_DEV void dft3(C2<F>* v) {
    const F c = F(-0.5);
    const F s = F(DIR) * F(0.86602540378443864676);
    const C2<F> t = cadd(v[1], v[2]);
    const C2<F> u = csub(v[1], v[2]);
    const C2<F> m = mk<F>(fmad(c, t.x, v[0].x), fmad(c, t.y, v[0].y));
    const C2<F> iu = mk<F>(-s * u.y, s * u.x);  // i*s*u
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, iu);
    v[2] = csub(m, iu);
}

template <typename F, int DIR> MCS_DEV void dft4(C2<F>* v) {
    const C2<F> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const C2<F> t2 = cadd(v[1], v[3]);
    const C2<F> t3 = cmuli<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <typename F, int DIR> MCS_DEV void dft5(C2<F>* v) {
    const F c1 = F(0.30901699437494742410), c2 = F(-0.80901699437494742410);
    const F s1 = F(DIR) * F(0.95105651629515357212), s2 = F(DIR) * F(0.58778525229247312917);
    const C2<F> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const C2<F> u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    const C2<F> a0 = v[0];
    const C2<F> m1 = mk<F>(fmad(c2, t2.x, fmad(c1, t1.x, a0.x)), fmad(c2, t2.y, fmad(c1, t1.y, a0.y)));
    const C2<F> m2 = mk<F>(fmad(c1, t2.x, fmad(c2, t1.x, a0.x)), fmad(c1, t2.y, fmad(c2, t1.y, a0.y)));
    const C2<F> n1 = mk<F>(fmad(s1, u1.x, s2 * u2.x), fmad(s1, u1.y, s2 * u2.y));
    const C2<F> n2 = mk<F>(fmad(s2, u1.x, -(s1 * u2.x)), fmad(s2, u1.y, -(s1 * u2.y)));
    const C2<F> in1 = mk<F>(-n1.y, n1.x), in2 = mk<F>(-n2.y, n2.x);
    v[0] = mk<F>(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
    v[1] = cadd(m1, in1);
    v[4] = csub(m1, in1);
    v[2] = cadd(m2, in2);
    v[3] = csub(m2, in2);
}

template <typename F, int DIR> MCS_DEV void dft8(C2<F>* v) {
    C2<F> e[4] = {v[0], v[2], v[4], v[6]};
    C2<F> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<F, DIR>(e);
    dft4<F, DIR>(o);
    const F r = F(0.70710678118654752440);
    // o1 *= w8, o2 *= w8^2 = DIR*i, o3 *= w8^3
    const C2<F> o1 = mk<F>(r * (o[1].x - F(DIR) * o[1].y), r * (o[1].y + F(DIR) * o[1].x));
    const C2<F> o2 = cmuli<DIR>(o[2]);
    const C2<F> o3 = mk<F>(r * (-o[3].x - F(DIR) * o[3].y), r * (-o[3].y + F(DIR) * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

// compile-time cos/sin of 2 pi j / n (Taylor series in double after reduction to [-pi, pi])
constexpr double ce_cos2pi(long j, long n) {
    j %= n;
    if (j < 0) j += n;
    if (2 * j > n) j -= n;
    const double x = 6.283185307179586476925286766559 * (double)j / (double)n;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; ++k) { term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k)); sum += term; }
    return sum;
}
// t * exp(DIR * 2 pi i j / N).  Called from fully unrolled loops, so j is a constant and the
// branches fold: multiples of N/4 become sign changes and swaps, odd multiples of N/8 two
// adds and two multiplies, the rest a full product with constants from the FP64 series above
// (IEEE arithmetic cannot drop a product with an exact 0 or 1 by itself)
template <typename F, int DIR, int N> MCS_DEV C2<F> twiddle_n(C2<F> t, int j) {
    j %= N;
    if (j == 0) return t;
    if ((4 * j) % N == 0) {
        const int q = 4 * j / N;  // quarter turns
        if (q == 2) return mk<F>(-t.x, -t.y);
        return q == 1 ? cmuli<DIR>(t) : cmuli<-DIR>(t);
    }
    if ((8 * j) % N == 0) {
        const int e = 8 * j / N;  // odd eighth turns: w = (c + i s) / sqrt(2)
        const F c = (e == 1 || e == 7) ? F(1) : F(-1);
        const F sn = F(DIR) * ((e == 1 || e == 3) ? F(1) : F(-1));
        const F r = F(0.70710678118654752440);
        return mk<F>(r * (c * t.x - sn * t.y), r * (sn * t.x + c * t.y));
    }
    return cmul(t, mk<F>((F)ce_cos2pi(j, N), (F)(DIR * ce_cos2pi(j - N / 4, N))));
}

// DFT of size N = N1 * N2 from N2 DFTs of size N1 (stride N2), twiddles W_N^{n2 k1}, and N1 DFTs of size N2:
//   X[k1 + N1 k2] = sum_{n2} W_N2^{n2 k2} W_N^{n2 k1} sum_{n1} W_N1^{n1 k1} x[N2 n1 + n2]
template <typename F, int N1, int N2, int DIR, typename D1, typename D2>
MCS_DEV void dft_split(C2<F>* v, D1 inner, D2 outer) {
    constexpr int N = N1 * N2;
    C2<F> y[N];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
        C2<F> t[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) t[n1] = v[N2 * n1 + n2];
        inner(t);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) y[n2 * N1 + k1] = twiddle_n<F, DIR, N>(t[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        C2<F> t[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) t[n2] = y[n2 * N1 + k1];
        outer(t);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = t[k2];
    }
}

template <typename F, int DIR> MCS_DEV void dft16(C2<F>* v) {
    dft_split<F, 4, 4, DIR>(v, [](C2<F>* t) { dft4<F, DIR>(t); }, [](C2<F>* t) { dft4<F, DIR>(t); });
}

template <typename F, int DIR> MCS_DEV void dft32(C2<F>* v) {
    dft_split<F, 8, 4, DIR>(v, [](C2<F>* t) { dft8<F, DIR>(t); }, [](C2<F>* t) { dft4<F, DIR>(t); });
}

// direct DFT for the odd primes 7, 11, 13 (generic engine only); roots in FP64
template <typename F, int R, int DIR> MCS_DEV void dftp(C2<F>* v) {
    C2<F> w[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double sn, cs;
        sincospi(2.0 * k / R, &sn, &cs);
        w[k] = mk<F>((F)cs, (F)(DIR * sn));
    }
    C2<F> out[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        C2<F> acc = v[0];
#pragma unroll
        for (int q = 1; q < R; ++q) acc = cadd(acc, cmul(v[q], w[(q * s) % R]));
        out[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = out[s];
}

template <typename F, int R, int DIR> MCS_DEV void dft(C2<F>* v) {
    if constexpr (R == 7 || R == 11 || R == 13) dftp<F, R, DIR>(v);
    else if constexpr (R == 2) dft2<F, DIR>(v);
    else if constexpr (R == 3) dft3<F, DIR>(v);
    else if constexpr (R == 4) dft4<F, DIR>(v);
    else if constexpr (R == 5) dft5<F, DIR>(v);
    else if constexpr (R == 8) dft8<F, DIR>(v);
    else if constexpr (R == 16) dft16<F, DIR>(v);
    else if constexpr (R == 32) dft32<F, DIR>(v);
}

// twiddle for pass with sub-length Ns: tw[k*(R-1) + q-1] = exp(-2 pi i q k / (Ns R)) (forward)
template <typename F, int R, int DIR>
MCS_DEV void apply_tw(C2<F>* v, const C2<F>* __restrict__ tw, int k) {
#pragma unroll
    for (int q = 1; q < R; ++q) {
        C2<F> w = tw[k * (R - 1) + (q - 1)];
        if (DIR > 0) w.y = -w.y;
        v[q] = cmul(v[q], w);
    }
}

// ------------------------------------------------------------ register engine
template <typename F, int P, int R, int DIR>
MCS_DEV void rpass(C2<F> (&a)[P], int t, int S, int Ns, unsigned nsm, const C2<F>* __restrict__ tw, C2<F>* lds,
                   int lbase, int lps, bool last) {
    constexpr int B = P / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        C2<F> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = a[b + q * B];
        const int j = t + b * S;
        const int k = Ns > 1 ? j - Ns * (int)__umulhi((unsigned)j, nsm) : 0;
        if (Ns > 1) apply_tw<F, R, DIR>(v, tw, k);
        dft<F, R, DIR>(v);
        if (last) {
#pragma unroll
            for (int q = 0; q < R; ++q) a[b + q * B] = v[q];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lbase + (base + q * Ns) * lps] = v[q];
        }
    }
}

// Synchronisation of an LDS exchange.  When all threads of a line sit in one
// wavefront (S divides 64) the exchange is wave-local: LDS operations of a
// wavefront complete in issue order, so a compiler fence at wavefront scope
// suffices and no workgroup barrier is needed.
template <bool WAVE> MCS_DEV void lds_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

template <typename F, int P, int DIR, bool WAVE = false>
MCS_DEV void reg_fft(C2<F> (&a)[P], const FftPlanDev& pl, const C2<F>* __restrict__ tw, int t, C2<F>* lds,
                     int lbase, int lps) {
    const int S = pl.L / P;
    for (int p = 0; p < pl.np; ++p) {
        const bool last = (p == pl.np - 1);
        const int R = pl.radix[p];
        const int Ns = pl.ns[p];
        const unsigned M = pl.nsm[p];
        const C2<F>* twp = tw + pl.twoff[p];
        switch (R) {
            case 2: if constexpr (P % 2 == 0) rpass<F, P, 2, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 3: if constexpr (P % 3 == 0) rpass<F, P, 3, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 4: if constexpr (P % 4 == 0) rpass<F, P, 4, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 5: if constexpr (P % 5 == 0) rpass<F, P, 5, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 8: if constexpr (P % 8 == 0) rpass<F, P, 8, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 16: if constexpr (P % 16 == 0) rpass<F, P, 16, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            case 32: if constexpr (P % 32 == 0) rpass<F, P, 32, DIR>(a, t, S, Ns, M, twp, lds, lbase, lps, last); break;
            default: break;
        }
        if (!last) {
            lds_sync<WAVE>();
#pragma unroll
            for (int m = 0; m < P; ++m) a[m] = lds[lbase + (t + m * S) * lps];
            lds_sync<WAVE>();
        }
    }
}

// entries of a runtime plan's twiddle table (passes with Ns > 1 hold Ns (R - 1) each)
MCS_DEV int plan_tw_count(const FftPlanDev& pl) {
    int n = 0;
    for (int p = 0; p < pl.np; ++p)
        if (pl.ns[p] > 1) n = max(n, pl.twoff[p] + pl.ns[p] * (pl.radix[p] - 1));
    return n;
}

// Copies a twiddle table of n entries into LDS at dst with every thread of the workgroup and
// returns dst (ends with a barrier); n = 0 (the launcher found no room for the table in LDS)
// returns tw itself.  Each pass of a transform reads its twiddles in the middle of the
// transform; from global memory that is an L2 round trip per pass, a chain that short launches
// (few workgroups, nothing to hide it behind) wait on in full.
template <typename F> MCS_DEV const C2<F>* stage_twiddles(const C2<F>* __restrict__ tw, int n, C2<F>* dst) {
    if (n == 0) return tw;
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) dst[i] = tw[i];
    __syncthreads();
    return dst;
}

// ------------------------------------------------------------ lane exchanges
// A swap stage of a register/lane transpose within a wavefront: the lanes whose bit D is clear
// take the lo value of their partner lane l ^ D into hi; the lanes whose bit D is set take the
// partner's hi into lo.  D = 1, 2: DPP quad permutations; 4, 8: DPP rotations within a row of 16
// lanes, each written only to the lanes (banks of 4) that need it; 16, 32: gfx950's
// v_permlane16_swap / v_permlane32_swap, which exchange exactly those halves between two
// registers.  No LDS, no barrier.
template <int D> MCS_DEV void lane_swap32(unsigned& lo, unsigned& hi) {
    static_assert(D == 1 || D == 2 || D == 4 || D == 8 || D == 16 || D == 32, "lane distance");
    if constexpr (D == 1 || D == 2) {
        constexpr int qp = D == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] and [2,3,0,1]
        const unsigned plo = (unsigned)__builtin_amdgcn_mov_dpp((int)lo, qp, 0xF, 0xF, false);
        const unsigned phi = (unsigned)__builtin_amdgcn_mov_dpp((int)hi, qp, 0xF, 0xF, false);
        const bool up = (__lane_id() & D) != 0;
        lo = up ? phi : lo;
        hi = up ? hi : plo;
    } else if constexpr (D == 4 || D == 8) {
        // row_ror:n gives lane l the value of lane (l - n) mod 16 within its row
        constexpr int from_up = D == 4 ? 0x12C : 0x128;  // from lane l + D (ror 12, ror 8)
        constexpr int from_dn = D == 4 ? 0x124 : 0x128;  // from lane l - D (ror 4, ror 8)
        constexpr int clear = D == 4 ? 0x5 : 0x3;        // banks whose lanes have bit D clear
        const unsigned nhi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)lo, from_up, 0xF, clear, false);
        const unsigned nlo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)hi, from_dn, 0xF, 0xF ^ clear, false);
        lo = nlo;
        hi = nhi;
    } else if constexpr (D == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    }
}

template <int D> MCS_DEV void lane_swap(float& lo, float& hi) {
    unsigned a = __float_as_uint(lo), b = __float_as_uint(hi);
    lane_swap32<D>(a, b);
    lo = __uint_as_float(a);
    hi = __uint_as_float(b);
}

template <int D> MCS_DEV void lane_swap(double& lo, double& hi) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(lo);
    const unsigned long long b = (unsigned long long)__double_as_longlong(hi);
    unsigned al = (unsigned)a, ah = (unsigned)(a >> 32), bl = (unsigned)b, bh = (unsigned)(b >> 32);
    lane_swap32<D>(al, bl);
    lane_swap32<D>(ah, bh);
    lo = __longlong_as_double((long long)(((unsigned long long)ah << 32) | al));
    hi = __longlong_as_double((long long)(((unsigned long long)bh << 32) | bl));
}

// The exchange before a last pass of radix P that follows a pass of radix P, done in registers.
// There thread t = a NS + k (a in [0, P), k in [0, NS), NS = L / P^2) holds output q of its
// butterfly in v[q], and thread q NS + k needs it in slot a: a P x P transpose between the lane
// bits of a and the register index, in log2(P) swap stages (stage i pairs lane bit NS 2^i with
// register bit i).  Needs the line's S = L/P threads to be lanes of one wavefront in order.
template <typename F, int P, int NS, int I = 1> MCS_DEV void lane_transpose(C2<F> (&v)[P]) {
    if constexpr (I < P) {
#pragma unroll
        for (int j = 0; j < P; ++j)
            if ((j & I) == 0) {
                lane_swap<NS * I>(v[j].x, v[j + I].x);
                lane_swap<NS * I>(v[j].y, v[j + I].y);
            }
        lane_transpose<F, P, NS, 2 * I>(v);
    }
}

// Two lines transformed together pass by pass (the x pass's row and mirror row, both owned by
// the same threads): every pass writes both lines' butterfly outputs, then one synchronisation,
// then both read back, so a pair takes half the synchronisations (and their latency) of two
// reg_fft calls.  Each line sees the same operations as in reg_fft, so the results are the same.
template <typename F, int P, int R, int DIR>
MCS_DEV void rpass2(C2<F> (&a)[P], C2<F> (&b)[P], int t, int S, int Ns, unsigned nsm, const C2<F>* __restrict__ tw,
                    C2<F>* lds, int la, int lb, int lps, bool last) {
    rpass<F, P, R, DIR>(a, t, S, Ns, nsm, tw, lds, la, lps, last);
    rpass<F, P, R, DIR>(b, t, S, Ns, nsm, tw, lds, lb, lps, last);
}

template <typename F, int P, int DIR, bool WAVE = false>
MCS_DEV void reg_fft2(C2<F> (&a)[P], C2<F> (&b)[P], const FftPlanDev& pl, const C2<F>* __restrict__ tw, int t,
                      C2<F>* lds, int la, int lb, int lps) {
    const int S = pl.L / P;
    for (int p = 0; p < pl.np; ++p) {
        const bool last = (p == pl.np - 1);
        const int R = pl.radix[p];
        const int Ns = pl.ns[p];
        const unsigned M = pl.nsm[p];
        const C2<F>* twp = tw + pl.twoff[p];
        switch (R) {
            case 2: if constexpr (P % 2 == 0) rpass2<F, P, 2, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 3: if constexpr (P % 3 == 0) rpass2<F, P, 3, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 4: if constexpr (P % 4 == 0) rpass2<F, P, 4, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 5: if constexpr (P % 5 == 0) rpass2<F, P, 5, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 8: if constexpr (P % 8 == 0) rpass2<F, P, 8, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 16: if constexpr (P % 16 == 0) rpass2<F, P, 16, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            case 32: if constexpr (P % 32 == 0) rpass2<F, P, 32, DIR>(a, b, t, S, Ns, M, twp, lds, la, lb, lps, last); break;
            default: break;
        }
        if (!last) {
            lds_sync<WAVE>();
#pragma unroll
            for (int m = 0; m < P; ++m) {
                a[m] = lds[la + (t + m * S) * lps];
                b[m] = lds[lb + (t + m * S) * lps];
            }
            lds_sync<WAVE>();
        }
    }
}

// ------------------------------------------------------------ fixed plans
// The register engine with the whole plan known at compile time: length L and
// radices Rs... in pass order (sub-lengths and twiddle-table offsets follow as
// in the runtime plan, so both read the same table and compute the same
// values).  Index arithmetic folds to constants and the passes unroll into one
// straight-line body, so no register array is copied between passes.
struct RunPlan {
    static constexpr bool fixed = false;
    static constexpr int L = 0;
};
template <int L_, int... Rs> struct FixPlan {
    static constexpr bool fixed = true;
    static constexpr int L = L_;
};

template <int R, int... Rest> constexpr int first_radix() { return R; }

// whether the exchange before the last pass of plan (L, Rs...) can run as lane_transpose: the
// last two radices equal P and a line's S = L/P threads are lanes of one wavefront
template <int P, int L, int... Rs> constexpr bool fix_shuffle_ok() {
    constexpr int n = sizeof...(Rs);
    constexpr int r[] = {Rs...};
    return n >= 2 && r[n - 1] == P && r[n - 2] == P && 64 % (L / P) == 0 && L % (P * P) == 0;
}

// SHUF: the exchange before the last pass through lane_transpose where fix_shuffle_ok holds
// (the same values land in the same registers, so results are the same bits)
template <typename F, int P, int DIR, bool WAVE, bool SHUF, int L, int Ns, int TwOff, int R, int... Rest>
MCS_DEV void fix_pass(C2<F> (&a)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int lbase, int lps) {
    constexpr int S = L / P;
    constexpr int B = P / R;
    constexpr bool last = sizeof...(Rest) == 0;
    static_assert(P % R == 0, "every radix of a register plan divides P");
    if constexpr (!last) {
        constexpr bool shuffle = SHUF && WAVE && sizeof...(Rest) == 1 && R == P && first_radix<Rest...>() == P &&
                                 64 % S == 0;
        if constexpr (shuffle) {
            const int k = Ns > 1 ? t % Ns : 0;
            if constexpr (Ns > 1) apply_tw<F, R, DIR>(a, tw + TwOff, k);
            dft<F, R, DIR>(a);
            lane_transpose<F, P, Ns>(a);
            fix_pass<F, P, DIR, WAVE, SHUF, L, Ns * R, TwOff + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>(a, tw, t, lds,
                                                                                               lbase, lps);
            return;
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        C2<F> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = a[b + q * B];
        const int j = t + b * S;
        const int k = Ns > 1 ? j % Ns : 0;
        if constexpr (Ns > 1) apply_tw<F, R, DIR>(v, tw + TwOff, k);
        dft<F, R, DIR>(v);
        if constexpr (last) {
#pragma unroll
            for (int q = 0; q < R; ++q) a[b + q * B] = v[q];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lbase + (base + q * Ns) * lps] = v[q];
        }
    }
    if constexpr (!last) {
        lds_sync<WAVE>();
#pragma unroll
        for (int m = 0; m < P; ++m) a[m] = lds[lbase + (t + m * S) * lps];
        lds_sync<WAVE>();
        fix_pass<F, P, DIR, WAVE, SHUF, L, Ns * R, TwOff + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>(a, tw, t, lds,
                                                                                           lbase, lps);
    }
}

// the butterflies of one fixed pass over one line: outputs to LDS, or back to a on the last pass
template <typename F, int P, int DIR, int L, int Ns, int TwOff, int R, bool LAST>
MCS_DEV void fix_bfly(C2<F> (&a)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int lbase, int lps) {
    constexpr int S = L / P;
    constexpr int B = P / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        C2<F> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = a[b + q * B];
        const int j = t + b * S;
        const int k = Ns > 1 ? j % Ns : 0;
        if constexpr (Ns > 1) apply_tw<F, R, DIR>(v, tw + TwOff, k);
        dft<F, R, DIR>(v);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) a[b + q * B] = v[q];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lbase + (base + q * Ns) * lps] = v[q];
        }
    }
}

// two lines together pass by pass, as reg_fft2 (SHUF as for fix_pass)
template <typename F, int P, int DIR, bool WAVE, bool SHUF, int L, int Ns, int TwOff, int R, int... Rest>
MCS_DEV void fix_pass2(C2<F> (&a)[P], C2<F> (&b)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int la, int lb,
                       int lps) {
    constexpr int S = L / P;
    constexpr bool last = sizeof...(Rest) == 0;
    static_assert(P % R == 0, "every radix of a register plan divides P");
    if constexpr (!last) {
        constexpr bool shuffle = SHUF && WAVE && sizeof...(Rest) == 1 && R == P && first_radix<Rest...>() == P &&
                                 64 % S == 0;
        if constexpr (shuffle) {
            const int k = Ns > 1 ? t % Ns : 0;
            if constexpr (Ns > 1) {
                apply_tw<F, R, DIR>(a, tw + TwOff, k);
                apply_tw<F, R, DIR>(b, tw + TwOff, k);
            }
            dft<F, R, DIR>(a);
            dft<F, R, DIR>(b);
            lane_transpose<F, P, Ns>(a);
            lane_transpose<F, P, Ns>(b);
            fix_pass2<F, P, DIR, WAVE, SHUF, L, Ns * R, TwOff + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>(a, b, tw, t, lds,
                                                                                                la, lb, lps);
            return;
        }
    }
    fix_bfly<F, P, DIR, L, Ns, TwOff, R, last>(a, tw, t, lds, la, lps);
    fix_bfly<F, P, DIR, L, Ns, TwOff, R, last>(b, tw, t, lds, lb, lps);
    if constexpr (!last) {
        lds_sync<WAVE>();
#pragma unroll
        for (int m = 0; m < P; ++m) {
            a[m] = lds[la + (t + m * S) * lps];
            b[m] = lds[lb + (t + m * S) * lps];
        }
        lds_sync<WAVE>();
        fix_pass2<F, P, DIR, WAVE, SHUF, L, Ns * R, TwOff + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>(a, b, tw, t, lds, la,
                                                                                            lb, lps);
    }
}

template <typename F, int P, int DIR, bool WAVE, int L, int... Rs>
MCS_DEV void fix_fft(C2<F> (&a)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int lbase, int lps,
                     FixPlan<L, Rs...>) {
    fix_pass<F, P, DIR, WAVE, false, L, 1, 0, Rs...>(a, tw, t, lds, lbase, lps);
}

// the same with the last exchange through lane shuffles where the plan allows it
template <typename F, int P, int DIR, bool WAVE, int L, int... Rs>
MCS_DEV void fix_fft_shuf(C2<F> (&a)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int lbase, int lps,
                          FixPlan<L, Rs...>) {
    fix_pass<F, P, DIR, WAVE, true, L, 1, 0, Rs...>(a, tw, t, lds, lbase, lps);
}

// entries of a fixed plan's twiddle table (Ns (R - 1) per pass with Ns > 1)
template <int Ns, int R, int... Rest> constexpr int fix_tw_count_from() {
    if constexpr (sizeof...(Rest) == 0) return Ns > 1 ? Ns * (R - 1) : 0;
    else return (Ns > 1 ? Ns * (R - 1) : 0) + fix_tw_count_from<Ns * R, Rest...>();
}
template <int L, int... Rs> constexpr int fix_tw_count(FixPlan<L, Rs...>) { return fix_tw_count_from<1, Rs...>(); }

template <typename F, int P, int DIR, bool WAVE, bool SHUF, int L, int... Rs>
MCS_DEV void fix_fft2(C2<F> (&a)[P], C2<F> (&b)[P], const C2<F>* __restrict__ tw, int t, C2<F>* lds, int la, int lb,
                      int lps, FixPlan<L, Rs...>) {
    fix_pass2<F, P, DIR, WAVE, SHUF, L, 1, 0, Rs...>(a, b, tw, t, lds, la, lb, lps);
}

// two lines through line_fft's transform together, pass by pass (LDS areas la and lb)
template <typename F, int P, int DIR, bool WAVE, typename FP, bool SHUF = false>
MCS_DEV void line_fft2(C2<F> (&a)[P], C2<F> (&b)[P], const FftPlanDev& pl, const C2<F>* __restrict__ tw, int t,
                       C2<F>* lds, int la, int lb, int lps) {
    if constexpr (FP::fixed) fix_fft2<F, P, DIR, WAVE, SHUF>(a, b, tw, t, lds, la, lb, lps, FP{});
    else reg_fft2<F, P, DIR, WAVE>(a, b, pl, tw, t, lds, la, lb, lps);
}

// twiddle entries of the runtime plan pl or of the fixed plan FP
template <typename FP> MCS_DEV int line_tw_count(const FftPlanDev& pl) {
    if constexpr (FP::fixed) return fix_tw_count(FP{});
    else return plan_tw_count(pl);
}

// the line transform of a pass body: runtime plan pl, or the fixed plan FP (pl unused); SHUF:
// a fixed plan's last exchange through lane shuffles where the plan allows it
template <typename F, int P, int DIR, bool WAVE, typename FP, bool SHUF = false>
MCS_DEV void line_fft(C2<F> (&a)[P], const FftPlanDev& pl, const C2<F>* __restrict__ tw, int t, C2<F>* lds,
                      int lbase, int lps) {
    if constexpr (FP::fixed && SHUF) fix_fft_shuf<F, P, DIR, WAVE>(a, tw, t, lds, lbase, lps, FP{});
    else if constexpr (FP::fixed) fix_fft<F, P, DIR, WAVE>(a, tw, t, lds, lbase, lps, FP{});
    else reg_fft<F, P, DIR, WAVE>(a, pl, tw, t, lds, lbase, lps);
}

// ------------------------------------------------------------ generic engine
// Lines live in LDS as buf[c*Lp + pos], c in [0, G).
template <typename F, int R, int DIR>
MCS_DEV void gpass(const C2<F>* in, C2<F>* out, int L, int Lp, int G, int Ns, const C2<F>* __restrict__ tw, int tid,
                   int nthr) {
    const int nb = L / R;
    for (int w = tid; w < G * nb; w += nthr) {
        const int c = w / nb, j = w - c * nb;
        C2<F> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = in[c * Lp + j + q * nb];
        const int k = j % Ns;
        if (Ns > 1) apply_tw<F, R, DIR>(v, tw, k);
        dft<F, R, DIR>(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) out[c * Lp + base + q * Ns] = v[q];
    }
}

// returns the buffer holding the result
template <typename F, int DIR>
MCS_DEV C2<F>* lds_fft(C2<F>* b0, C2<F>* b1, const FftPlanDev& pl, const C2<F>* __restrict__ tw, int Lp, int G,
                       int tid, int nthr) {
    C2<F>* in = b0;
    C2<F>* out = b1;
    for (int p = 0; p < pl.np; ++p) {
        const int R = pl.radix[p], Ns = pl.ns[p];
        const C2<F>* twp = tw + pl.twoff[p];
        switch (R) {
            case 2: gpass<F, 2, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 3: gpass<F, 3, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 4: gpass<F, 4, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 5: gpass<F, 5, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 8: gpass<F, 8, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 7: gpass<F, 7, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 11: gpass<F, 11, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            case 13: gpass<F, 13, DIR>(in, out, pl.L, Lp, G, Ns, twp, tid, nthr); break;
            default: break;
        }
        __syncthreads();
        C2<F>* tmp = in; in = out; out = tmp;
    }
    return in;
}
