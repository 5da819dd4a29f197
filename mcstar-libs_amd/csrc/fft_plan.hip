// Line plans: factorisation, radix ordering and twiddle generation.
#include "fft_plan.hpp"

#include <cmath>
#include <vector>

// L = 2^a 3^b 5^c * rest; returns rest
static int factor235(int L, int& a, int& b, int& c) {
    a = b = c = 0;
    int m = L;
    while (m % 2 == 0) { m /= 2; ++a; }
    while (m % 3 == 0) { m /= 3; ++b; }
    while (m % 5 == 0) { m /= 5; ++c; }
    return m;
}

// the generic engine also takes the primes 7, 11 and 13
static bool generic_ok(int L, std::vector<int>* extra) {
    int a, b, c;
    int m = factor235(L, a, b, c);
    for (int p : {7, 11, 13})
        while (m % p == 0) { m /= p; if (extra) extra->push_back(p); }
    return m == 1;
}

// Register-engine plans: P points per thread, every radix divides P, and the S = L/P threads of
// a line within the launch bound of the register kernels (mcs_reg_max_threads(P, prec): 256 when a
// thread holds more than 64 bytes of data or 16+ floats, else 1024).

// power-of-two L with P points per thread: radix-P passes after one remainder pass
static bool reg_plan_pow2(int L, int P, int prec, std::vector<int>* rad) {
    int a, b, c;
    if (L < 4 || factor235(L, a, b, c) != 1 || b || c) return false;
    if (P < 4 || (P & (P - 1)) || P > L || L / P > mcs_reg_max_threads(P, prec)) return false;
    int lp = 0;
    while ((1 << lp) < P) ++lp;
    std::vector<int> r;
    if (a % lp) r.push_back(1 << (a % lp));
    for (int i = 0; i < a / lp; ++i) r.push_back(P);
    if (rad) *rad = r;
    return true;
}

// 3^b*2^a with 6 points per thread or 5^c*2^a with 10: the radix-3 or radix-5 passes, then radix-2
// passes (twice the threads of the 12- and 20-point plans, for small grids)
static bool reg_plan_small(int L, int P, int prec, std::vector<int>* rad) {
    int a, b, c;
    if (L < 6 || factor235(L, a, b, c) != 1 || a < 1) return false;
    const bool three = P == 6 && b >= 1 && c == 0, five = P == 10 && b == 0 && c >= 1;
    if (!(three || five) || a + b + c > MCS_MAX_PASSES || L / P > mcs_reg_max_threads(P, prec)) return false;
    std::vector<int> r;
    for (int i = 0; i < (three ? b : c); ++i) r.push_back(three ? 3 : 5);
    for (int i = 0; i < a; ++i) r.push_back(2);
    if (rad) *rad = r;
    return true;
}

// The default plan.  Powers of two take in-register radix-16/32 butterflies so that 256, 512
// and 1024 need a single LDS exchange per transform; 3^b*2^a and 5^c*2^a use radix-3/5 passes and
// radix-4 passes (12 and 20 points), and lengths with both 3 and 5 radix-3, -5 and -2 passes on
// 30 points (the sizes cuda/math.hpp:139-149's roundup235 yields, as 1200 = 3 5^2 2^4).  FP64 keeps P = 8 for powers of two (radix-8 passes): larger register blocks of
// doubles drop the kernels to one wavefront per SIMD.  Where that plan's line would exceed the
// launch bound, the next plan within it is taken (32 points for 8192 in FP32, 6 or 10 points for
// the long 3*2^a and 5*2^a lines); none within it means no register plan.
static int reg_plan(int L, int prec, std::vector<int>* rad, int max_p = 32) {
    int a, b, c;
    if (L < 4 || factor235(L, a, b, c) != 1) return 0;
    std::vector<int> r;
    int P = 0;
    if (b == 0 && c == 0) {
        if (a == 2) { P = 4; r = {4}; }
        else if (a <= 6 || prec == 64 || max_p < 16) {  // radix-8 passes
            P = 8;
            if (a % 3 == 1) r.push_back(2);
            if (a % 3 == 2) r.push_back(4);
            for (int i = 0; i < a / 3; ++i) r.push_back(8);
        } else if (a == 7) { P = 16; r = {8, 16}; }
        else if (a == 8) { P = 16; r = {16, 16}; }
        else if (a == 9) { P = 32; r = {16, 32}; }
        else if (a == 10) { P = 32; r = {32, 32}; }
        else {  // radix-16 passes with the remainder first
            P = 16;
            if (a % 4 == 1) r.push_back(2);
            if (a % 4 == 2) r.push_back(4);
            if (a % 4 == 3) r.push_back(8);
            for (int i = 0; i < a / 4; ++i) r.push_back(16);
        }
    } else if (b >= 1 && c == 0 && a >= 2) {  // 3^b 2^a: radix-3 passes, then radix-4 (and one radix-2)
        P = 12;
        for (int i = 0; i < b; ++i) r.push_back(3);
        if (a % 2 == 1) r.push_back(2);
        for (int i = 0; i < a / 2; ++i) r.push_back(4);
    } else if (b == 0 && c >= 1 && a >= 2) {  // 5^c 2^a: radix-5 passes, then radix-4 (and one radix-2)
        P = 20;
        for (int i = 0; i < c; ++i) r.push_back(5);
        if (a % 2 == 1) r.push_back(2);
        for (int i = 0; i < a / 2; ++i) r.push_back(4);
    } else if (b >= 1 && c >= 1 && a >= 1) {  // 3^b 5^c 2^a: 30 points, which radices 2, 3 and 5 divide
        P = 30;
        for (int i = 0; i < b; ++i) r.push_back(3);
        for (int i = 0; i < c; ++i) r.push_back(5);
        for (int i = 0; i < a; ++i) r.push_back(2);
    }
    if (P == 0 || (int)r.size() > MCS_MAX_PASSES) return 0;
    if (L / P > mcs_reg_max_threads(P, prec)) {
        P = 0;
        for (int Q : {32, 16, 8}) {
            if (Q > max_p || (prec == 64 && Q > 16)) continue;
            if (reg_plan_pow2(L, Q, prec, &r)) { P = Q; break; }
        }
        if (!P)
            for (int Q : {6, 10})
                if (reg_plan_small(L, Q, prec, &r)) { P = Q; break; }
        if (!P) return 0;
    }
    if (rad) *rad = r;
    return P;
}

int line_plan_reg_P(int L, int prec) { return reg_plan(L, prec, nullptr); }

std::vector<int> line_plan_reg_options(int L, int prec) {
    std::vector<int> out;
    const int def = reg_plan(L, prec, nullptr);
    if (!def) return out;
    out.push_back(def);
    const size_t e = prec == 64 ? 16 : 8;
    for (int P : {4, 8, 16, 32}) {
        if (P == def || (prec == 64 && P > 16) || (size_t)L * e > MCS_LDS_MAX) continue;
        if (reg_plan_pow2(L, P, prec, nullptr)) out.push_back(P);
    }
    for (int P : {6, 10})
        if (P != def && (size_t)L * e <= MCS_LDS_MAX && reg_plan_small(L, P, prec, nullptr)) out.push_back(P);
    return out;
}

// exchange LDS of a tile: within 64 KB, or a single line (row pair) up to the workgroup maximum
static bool tile_lds_ok(size_t bytes, int G) { return bytes <= 65536 || (G == 1 && bytes <= MCS_LDS_MAX); }

bool col_tile_ok(int L, int P, int prec, int G) {
    const size_t esz = prec == 64 ? 16 : 8;
    return P > 0 && L % P == 0 && G >= 1 && (long)G * (L / P) <= mcs_reg_max_threads(P, prec) &&
           tile_lds_ok((size_t)G * L * esz, G);
}

bool x_tile_ok(int L, int P, int prec, int G) {
    const size_t esz = prec == 64 ? 16 : 8;
    return P > 0 && L % P == 0 && G >= 1 && (long)G * (L / P) <= MCS_XPASS_MAX_THREADS &&
           tile_lds_ok((size_t)G * 2 * L * esz, G);
}

int col_lines_per_wg_default(int L, int P, int prec) {
    const int S = L / P;
    const size_t esz = prec == 64 ? 16 : 8;
    int G = (int)(128 / esz);
    if (G > mcs_reg_max_threads(P, prec) / S) G = mcs_reg_max_threads(P, prec) / S;
    while (G > 1 && (size_t)G * L * esz > 65536) --G;
    return G < 1 ? 1 : G;
}

int xupd_pairs_per_wg_default(int L, int P, int prec) {
    const int S = L / P;
    const size_t esz = prec == 64 ? 16 : 8;
    int G = S >= MCS_XPASS_MAX_THREADS ? 1 : MCS_XPASS_MAX_THREADS / S;
    while (G > 1 && (size_t)G * 2 * L * esz > 65536) --G;
    return G;
}

int col_lines_per_wg(const LinePlan& lp) {
    return lp.G > 0 ? lp.G : col_lines_per_wg_default(lp.L, lp.P, lp.prec);
}

int xupd_pairs_per_wg(const LinePlan& lp) {
    return lp.G > 0 ? lp.G : xupd_pairs_per_wg_default(lp.L, lp.P, lp.prec);
}

static void gen_radices(int L, std::vector<int>& r) {
    int a, b, c;
    factor235(L, a, b, c);
    r.clear();
    if (a % 3 == 1) r.push_back(2);
    if (a % 3 == 2) r.push_back(4);
    for (int i = 0; i < a / 3; ++i) r.push_back(8);
    for (int i = 0; i < b; ++i) r.push_back(3);
    for (int i = 0; i < c; ++i) r.push_back(5);
    generic_ok(L, &r);
}

int line_plan_gen_Lp(const LinePlan& lp) { return lp.L + 1; }

size_t line_plan_gen_lds(const LinePlan& lp, int G) {
    const size_t e = lp.prec == 64 ? 16 : 8;
    return 2 * (size_t)G * (size_t)line_plan_gen_Lp(lp) * e;
}

template <typename F> static int upload(void** dev, const std::vector<double>& re, const std::vector<double>& im) {
    std::vector<C2<F>> h(re.size());
    for (size_t i = 0; i < re.size(); ++i) { h[i].x = (F)re[i]; h[i].y = (F)im[i]; }
    if (h.empty()) { *dev = nullptr; return MCS_OK; }
    MCS_HIP_TRY(hipMalloc(dev, h.size() * sizeof(C2<F>)));
    MCS_HIP_TRY(hipMemcpy(*dev, h.data(), h.size() * sizeof(C2<F>), hipMemcpyHostToDevice));
    return MCS_OK;
}

// exp(-2 pi i j / n) in FP64, exact at the quarter and half turns
static void phase(long j, long n, double& c, double& s) {
    j %= n;
    if (j < 0) j += n;
    const double a = 2.0 * M_PI * (double)j / (double)n;
    c = std::cos(a);
    s = -std::sin(a);
    if (n % 4 == 0) {
        const long q = n / 4;
        if (j == 0) { c = 1; s = 0; }
        else if (j == q) { c = 0; s = -1; }
        else if (j == 2 * q) { c = -1; s = 0; }
        else if (j == 3 * q) { c = 0; s = 1; }
    } else if (n % 2 == 0 && j == n / 2) { c = -1; s = 0; }
}

static int plan_finish(LinePlan* lp, const std::vector<int>& rad);

int line_plan_build(LinePlan* lp, int L, int prec, bool allow_reg, int max_p) {
    MCS_CHECK(L >= 1, MCS_ERR_INVALID, "FFT length must be >= 1");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(generic_ok(L, nullptr), MCS_ERR_UNSUPPORTED, "FFT length must have no prime factor above 13");
    lp->L = L;
    lp->prec = prec;
    lp->G = 0;
    std::vector<int> rad;
    const int P = allow_reg ? reg_plan(L, prec, nullptr, max_p) : 0;
    if (P) {
        lp->engine = ENGINE_REG;
        lp->P = P;
        reg_plan(L, prec, &rad, max_p);
        const size_t e = prec == 64 ? 16 : 8;
        MCS_CHECK((size_t)L * e <= MCS_LDS_MAX, MCS_ERR_UNSUPPORTED, "FFT length too large for LDS");
    } else {
        lp->engine = ENGINE_GEN;
        lp->P = 0;
        gen_radices(L, rad);
        MCS_CHECK(line_plan_gen_lds(*lp, 1) <= MCS_LDS_MAX, MCS_ERR_UNSUPPORTED, "FFT length too large for LDS");
    }
    return plan_finish(lp, rad);
}

int line_plan_build_reg(LinePlan* lp, int L, int prec, int P) {
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    std::vector<int> rad;
    if (P == reg_plan(L, prec, nullptr)) {
        reg_plan(L, prec, &rad);
    } else {
        MCS_CHECK(reg_plan_pow2(L, P, prec, &rad) || reg_plan_small(L, P, prec, &rad), MCS_ERR_UNSUPPORTED,
                  "no register plan with this many points per thread");
    }
    const size_t e = prec == 64 ? 16 : 8;
    MCS_CHECK((size_t)L * e <= MCS_LDS_MAX, MCS_ERR_UNSUPPORTED, "FFT length too large for LDS");
    lp->L = L;
    lp->prec = prec;
    lp->G = 0;
    lp->engine = ENGINE_REG;
    lp->P = P;
    return plan_finish(lp, rad);
}

static int plan_finish(LinePlan* lp, const std::vector<int>& rad) {
    const int L = lp->L, prec = lp->prec;
    MCS_CHECK((int)rad.size() <= MCS_MAX_PASSES, MCS_ERR_UNSUPPORTED, "too many FFT passes");
    FftPlanDev& d = lp->dev;
    d = FftPlanDev{};
    d.L = L;
    d.np = (int)rad.size();
    std::vector<double> re, im;
    int Ns = 1;
    for (int p = 0; p < d.np; ++p) {
        const int R = rad[p];
        d.radix[p] = R;
        d.ns[p] = Ns;
        d.nsm[p] = 0;
        if (Ns > 1) {
            const uint64_t m = ((uint64_t)1 << 32) / (uint64_t)Ns + (((uint64_t)1 << 32) % (uint64_t)Ns ? 1 : 0);
            d.nsm[p] = (unsigned)m;
            for (int j = 0; j < L; ++j)  // the identity is exact for every index of a line; checked, not assumed
                MCS_CHECK((int)(((uint64_t)j * m) >> 32) == j / Ns, MCS_ERR_UNSUPPORTED, "index reciprocal inexact");
        }
        d.twoff[p] = (int)re.size();
        if (Ns > 1) {
            for (int k = 0; k < Ns; ++k)
                for (int q = 1; q < R; ++q) {
                    double cr, ci;
                    phase((long)q * k, (long)Ns * R, cr, ci);
                    re.push_back(cr);
                    im.push_back(ci);
                }
        }
        Ns *= R;
    }
    lp->twcount = re.size();
    int st = prec == 64 ? upload<double>(&lp->tw, re, im) : upload<float>(&lp->tw, re, im);
    return st;
}

void line_plan_free(LinePlan* lp) {
    if (lp->tw) (void)hipFree(lp->tw);
    lp->tw = nullptr;
}

int make_phase_table(void** dev, int n, int count, int prec) {
    std::vector<double> re(count), im(count);
    for (int k = 0; k < count; ++k) phase(k, n, re[k], im[k]);
    return prec == 64 ? upload<double>(dev, re, im) : upload<float>(dev, re, im);
}
