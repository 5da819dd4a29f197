/*
 * FP64 CPU oracle for the PFC spectral time step (TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP path, never part of the product).
 *
 * Each function names the reference file:line it restates.  Parity pins:
 *   FFT            numpy.fft goldens (tests/golden), closed forms, Parseval
 *   Philox         Random123 known-answer vectors
 *   noise          exact Hermitian invariants + spectrum statistics
 *   update / step  single-mode closed form, k=0 invariant, both splits
 *   peaks          brute force on tiny grids, special cases
 *   bond order     closed-form lattice values (fcc/hcp/bcc/sc/icosahedron),
 *                  the reference's own bond-order.cpp built into oracle/_ref
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time.  Radix 4 first, then any  */
/* prime factor (small DFT by direct sum).  Unnormalised, sign = -1 fwd.    */
/* Restates the transform cuFFT computes at cuda/cufft_wrapper.hpp:50-58.   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    int nf;
    int f[64];
    cplx* w; /* w[j] = exp(sign * 2 pi i j / n) */
} fft_plan;

static void plan_init(fft_plan* p, int n, int sign) {
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    for (int q = 2; m > 1; q++) {
        while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
    }
    p->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; j++) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        p->w[j] = cos(a) + (double)sign * I * sin(a);
    }
}

static void plan_free(fft_plan* p) { free(p->w); }

static void fft_rec(const cplx* x, long xs, cplx* y, int n, const int* f, const fft_plan* p, long wstride) {
    if (n == 1) { y[0] = x[0]; return; }
    const int r = f[0];
    const int m = n / r;
    for (int q = 0; q < r; q++) fft_rec(x + q * xs, xs * r, y + (long)q * m, m, f + 1, p, wstride * r);
    cplx t[64], s[64];
    cplx* tt = t; cplx* ss = s;
    cplx* heap = NULL;
    if (r > 64) { heap = (cplx*)malloc(sizeof(cplx) * 2 * (size_t)r); tt = heap; ss = heap + r; }
    const long N = p->n;
    const long wr = N / r; /* exp(sign 2 pi i / r) = w[N/r] */
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < r; q++) tt[q] = y[(long)q * m + k] * p->w[((long)q * k * wstride) % N];
        for (int s2 = 0; s2 < r; s2++) {
            cplx acc = 0;
            for (int q = 0; q < r; q++) acc += tt[q] * p->w[(((long)q * s2) % r) * wr];
            ss[s2] = acc;
        }
        for (int s2 = 0; s2 < r; s2++) y[k + (long)s2 * m] = ss[s2];
    }
    free(heap);
}

static void fft_line(const fft_plan* p, cplx* buf, cplx* tmp) {
    fft_rec(buf, 1, tmp, p->n, p->f, p, 1);
    memcpy(buf, tmp, sizeof(cplx) * (size_t)p->n);
}

int orc_fft_c2c(double* data, int n, int sign) {
    if (n < 1) return -1;
    fft_plan p; plan_init(&p, n, sign);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    fft_line(&p, (cplx*)data, tmp);
    free(tmp); plan_free(&p);
    return 0;
}

/* transform every line of a complex 3D array [nz][ny][nxc] along axis (1 = y, 2 = z) */
static void fft_axis(cplx* a, int nxc, int ny, int nz, int axis, int sign) {
    const int L = axis == 1 ? ny : nz;
    if (L == 1) return;
    const long stride = axis == 1 ? nxc : (long)nxc * ny;
    const long nlines = axis == 1 ? (long)nxc * nz : (long)nxc * ny;
    fft_plan p; plan_init(&p, L, sign);
#pragma omp parallel
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)L);
        cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)L);
#pragma omp for schedule(static)
        for (long line = 0; line < nlines; line++) {
            long base;
            if (axis == 1) { long i2 = line / nxc, i0 = line % nxc; base = i2 * (long)nxc * ny + i0; }
            else { base = line; }
            for (int j = 0; j < L; j++) buf[j] = a[base + j * stride];
            fft_line(&p, buf, tmp);
            for (int j = 0; j < L; j++) a[base + j * stride] = buf[j];
        }
        free(buf); free(tmp);
    }
    plan_free(&p);
}

int orc_r2c(const double* in, double* outd, const int n[3]) {
    const int nx = n[0], ny = n[1], nz = n[2];
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    const int nxc = nx / 2 + 1;
    cplx* out = (cplx*)outd;
    const long rows = (long)ny * nz;
    fft_plan p; plan_init(&p, nx, -1);
#pragma omp parallel
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
        cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
        for (long r = 0; r < rows; r++) {
            for (int j = 0; j < nx; j++) buf[j] = in[r * nx + j];
            fft_line(&p, buf, tmp);
            for (int j = 0; j < nxc; j++) out[r * nxc + j] = buf[j];
        }
        free(buf); free(tmp);
    }
    plan_free(&p);
    fft_axis(out, nxc, ny, nz, 1, -1);
    fft_axis(out, nxc, ny, nz, 2, -1);
    return 0;
}

int orc_c2r(const double* ind, double* out, const int n[3]) {
    const int nx = n[0], ny = n[1], nz = n[2];
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    const int nxc = nx / 2 + 1;
    const long rows = (long)ny * nz;
    cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)(nxc * rows));
    memcpy(a, ind, sizeof(cplx) * (size_t)(nxc * rows));
    fft_axis(a, nxc, ny, nz, 2, +1);
    fft_axis(a, nxc, ny, nz, 1, +1);
    fft_plan p; plan_init(&p, nx, +1);
#pragma omp parallel
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
        cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
        for (long r = 0; r < rows; r++) {
            const cplx* row = a + r * nxc;
            for (int j = 0; j < nxc; j++) buf[j] = row[j];
            for (int j = nxc; j < nx; j++) buf[j] = conj(row[nx - j]);
            /* numpy irfft semantics: imaginary parts of the self-conjugate x modes are ignored */
            buf[0] = creal(buf[0]);
            if (nx % 2 == 0) buf[nx / 2] = creal(buf[nx / 2]);
            fft_line(&p, buf, tmp);
            for (int j = 0; j < nx; j++) out[r * nx + j] = creal(buf[j]);
        }
        free(buf); free(tmp);
    }
    plan_free(&p);
    free(a);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11.  Replaces the cuRAND XORWOW generator of        */
/* cuda/noise.hpp:103-105,145-150 (that stream is not reproducible).        */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_stream(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int64_t n, uint32_t* out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const uint64_t c = (uint64_t)i + c0;
        const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)(c >> 32), c2, c3};
        orc_philox4x32_10(ctr, key, out + 4 * i);
    }
}

/* stream tags in ctr[3] */
#define TAG_KNOISE 0u
#define TAG_INIT 1u

/* uniform in (0,1) from the top 24 bits: exactly representable in float and double */
static inline double u01_24(uint32_t x) { return ((double)(x >> 8) + 0.5) * 5.9604644775390625e-08; }

/* raw unit complex Gaussian for half-spectrum index h: Box-Muller on two 32-bit uniforms */
static cplx raw_xi(uint64_t h, uint64_t seed, uint32_t step, double* z0_out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)h, (uint32_t)(h >> 32), step, TAG_KNOISE};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const double u1 = u01_24(o[0]), u2 = u01_24(o[1]);
    const double r = sqrt(-2.0 * log(u1));
    const double th = 2.0 * M_PI * u2;
    const double z0 = r * cos(th), z1 = r * sin(th);
    if (z0_out) *z0_out = z0;
    return (z0 + I * z1) * M_SQRT1_2;
}

/* Hermitian-consistent unit noise at half-spectrum mode (i0,i1,i2).  Restates the
   repair of cuda/noise.hpp:47-91 (self-conjugate modes real with unit variance, a
   mode and its mirror on the i0 in {0, nx/2} planes complex conjugates), with the
   canonical member (smaller (i2*ny+i1)) drawing and its mirror conjugating. */
static cplx noise_mode(int i0, int i1, int i2, const int n[3], uint64_t seed, uint32_t step) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const uint64_t h = ((uint64_t)i2 * ny + i1) * nxc + i0;
    if (i0 == 0 || (nx % 2 == 0 && i0 == nx / 2)) {
        const int m1 = (ny - i1) % ny, m2 = (nz - i2) % nz;
        if (m1 == i1 && m2 == i2) {
            double z0;
            raw_xi(h, seed, step, &z0);
            return z0;
        }
        const uint64_t lin = (uint64_t)i2 * ny + i1, mlin = (uint64_t)m2 * ny + m1;
        if (lin < mlin) return raw_xi(h, seed, step, NULL);
        return conj(raw_xi(mlin * nxc + i0, seed, step, NULL));
    }
    return raw_xi(h, seed, step, NULL);
}

/* Noise layout of the fused stepper (even nx, L = nx/2).  One Philox block at
   counter (c*L + p, step, TAG_KROW) of row c = i2*ny + i1 gives two unit complex
   Gaussians: a from (o0, o1) for the full x mode p of row c, b from (o2, o3) for
   the full x mode p + L.  Of a mirror pair of rows (c, m = -c), the smaller index
   draws and the other takes complex conjugates, since full mode (k, m) is the
   Hermitian partner of (2L - k, c).  In a self-mirrored row (c == m) mode k and
   2L - k are partners: 0 < k < L draws a at p = k, the modes 0 and L are real
   (the first normal of a and of b at p = 0).  Same statistics as noise_mode. */
#define TAG_KROW 3u

static void xi_block(uint64_t c, uint64_t seed, uint32_t step, cplx* a, cplx* b, double* za, double* zb) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)(c >> 32), step, TAG_KROW};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    for (int j = 0; j < 2; j++) {
        const double u1 = u01_24(o[2 * j]), u2 = u01_24(o[2 * j + 1]);
        const double r = sqrt(-2.0 * log(u1)), th = 2.0 * M_PI * u2;
        const double z0 = r * cos(th), z1 = r * sin(th);
        if (j == 0) { *a = (z0 + I * z1) * M_SQRT1_2; *za = z0; }
        else { *b = (z0 + I * z1) * M_SQRT1_2; *zb = z0; }
    }
}

static cplx row_noise_mode(int i0, int i1, int i2, const int n[3], uint64_t seed, uint32_t step) {
    const int nx = n[0], ny = n[1], nz = n[2], L = nx / 2;
    const uint64_t r = (uint64_t)i2 * ny + i1;
    const uint64_t m = (uint64_t)((nz - i2) % nz) * ny + (uint64_t)((ny - i1) % ny);
    cplx a, b;
    double za, zb;
    if (r == m) {
        if (i0 == 0 || i0 == L) {
            xi_block(r * L, seed, step, &a, &b, &za, &zb);
            return i0 == 0 ? za : zb;
        }
        xi_block(r * L + i0, seed, step, &a, &b, &za, &zb);
        return a;
    }
    if (r < m) {
        xi_block(r * L + (i0 == L ? 0 : i0), seed, step, &a, &b, &za, &zb);
        return i0 == L ? b : a;
    }
    if (i0 == 0) {
        xi_block(m * L, seed, step, &a, &b, &za, &zb);
        return conj(a);
    }
    xi_block(m * L + (L - i0), seed, step, &a, &b, &za, &zb);
    return conj(b);
}

void orc_stepper_noise(double* xid, const int n[3], uint64_t seed, uint32_t step) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    cplx* xi = (cplx*)xid;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++)
                xi[((long)i2 * ny + i1) * nxc + i0] = row_noise_mode(i0, i1, i2, n, seed, step);
}

void orc_kspace_noise(double* xid, const int n[3], uint64_t seed, uint32_t step) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    cplx* xi = (cplx*)xid;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++)
                xi[((long)i2 * ny + i1) * nxc + i0] = noise_mode(i0, i1, i2, n, seed, step);
}

void orc_init_field(double* psi, const int n[3], double psi0, double amp, uint64_t seed) {
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; c++) {
        const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)((uint64_t)c >> 32), 0u, TAG_INIT};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        /* 53-bit uniform in [0,1): exact in double */
        const double u = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
        const double t = 2.0 * u - 1.0;
        const double s = amp * t;
        psi[c] = psi0 + s;
    }
}

/* ------------------------------------------------------------------------ */
/* PFC update.  Restates cuda/kernels/eom.hpp:60-72 (reference split) and   */
/* the cubic split: N = psi^3, D = 1 + dt k^2 (-eps + (1-k^2)^2).       */
/* ------------------------------------------------------------------------ */

static inline double kval(int i, int nn, double len) {
    return (double)(i < nn / 2 + 1 ? i : i - nn) * (2.0 * M_PI) / len;
}

void orc_update(double* kpsid, const double* knld, const double* xid, const int n[3], const double lens[3],
                double dt, double eps, double noise_amp, int split) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    const cplx* xi = (const cplx*)xid;
    const double lin0 = split == ORC_SPLIT_REFERENCE ? 1.0 - eps : -eps;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++) {
            const double k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]);
                const double ksq = k0 * k0 + k1 * k1 + k2 * k2;
                const double om = 1.0 - ksq;
                const double D = 1.0 + dt * ksq * (lin0 + om * om);
                cplx num = kpsi[idx] - dt * ksq * knl[idx];
                if (xi) num += noise_amp * sqrt(ksq) * xi[idx];
                kpsi[idx] = num / D;
            }
        }
}

/* Keeps the half spectrum Hermitian on the self-conjugate x planes (i0 = 0, nx/2).
   c2r ignores the anti-Hermitian part there, so the nonlinearity never sees it,
   and round-off in it would grow at the linear rate of the unstable band; the
   projection is what a round trip through the real field does. */
static void hermitian_planes(cplx* k, const int n[3]) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const int planes[2] = {0, nx / 2};
    const int np = (nx % 2 == 0 && nx > 1) ? 2 : 1;
    for (int pi = 0; pi < np; pi++) {
        const int i0 = planes[pi];
        for (int i2 = 0; i2 < nz; i2++)
            for (int i1 = 0; i1 < ny; i1++) {
                const int m1 = (ny - i1) % ny, m2 = (nz - i2) % nz;
                const long a = ((long)i2 * ny + i1) * nxc + i0, b = ((long)m2 * ny + m1) * nxc + i0;
                if (a > b) continue;
                const cplx v = 0.5 * (k[a] + conj(k[b]));
                k[a] = v;
                k[b] = conj(v);
            }
    }
}

/* cuda/kernels/eom.hpp:88-101 (kernel_update_eom_with_fen), in FP64: the reference split's
   update written with lk = (2 - eps) + 2 dop.x + dop.y and denom = 1 - dt dop.x lk (dop.x = -k^2,
   dop.y = k^4), noise namp*sqrt(N)*|k|*xi, and knoise <- lk * kpsi (the kpsi before the update). */
void orc_update_eom_fen(double* kpsid, const double* knld, double* knoised, const int n[3], const double lens[3],
                        double namp, double dt, double eps) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const double na = namp * sqrt((double)nx * ny * nz);
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    cplx* knoise = (cplx*)knoised;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double dx = -(k0 * k0 + k1 * k1 + k2 * k2), dy = dx * dx;
                const double lk = (2.0 - eps) + 2.0 * dx + dy;
                const double denom = 1.0 - dt * dx * lk;
                const cplx p = kpsi[idx];
                kpsi[idx] = (p + dt * dx * knl[idx] + dt * na * sqrt(-dx) * knoise[idx]) / denom;
                knoise[idx] = lk * p;
            }
}

/* cuda/kernels/eom.hpp:111-141 (kernel_calc_fen): f = 0.5 psi (lpsi - psi + 0.5 psi^3) per cell,
   written into lpsi; returns h0 h1 h2 * sum f. */
double orc_calc_fen(const double* psi, double* lpsi, const int n[3], const double h[3]) {
    const long N = (long)n[0] * n[1] * n[2];
    double sum = 0;
    for (long c = 0; c < N; c++) {
        const double p = psi[c];
        lpsi[c] = 0.5 * p * (lpsi[c] - p + 0.5 * p * p * p);
        sum += lpsi[c];
    }
    return sum * h[0] * h[1] * h[2];
}

/* cuda/kernels/el.hpp:3-45 (kernel_update_el): non-conserved relaxation
   kpsi' = (kpsi - dt knl) / (1 + dt ((2 - eps) + 2 dop.x + dop.y)); returns max |kpsi' - kpsi|. */
double orc_update_el(double* kpsid, const double* knld, const int n[3], const double lens[3], double dt, double eps) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    double mx = 0;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double dx = -(k0 * k0 + k1 * k1 + k2 * k2), dy = dx * dx;
                const double denom = 1.0 + dt * ((2.0 - eps) + 2.0 * dx + dy);
                const cplx pn = (kpsi[idx] - dt * knl[idx]) / denom;
                const double d = cabs(kpsi[idx] - pn);
                if (d > mx) mx = d;
                kpsi[idx] = pn;
            }
    return mx;
}

/* cuda/kernels/eom-exp.hpp:3-10 and el-exp.hpp:3-9: the ETD tables (E, S) per half-spectrum mode.
   conserved (el = 0): linop = k^2 L', S = (E - 1) / linop or -dt where that is not a number;
   non-conserved (el = 1): linop = L'; L' = (2 - eps) + 2 dop.x + dop.y.  E = expf(-linop dt): the
   reference takes the single-precision exponential in both builds. */
void orc_etd_op(double* kopd, const int n[3], const double lens[3], double dt, double eps, int el) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double dx = -(k0 * k0 + k1 * k1 + k2 * k2), dy = dx * dx;
                const double lp = (2.0 - eps) + 2.0 * dx + dy;
                const double linop = el ? lp : -dx * lp;
                const double e = (double)expf((float)(-linop * dt));
                double second = (e - 1.0) / linop;
                if (second != second) second = -dt;
                kopd[2 * idx] = e;
                kopd[2 * idx + 1] = second;
            }
}

/* eom-exp.hpp:18-34: kpsi <- kpsi E - dop.x knl S - namp sqrt(N) |k| knoise S (knoise may be NULL) */
void orc_update_eom_etd(const double* kopd, double* kpsid, const double* knld, const double* knoised, const int n[3],
                        const double lens[3], double namp) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const double na = namp * sqrt((double)nx * ny * nz);
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    const cplx* knoise = (const cplx*)knoised;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double dx = -(k0 * k0 + k1 * k1 + k2 * k2);
                const double e = kopd[2 * idx], sv = kopd[2 * idx + 1];
                cplx v = kpsi[idx] * e - dx * knl[idx] * sv;
                if (knoise) v -= na * sqrt(-dx) * knoise[idx] * sv;
                kpsi[idx] = v;
            }
}

/* el-exp.hpp:17-34: kpsi <- kpsi E + knl S; returns max |change| */
double orc_update_el_etd(const double* kopd, double* kpsid, const double* knld, const int n[3]) {
    const long nc = (long)(n[0] / 2 + 1) * n[1] * n[2];
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    double mx = 0;
    for (long idx = 0; idx < nc; idx++) {
        const cplx pn = kpsi[idx] * kopd[2 * idx] + knl[idx] * kopd[2 * idx + 1];
        const double d = cabs(kpsi[idx] - pn);
        if (d > mx) mx = d;
        kpsi[idx] = pn;
    }
    return mx;
}

/* eom.hpp:16-46 (2D): semi-implicit update with noise (c1 kx^2 + c2 ky^2) xi1 + (c2 kx^2 + c1 ky^2) xi2
   + 2 kx ky xi3 for |k| <= 1/4, kappa = 4/3, c2 = sqrt(1 + kappa/2 + sqrt(1 + kappa)), c1 = kappa/(2 c2) */
void orc_update_eom_hpfc(double* kpsid, const double* knld, const double* x1d, const double* x2d, const double* x3d,
                         const int n[3], const double lens[3], double namp, double dt, double eps) {
    const int nx = n[0], ny = n[1], nxc = nx / 2 + 1;
    const double na = namp * sqrt((double)nx * ny);
    const double kappa = 4.0 / 3.0, c2 = sqrt(1.0 + kappa / 2.0 + sqrt(1.0 + kappa)), c1 = kappa / c2 / 2.0;
    cplx* kpsi = (cplx*)kpsid;
    const cplx *knl = (const cplx*)knld, *x1 = (const cplx*)x1d, *x2 = (const cplx*)x2d, *x3 = (const cplx*)x3d;
    for (int i1 = 0; i1 < ny; i1++)
        for (int i0 = 0; i0 < nxc; i0++) {
            const long idx = (long)i1 * nxc + i0;
            const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]);
            const double dx = -(k0 * k0 + k1 * k1), dy = dx * dx;
            const double denom = 1.0 - dt * ((2.0 - eps) * dx + 2.0 * dy + dx * dy);
            cplx noise = (c1 * k0 * k0 + c2 * k1 * k1) * x1[idx] + (c2 * k0 * k0 + c1 * k1 * k1) * x2[idx] +
                         2.0 * k0 * k1 * x3[idx];
            if (!(sqrt(-dx) <= 0.25)) noise = 0;
            kpsi[idx] = (kpsi[idx] + dt * dx * knl[idx] + dt * na * noise) / denom;
        }
}

/* eom.hpp:143-149: the fitted coefficients c_0, c_2, ..., c_12 of P (data of that model) */
static const double poly_ref[7] = {-57.94636917114257, 164.4074088310241, -181.1481995196213, 100.6942650417575,
                                   -30.20985042718309, 4.678065276755401, -0.293423255328944};

void orc_poly_coefficients(double out[7]) {
    for (int j = 0; j < 7; j++) out[j] = poly_ref[j];
}

/* eom.hpp:151-171: operator 1 - eps - P(k^2), P(x) = sum_j c[j] x^j (coef[0..ncoef), NULL: the
   reference's fitted coefficients), denom = 1 - dt dop.x (1 - eps - P); knoise may be NULL */
void orc_update_eom_poly(double* kpsid, const double* knld, const double* knoised, const int n[3],
                         const double lens[3], double namp, double dt, double eps, const double* coef, int ncoef) {
    const double* c = coef ? coef : poly_ref;
    const int nc = coef ? ncoef : 7;
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const double na = namp * sqrt((double)nx * ny * nz);
    cplx* kpsi = (cplx*)kpsid;
    const cplx* knl = (const cplx*)knld;
    const cplx* knoise = (const cplx*)knoised;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const long idx = ((long)i2 * ny + i1) * nxc + i0;
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double ksq = k0 * k0 + k1 * k1 + k2 * k2;
                double p = 0, xp = 1;
                for (int j = 0; j < nc; j++) { p += c[j] * xp; xp *= ksq; }
                const double denom = 1.0 + dt * ksq * (1.0 - eps - p);
                cplx v = kpsi[idx] - dt * ksq * knl[idx];
                if (knoise) v += dt * na * sqrt(ksq) * knoise[idx];
                kpsi[idx] = v / denom;
            }
}

int orc_pfc_run(double* psi, const int n[3], const double h[3], double eps, double dt, double temp,
                uint64_t seed, int step0, int nsteps, int split, double* mass_hist) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const long N = (long)nx * ny * nz, Nc = (long)nxc * ny * nz;
    const double lens[3] = {nx * h[0], ny * h[1], nz * h[2]};
    const double dV = h[0] * (ny > 1 ? h[1] : 1.0) * (nz > 1 ? h[2] : 1.0);
    const double namp = temp > 0 ? sqrt(2.0 * temp * dt / dV) * sqrt((double)N) : 0.0;
    double* kpsi = (double*)malloc(sizeof(double) * 2 * (size_t)Nc);
    double* knl = (double*)malloc(sizeof(double) * 2 * (size_t)Nc);
    double* xi = temp > 0 ? (double*)malloc(sizeof(double) * 2 * (size_t)Nc) : NULL;
    double* nl = (double*)malloc(sizeof(double) * (size_t)N);
    orc_r2c(psi, kpsi, n);
    if (mass_hist) mass_hist[0] = kpsi[0] / (double)N;
    for (int s = 0; s < nsteps; s++) {
        orc_c2r(kpsi, nl, n);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < N; c++) {
            const double v = nl[c] / (double)N;
            nl[c] = split == ORC_SPLIT_REFERENCE ? v * v * v - v : v * v * v;
        }
        orc_r2c(nl, knl, n);
        if (xi) {
            if (nx % 2 == 0) orc_stepper_noise(xi, n, seed, (uint32_t)(step0 + s));
            else orc_kspace_noise(xi, n, seed, (uint32_t)(step0 + s));
        }
        orc_update(kpsi, knl, xi, n, lens, dt, eps, namp, split);
        hermitian_planes((cplx*)kpsi, n);
        if (mass_hist) mass_hist[s + 1] = kpsi[0] / (double)N;
    }
    orc_c2r(kpsi, psi, n);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; c++) psi[c] /= (double)N;
    free(kpsi); free(knl); free(xi); free(nl);
    return 0;
}

double orc_free_energy(const double* psi, const int n[3], const double h[3], double eps) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const long N = (long)nx * ny * nz, Nc = (long)nxc * ny * nz;
    const double lens[3] = {nx * h[0], ny * h[1], nz * h[2]};
    double* k = (double*)malloc(sizeof(double) * 2 * (size_t)Nc);
    double* lp = (double*)malloc(sizeof(double) * (size_t)N);
    orc_r2c(psi, k, n);
    cplx* kc = (cplx*)k;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const double k0 = kval(i0, nx, lens[0]), k1 = kval(i1, ny, lens[1]), k2 = kval(i2, nz, lens[2]);
                const double ksq = k0 * k0 + k1 * k1 + k2 * k2;
                const double om = 1.0 - ksq;
                kc[((long)i2 * ny + i1) * nxc + i0] *= (-eps + om * om) / (double)N;
            }
    orc_c2r(k, lp, n);
    double sum = 0.0;
    for (long c = 0; c < N; c++) {
        const double v = psi[c];
        sum += 0.5 * v * lp[c] + 0.25 * v * v * v * v;
    }
    free(k); free(lp);
    const double dV = h[0] * (ny > 1 ? h[1] : 1.0) * (nz > 1 ? h[2] : 1.0);
    return sum * dV;
}

/* ------------------------------------------------------------------------ */
/* Peaks: restates cuda/kernels/findpeaks.hpp:7-29 (26 periodic neighbours, */
/* strict >, v >= ths) and findpeaks_2d.hpp:7-27 (8 neighbours).  Output is */
/* ordered by linear index (the reference's atomic order is arbitrary).     */
/* ------------------------------------------------------------------------ */

static inline int wrapi(int i, int nn) { return i < 0 ? i + nn : (i >= nn ? i - nn : i); }

int64_t orc_findpeaks(const double* v, const int n[3], double ths, int64_t* idx_out, int64_t cap) {
    const int nx = n[0], ny = n[1], nz = n[2];
    const int is3d = nz > 1;
    int64_t cnt = 0;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nx; i0++) {
                const int64_t c = ((int64_t)i2 * ny + i1) * nx + i0;
                const double lv = v[c];
                int peak = 1;
                for (int d0 = -1; d0 <= 1 && peak; d0++)
                    for (int d1 = -1; d1 <= 1 && peak; d1++)
                        for (int d2 = is3d ? -1 : 0; d2 <= (is3d ? 1 : 0) && peak; d2++) {
                            if (d0 == 0 && d1 == 0 && d2 == 0) continue;
                            const int64_t o = ((int64_t)wrapi(i2 + d2, nz) * ny + wrapi(i1 + d1, ny)) * nx + wrapi(i0 + d0, nx);
                            if (!(lv > v[o])) peak = 0;
                        }
                if (peak && lv >= ths) {
                    if (cnt < cap) idx_out[cnt] = c;
                    cnt++;
                }
            }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Bond order.                                                              */
/* ------------------------------------------------------------------------ */

static double lfact(int k) { double r = 1; for (int i = 2; i <= k; i++) r *= i; return r; }

void orc_ylm(int l, int m, double ct, double phi, double* re, double* im) {
    const int am = m < 0 ? -m : m;
    /* associated Legendre P_l^am(ct) with Condon-Shortley phase, by upward recurrence */
    const double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    double pmm = 1.0;
    for (int i = 1; i <= am; i++) pmm *= -(2.0 * i - 1.0) * st;
    double plm = pmm;
    if (l > am) {
        double pm1 = ct * (2.0 * am + 1.0) * pmm;
        double pm2 = pmm;
        plm = pm1;
        for (int ll = am + 2; ll <= l; ll++) {
            const double pl = ((2.0 * ll - 1.0) * ct * pm1 - (ll + am - 1.0) * pm2) / (double)(ll - am);
            pm2 = pm1; pm1 = pl; plm = pl;
        }
    }
    const double nrm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * lfact(l - am) / lfact(l + am));
    double r = nrm * plm * cos(am * phi), i = nrm * plm * sin(am * phi);
    if (m < 0) { /* Y_{l,-m} = (-1)^m conj(Y_lm) */
        const double sg = (am % 2) ? -1.0 : 1.0;
        r = sg * r; i = -sg * i;
    }
    *re = r; *im = i;
}

double orc_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    /* Racah formula */
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    const double tri = lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) * lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
    const double pre = sqrt(tri * lfact(j1 + m1) * lfact(j1 - m1) * lfact(j2 + m2) * lfact(j2 - m2) *
                            lfact(j3 + m3) * lfact(j3 - m3));
    double sum = 0.0;
    for (int k = 0; k <= j1 + j2 + j3; k++) {
        const int a1 = j3 - j2 + k + m1, a2 = j3 - j1 + k - m2, a3 = j1 + j2 - j3 - k, a4 = j1 - k - m1, a5 = j2 - k + m2;
        if (a1 < 0 || a2 < 0 || a3 < 0 || a4 < 0 || a5 < 0) continue;
        const double den = lfact(k) * lfact(a1) * lfact(a2) * lfact(a3) * lfact(a4) * lfact(a5);
        sum += ((k % 2) ? -1.0 : 1.0) / den;
    }
    const int ph = j1 - j2 - m3;
    return (((ph % 2) + 2) % 2 ? -1.0 : 1.0) * pre * sum;
}

/* minimum-image difference, restating bond-order-analysis/bond-order.cpp:53-64 (tie order kept) */
static inline double get_diff(double x1, double x2, double w) {
    const double d1 = x1 - x2, d2 = d1 - w, d3 = d1 + w;
    if (fabs(d1) < fabs(d2)) return fabs(d1) < fabs(d3) ? d1 : d3;
    return fabs(d2) < fabs(d3) ? d2 : d3;
}

/* near-pole rule of bond-order.cpp:123-133: m<0 terms vanish within 1e-3 of ct=-1, m>0 within 1e-3 of ct=+1 */
static void ylm_ref(int l, int m, double ct, double phi, cplx* y) {
    if ((m < 0 && fabs(1.0 + ct) < 1e-3) || (m > 0 && fabs(1.0 - ct) < 1e-3)) { *y = 0; return; }
    double re, im;
    orc_ylm(l, m, ct, phi, &re, &im);
    *y = re + I * im;
}

static double ql_of(const cplx* q, int l) {
    double s = 0;
    for (int m = 0; m <= 2 * l; m++) { const double a = cabs(q[m]); s += a * a; }
    return sqrt(4.0 * M_PI / (2.0 * l + 1.0) * s);
}

static double w4_of(const cplx* q4) {
    /* restates bond-order.cpp:193-203 with the 3j weights computed, not tabulated */
    cplx acc = 0;
    for (int m1 = -4; m1 <= 4; m1++)
        for (int m2 = -4; m2 <= 4; m2++) {
            const int m3 = -m1 - m2;
            if (m3 < -4 || m3 > 4) continue;
            acc += orc_wigner3j(4, 4, 4, m1, m2, m3) * q4[m1 + 4] * q4[m2 + 4] * q4[m3 + 4];
        }
    double s = 0;
    for (int m = 0; m < 9; m++) { const double a = cabs(q4[m]); s += a * a; }
    return creal(acc) / sqrt(s * s * s);
}

typedef struct { int cnt, cap; int* j; double* ct; double* ph; } nblist;

static void nb_push(nblist* L, int j, double ct, double ph) {
    if (L->cnt == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 16;
        L->j = (int*)realloc(L->j, sizeof(int) * (size_t)L->cap);
        L->ct = (double*)realloc(L->ct, sizeof(double) * (size_t)L->cap);
        L->ph = (double*)realloc(L->ph, sizeof(double) * (size_t)L->cap);
    }
    L->j[L->cnt] = j; L->ct[L->cnt] = ct; L->ph[L->cnt] = ph; L->cnt++;
}

void orc_bond_order_cutoff(const double* xyz, int P, double box, double rlim, int* nb, double* out) {
    nblist* L = (nblist*)calloc((size_t)(P > 0 ? P : 1), sizeof(nblist));
    /* per point: q4m (9) | q6m (13) | q8m (17), padded to 40 */
    cplx* q = (cplx*)calloc((size_t)(P > 0 ? P : 1) * 40, sizeof(cplx));
    /* find_nbors: bond-order.cpp:68-111 (j ascending, r <= cutoff) */
#pragma omp parallel for schedule(dynamic, 2)
    for (int i = 0; i < P; i++) {
        for (int j = 0; j < P; j++) {
            if (i == j) continue;
            const double dx = get_diff(xyz[3 * j], xyz[3 * i], box);
            const double dy = get_diff(xyz[3 * j + 1], xyz[3 * i + 1], box);
            const double dz = get_diff(xyz[3 * j + 2], xyz[3 * i + 2], box);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= rlim) nb_push(&L[i], j, dz / r, atan2(dy, dx));
        }
    }
    /* q_lm: bond-order.cpp:252-265 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        cplx* qi = q + 40 * (long)i;
        if (L[i].cnt == 0) continue;
        for (int t = 0; t < L[i].cnt; t++) {
            cplx y;
            for (int m = -4; m <= 4; m++) { ylm_ref(4, m, L[i].ct[t], L[i].ph[t], &y); qi[m + 4] += y; }
            for (int m = -6; m <= 6; m++) { ylm_ref(6, m, L[i].ct[t], L[i].ph[t], &y); qi[9 + m + 6] += y; }
            for (int m = -8; m <= 8; m++) { ylm_ref(8, m, L[i].ct[t], L[i].ph[t], &y); qi[22 + m + 8] += y; }
        }
        for (int m = 0; m < 39; m++) qi[m] /= (double)L[i].cnt;
    }
    /* q_l, w4, averaged q_l (bond-order.cpp:177-250), xi (267-287) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        double* o = out + 9 * (long)i;
        nb[i] = L[i].cnt;
        if (L[i].cnt == 0) { for (int t = 0; t < 9; t++) o[t] = NAN; continue; }
        const cplx* qi = q + 40 * (long)i;
        o[0] = ql_of(qi, 4); o[1] = ql_of(qi + 9, 6); o[2] = ql_of(qi + 22, 8);
        o[3] = w4_of(qi);
        const int offs[3] = {0, 9, 22}, ls[3] = {4, 6, 8};
        for (int t = 0; t < 3; t++) {
            cplx bar[17];
            for (int m = 0; m <= 2 * ls[t]; m++) {
                cplx s = qi[offs[t] + m];
                for (int u = 0; u < L[i].cnt; u++) s += q[40 * (long)L[i].j[u] + offs[t] + m];
                bar[m] = s / (1.0 + (double)L[i].cnt);
            }
            o[4 + t] = ql_of(bar, ls[t]);
        }
        /* bond-order.cpp:205-219 assigns (not accumulates) each neighbour's q4m, so the
           averaged vector is the last (largest-index) neighbour's q4m over the count */
        {
            cplx bar[9];
            const long last = L[i].j[L[i].cnt - 1];
            for (int m = 0; m < 9; m++) bar[m] = q[40 * last + m] / (double)L[i].cnt;
            o[7] = w4_of(bar);
        }
        double den_i = 0;
        for (int m = 0; m < 13; m++) { const double a = cabs(qi[9 + m]); den_i += a * a; }
        den_i = sqrt(den_i);
        int s = 0;
        for (int u = 0; u < L[i].cnt; u++) {
            const cplx* qj = q + 40 * (long)L[i].j[u] + 9;
            cplx d6 = 0; double den_j = 0;
            for (int m = 0; m < 13; m++) { d6 += qi[9 + m] * conj(qj[m]); const double a = cabs(qj[m]); den_j += a * a; }
            den_j = sqrt(den_j);
            if (creal(d6) / (den_i * den_j) > 0.7) s++;
        }
        o[8] = s;
    }
    for (int i = 0; i < P; i++) { free(L[i].j); free(L[i].ct); free(L[i].ph); }
    free(L); free(q);
}

static inline int min_image(int d, int n) {
    /* integer minimum image into (-n/2, n/2] */
    d %= n;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

/* the k nearest neighbours of point i by an all-pairs scan and its q4 q6 q8 w4 (the
   12-NN form asked for beside bond-order.cpp's cutoff form); bestd/bestj: k slots */
static void knn_point(const int* ijk, int P, int n, double h, int k, int i, long* bestd, int* bestj, int* nbr,
                      double* o) {
    int cnt = 0;
    for (int j = 0; j < P; j++) {
        if (j == i) continue;
        const long dx = min_image(ijk[3 * j] - ijk[3 * i], n);
        const long dy = min_image(ijk[3 * j + 1] - ijk[3 * i + 1], n);
        const long dz = min_image(ijk[3 * j + 2] - ijk[3 * i + 2], n);
        const long d2 = dx * dx + dy * dy + dz * dz;
        /* insert (d2, j) keeping the k smallest, ordered by (d2, j) */
        if (cnt == k && !(d2 < bestd[k - 1] || (d2 == bestd[k - 1] && j < bestj[k - 1]))) continue;
        int pos = cnt < k ? cnt : k - 1;
        while (pos > 0 && (bestd[pos - 1] > d2 || (bestd[pos - 1] == d2 && bestj[pos - 1] > j))) {
            bestd[pos] = bestd[pos - 1]; bestj[pos] = bestj[pos - 1]; pos--;
        }
        bestd[pos] = d2; bestj[pos] = j;
        if (cnt < k) cnt++;
    }
    for (int t = 0; t < k; t++) nbr[t] = t < cnt ? bestj[t] : -1;
    cplx q4[9] = {0}, q6[13] = {0}, q8[17] = {0};
    for (int t = 0; t < cnt; t++) {
        const int j = bestj[t];
        const double dx = h * min_image(ijk[3 * j] - ijk[3 * i], n);
        const double dy = h * min_image(ijk[3 * j + 1] - ijk[3 * i + 1], n);
        const double dz = h * min_image(ijk[3 * j + 2] - ijk[3 * i + 2], n);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double ct = dz / r, ph = atan2(dy, dx);
        double re, im;
        for (int m = -4; m <= 4; m++) { orc_ylm(4, m, ct, ph, &re, &im); q4[m + 4] += re + I * im; }
        for (int m = -6; m <= 6; m++) { orc_ylm(6, m, ct, ph, &re, &im); q6[m + 6] += re + I * im; }
        for (int m = -8; m <= 8; m++) { orc_ylm(8, m, ct, ph, &re, &im); q8[m + 8] += re + I * im; }
    }
    if (cnt == 0) { o[0] = o[1] = o[2] = o[3] = NAN; return; }
    for (int m = 0; m < 9; m++) q4[m] /= cnt;
    for (int m = 0; m < 13; m++) q6[m] /= cnt;
    for (int m = 0; m < 17; m++) q8[m] /= cnt;
    o[0] = ql_of(q4, 4); o[1] = ql_of(q6, 6); o[2] = ql_of(q8, 8); o[3] = w4_of(q4);
}

/* queries qi[0..nq) (or every point when qi is NULL, nq = P): row q of nbr/out for point qi[q] */
void orc_bond_order_knn_points(const int* ijk, int P, int n, double h, int k, const int* qi, int nq, int* nbr,
                               double* out) {
#pragma omp parallel
    {
        long* bestd = (long*)malloc(sizeof(long) * (size_t)(k > 0 ? k : 1));
        int* bestj = (int*)malloc(sizeof(int) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 16)
        for (int q = 0; q < nq; q++)
            knn_point(ijk, P, n, h, k, qi ? qi[q] : q, bestd, bestj, nbr + (long)k * q, out + 4 * (long)q);
        free(bestd); free(bestj);
    }
}

void orc_bond_order_knn(const int* ijk, int P, int n, double h, int k, int* nbr, double* out) {
    orc_bond_order_knn_points(ijk, P, n, h, k, NULL, P, nbr, out);
}

/* ------------------------------------------------------------------------ */
/* Post-processing.                                                          */
/* ------------------------------------------------------------------------ */

/* Gaussian filter table, restating cuda/kernels/gaussian.hpp:23-33:
   e = exp(-(sx^2 kx^2 + sy^2 ky^2 + sz^2 kz^2)/2), s_i = 1.1 divs h_i, kg = (e, 1 - e) */
void orc_gaussian_init(double* kgd, const int n[3], const double lens[3], const double h[3], double divs) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const double s[3] = {1.1 * divs * h[0], 1.1 * divs * h[1], 1.1 * divs * h[2]};
    cplx* kg = (cplx*)kgd;
    for (int i2 = 0; i2 < nz; i2++)
        for (int i1 = 0; i1 < ny; i1++)
            for (int i0 = 0; i0 < nxc; i0++) {
                const double a = s[0] * kval(i0, nx, lens[0]), b = s[1] * kval(i1, ny, lens[1]),
                             c = s[2] * kval(i2, nz, lens[2]);
                const double e = exp(-0.5 * (a * a + b * b + c * c));
                kg[((long)i2 * ny + i1) * nxc + i0] = e + I * (1.0 - e);
            }
}

/* minimum image of x2 - x1 in a periodic box of width w, restating particlepairs/main.cpp:8-15 */
static inline double image_diff(double x2, double x1, double w) {
    const double dx = x2 - x1;
    if (fabs(dx) > w / 2) return dx > 0 ? dx - w : dx + w;
    return dx;
}

/* all ordered pairs i != j, i-major (particlepairs/main.cpp:41-50); returns P (P - 1) */
int64_t orc_pair_distances(const double* xyz, int P, double box, double* out) {
    int64_t t = 0;
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) {
            if (i == j) continue;
            const double dx = image_diff(xyz[3 * j], xyz[3 * i], box);
            const double dy = image_diff(xyz[3 * j + 1], xyz[3 * i + 1], box);
            const double dz = image_diff(xyz[3 * j + 2], xyz[3 * i + 2], box);
            const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
            out[t++] = sqrt(dx2 + dy2 + dz2);
        }
    return t;
}
