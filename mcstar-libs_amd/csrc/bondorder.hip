// Steinhardt bond-order analysis on the GPU (FP64).
//
// Cutoff mode restates bond_order_analysis<T> (bond-order-analysis/
// bond-order.cpp:289-349): neighbours j != i with |d| <= rlim under the
// minimum image of get_diff (53-64), q_lm averaged over the neighbours
// (252-265), q_l (177-191), w4 (193-203), neighbour-averaged q_l (221-250),
// the averaged w4 as that code computes it (205-219: the running vector is
// overwritten by each neighbour, leaving the last neighbour's q4m) and xi
// (267-287).  Y_lm terms with m < 0 vanish within 1e-3 of cos(theta) = -1 and
// with m > 0 within 1e-3 of +1 (123-133).
//
// k-nearest mode: integer squared distances on the periodic grid, ties broken
// by the smaller index, exact Y_lm, q4 q6 q8 w4 and histograms of q4, q6.
// One thread per point scans all points (pair distances are computed with
// contraction off so neighbour sets match the CPU bit for bit).
#include "mcs_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

struct Wig4 { double w[81]; };  // (4 4 4; m1 m2 -m1-m2), index (m1+4)*9 + (m2+4)

MCS_DEV double dfact(int k) { double r = 1; for (int i = 2; i <= k; ++i) r *= i; return r; }

// normalised Y_lm(ct, phi) with Condon-Shortley phase, l <= 8
MCS_DEV void ylm_dev(int l, int m, double ct, double phi, double* re, double* im) {
    const int am = m < 0 ? -m : m;
    const double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    double pmm = 1.0;
    for (int i = 1; i <= am; ++i) pmm *= -(2.0 * i - 1.0) * st;
    double plm = pmm;
    if (l > am) {
        double pm1 = ct * (2.0 * am + 1.0) * pmm, pm2 = pmm;
        plm = pm1;
        for (int ll = am + 2; ll <= l; ++ll) {
            const double pl = ((2.0 * ll - 1.0) * ct * pm1 - (ll + am - 1.0) * pm2) / (double)(ll - am);
            pm2 = pm1; pm1 = pl; plm = pl;
        }
    }
    const double nrm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * dfact(l - am) / dfact(l + am));
    double s, c;
    sincos(am * phi, &s, &c);
    double r = nrm * plm * c, i = nrm * plm * s;
    if (m < 0) {
        const double sg = (am & 1) ? -1.0 : 1.0;
        r = sg * r; i = -sg * i;
    }
    *re = r; *im = i;
}

// accumulate the 39 (l = 4, 6, 8) harmonics of one bond into q[78] (re, im interleaved)
MCS_DEV void add_bond(double* q, double ct, double phi, bool pole_rule) {
    const int ls[3] = {4, 6, 8}, off[3] = {0, 9, 22};
    for (int t = 0; t < 3; ++t) {
        const int l = ls[t];
        for (int m = -l; m <= l; ++m) {
            if (pole_rule && ((m < 0 && fabs(1.0 + ct) < 1e-3) || (m > 0 && fabs(1.0 - ct) < 1e-3))) continue;
            double re, im;
            ylm_dev(l, m, ct, phi, &re, &im);
            q[2 * (off[t] + m + l)] += re;
            q[2 * (off[t] + m + l) + 1] += im;
        }
    }
}

MCS_DEV double ql_dev(const double* q, int l) {
    double s = 0;
    for (int m = 0; m <= 2 * l; ++m) s += q[2 * m] * q[2 * m] + q[2 * m + 1] * q[2 * m + 1];
    return sqrt(4.0 * M_PI / (2.0 * l + 1.0) * s);
}

MCS_DEV double w4_dev(const double* q4, const Wig4& W) {
    double acc = 0;
    for (int m1 = -4; m1 <= 4; ++m1)
        for (int m2 = -4; m2 <= 4; ++m2) {
            const int m3 = -m1 - m2;
            if (m3 < -4 || m3 > 4) continue;
            const double* a = q4 + 2 * (m1 + 4);
            const double* b = q4 + 2 * (m2 + 4);
            const double* c = q4 + 2 * (m3 + 4);
            const double abr = a[0] * b[0] - a[1] * b[1], abi = a[0] * b[1] + a[1] * b[0];
            acc += W.w[(m1 + 4) * 9 + (m2 + 4)] * (abr * c[0] - abi * c[1]);
        }
    double s = 0;
    for (int m = 0; m < 9; ++m) s += q4[2 * m] * q4[2 * m] + q4[2 * m + 1] * q4[2 * m + 1];
    return acc / sqrt(s * s * s);
}

MCS_DEV double image_diff(double x1, double x2, double w) {
#pragma clang fp contract(off)
    const double d1 = x1 - x2, d2 = d1 - w, d3 = d1 + w;
    if (fabs(d1) < fabs(d2)) return fabs(d1) < fabs(d3) ? d1 : d3;
    return fabs(d2) < fabs(d3) ? d2 : d3;
}

MCS_DEV bool bond(const double* xyz, int i, int j, double box, double rlim, double* ct, double* ph) {
#pragma clang fp contract(off)
    const double dx = image_diff(xyz[3 * j], xyz[3 * i], box);
    const double dy = image_diff(xyz[3 * j + 1], xyz[3 * i + 1], box);
    const double dz = image_diff(xyz[3 * j + 2], xyz[3 * i + 2], box);
    const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
    const double r = sqrt(xx + yy + zz);
    if (!(r <= rlim)) return false;
    *ct = dz / r;
    *ph = atan2(dy, dx);
    return true;
}

static constexpr int QS = 80;  // doubles per point: 39 complex + pad

__global__ void __launch_bounds__(128) k_bo_qlm(const double* xyz, int P, double box, double rlim, double* q,
                                                int* nb, int* last) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    double acc[78];
    for (int t = 0; t < 78; ++t) acc[t] = 0;
    int cnt = 0, lj = -1;
    for (int j = 0; j < P; ++j) {
        if (j == i) continue;
        double ct, ph;
        if (bond(xyz, i, j, box, rlim, &ct, &ph)) { add_bond(acc, ct, ph, true); ++cnt; lj = j; }
    }
    double* qi = q + (long)QS * i;
    for (int t = 0; t < 78; ++t) qi[t] = cnt ? acc[t] / (double)cnt : 0.0;
    nb[i] = cnt;
    last[i] = lj;
}

__global__ void __launch_bounds__(128) k_bo_stats(const double* xyz, int P, double box, double rlim, const double* q,
                                                  const int* nb, const int* last, Wig4 W, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    double* o = out + 9L * i;
    const int cnt = nb[i];
    if (cnt == 0) {
        for (int t = 0; t < 9; ++t) o[t] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    const double* qi = q + (long)QS * i;
    o[0] = ql_dev(qi, 4);
    o[1] = ql_dev(qi + 18, 6);
    o[2] = ql_dev(qi + 44, 8);
    o[3] = w4_dev(qi, W);
    double bar[78];
    for (int t = 0; t < 78; ++t) bar[t] = qi[t];
    double den_i = 0;
    for (int m = 0; m < 13; ++m) den_i += qi[18 + 2 * m] * qi[18 + 2 * m] + qi[19 + 2 * m] * qi[19 + 2 * m];
    den_i = sqrt(den_i);
    int xi = 0;
    for (int j = 0; j < P; ++j) {
        if (j == i) continue;
        double ct, ph;
        if (!bond(xyz, i, j, box, rlim, &ct, &ph)) continue;
        const double* qj = q + (long)QS * j;
        for (int t = 0; t < 78; ++t) bar[t] += qj[t];
        double dr = 0, den_j = 0;
        for (int m = 0; m < 13; ++m) {
            const double ar = qi[18 + 2 * m], ai = qi[19 + 2 * m], br = qj[18 + 2 * m], bi = qj[19 + 2 * m];
            dr += ar * br + ai * bi;
            den_j += br * br + bi * bi;
        }
        if (dr / (den_i * sqrt(den_j)) > 0.7) ++xi;
    }
    const double inv = 1.0 / (1.0 + (double)cnt);
    for (int t = 0; t < 78; ++t) bar[t] *= inv;
    o[4] = ql_dev(bar, 4);
    o[5] = ql_dev(bar + 18, 6);
    o[6] = ql_dev(bar + 44, 8);
    const double* ql = q + (long)QS * last[i];
    double b4[18];
    for (int t = 0; t < 18; ++t) b4[t] = ql[t] / (double)cnt;
    o[7] = w4_dev(b4, W);
    o[8] = (double)xi;
}

MCS_DEV int min_image(int d, int n) {
    d %= n;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

static constexpr int KMAX = 32;

__global__ void __launch_bounds__(128) k_bo_knn(const int* ijk, int P, int n, double h, int k, Wig4 W, int* nbr,
                                                double* out, unsigned long long* h4, unsigned long long* h6,
                                                int nbins) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    long bd[KMAX];
    int bj[KMAX];
    int cnt = 0;
    const int xi = ijk[3 * i], yi = ijk[3 * i + 1], zi = ijk[3 * i + 2];
    for (int j = 0; j < P; ++j) {
        if (j == i) continue;
        const long dx = min_image(ijk[3 * j] - xi, n), dy = min_image(ijk[3 * j + 1] - yi, n),
                   dz = min_image(ijk[3 * j + 2] - zi, n);
        const long d2 = dx * dx + dy * dy + dz * dz;
        if (cnt == k && !(d2 < bd[k - 1] || (d2 == bd[k - 1] && j < bj[k - 1]))) continue;
        int pos = cnt < k ? cnt : k - 1;
        while (pos > 0 && (bd[pos - 1] > d2 || (bd[pos - 1] == d2 && bj[pos - 1] > j))) {
            bd[pos] = bd[pos - 1]; bj[pos] = bj[pos - 1]; --pos;
        }
        bd[pos] = d2; bj[pos] = j;
        if (cnt < k) ++cnt;
    }
    for (int t = 0; t < k; ++t) nbr[(long)k * i + t] = t < cnt ? bj[t] : -1;
    double acc[78];
    for (int t = 0; t < 78; ++t) acc[t] = 0;
    for (int t = 0; t < cnt; ++t) {
        const int j = bj[t];
        const double dx = h * min_image(ijk[3 * j] - xi, n), dy = h * min_image(ijk[3 * j + 1] - yi, n),
                     dz = h * min_image(ijk[3 * j + 2] - zi, n);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        add_bond(acc, dz / r, atan2(dy, dx), false);
    }
    double* o = out + 4L * i;
    if (cnt == 0) {
        for (int t = 0; t < 4; ++t) o[t] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    for (int t = 0; t < 78; ++t) acc[t] /= (double)cnt;
    o[0] = ql_dev(acc, 4);
    o[1] = ql_dev(acc + 18, 6);
    o[2] = ql_dev(acc + 44, 8);
    o[3] = w4_dev(acc, W);
    if (h4 && nbins > 0) {
        int b4 = (int)(o[0] * nbins), b6 = (int)(o[1] * nbins);
        b4 = b4 < 0 ? 0 : (b4 >= nbins ? nbins - 1 : b4);
        b6 = b6 < 0 ? 0 : (b6 >= nbins ? nbins - 1 : b6);
        atomicAdd(&h4[b4], 1ull);
        atomicAdd(&h6[b6], 1ull);
    }
}

// device-resident peaks (linear cell indices, count on the device) -> ijk
__global__ void __launch_bounds__(256) k_idx_to_ijk(const int64_t* idx, const int64_t* d_count, int64_t cap, int nx,
                                                    int ny, int* ijk) {
    const int64_t P = *d_count < cap ? *d_count : cap;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < P; i += (int64_t)gridDim.x * 256) {
        const int64_t c = idx[i];
        ijk[3 * i] = (int)(c % nx);
        ijk[3 * i + 1] = (int)((c / nx) % ny);
        ijk[3 * i + 2] = (int)(c / ((int64_t)nx * ny));
    }
}

// k nearest neighbours on a periodic nx*ny*nz grid (spacing h) and q4/q6 histograms only
__global__ void __launch_bounds__(128) k_bo_knn_hist(const int* ijk, const int64_t* d_count, int64_t cap, int nx,
                                                     int ny, int nz, double h, int k, Wig4 W,
                                                     unsigned long long* h4, unsigned long long* h6, int nbins) {
    const int P = (int)(*d_count < cap ? *d_count : cap);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    long bd[KMAX];
    int bj[KMAX];
    int cnt = 0;
    const int xi = ijk[3 * i], yi = ijk[3 * i + 1], zi = ijk[3 * i + 2];
    for (int j = 0; j < P; ++j) {
        if (j == i) continue;
        const long dx = min_image(ijk[3 * j] - xi, nx), dy = min_image(ijk[3 * j + 1] - yi, ny),
                   dz = min_image(ijk[3 * j + 2] - zi, nz);
        const long d2 = dx * dx + dy * dy + dz * dz;
        if (cnt == k && !(d2 < bd[k - 1] || (d2 == bd[k - 1] && j < bj[k - 1]))) continue;
        int pos = cnt < k ? cnt : k - 1;
        while (pos > 0 && (bd[pos - 1] > d2 || (bd[pos - 1] == d2 && bj[pos - 1] > j))) {
            bd[pos] = bd[pos - 1]; bj[pos] = bj[pos - 1]; --pos;
        }
        bd[pos] = d2; bj[pos] = j;
        if (cnt < k) ++cnt;
    }
    if (cnt == 0) return;
    double acc[78];
    for (int t = 0; t < 78; ++t) acc[t] = 0;
    for (int t = 0; t < cnt; ++t) {
        const int j = bj[t];
        const double dx = h * min_image(ijk[3 * j] - xi, nx), dy = h * min_image(ijk[3 * j + 1] - yi, ny),
                     dz = h * min_image(ijk[3 * j + 2] - zi, nz);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        add_bond(acc, dz / r, atan2(dy, dx), false);
    }
    for (int t = 0; t < 78; ++t) acc[t] /= (double)cnt;
    int b4 = (int)(ql_dev(acc, 4) * nbins), b6 = (int)(ql_dev(acc + 18, 6) * nbins);
    b4 = b4 < 0 ? 0 : (b4 >= nbins ? nbins - 1 : b4);
    b6 = b6 < 0 ? 0 : (b6 >= nbins ? nbins - 1 : b6);
    atomicAdd(&h4[b4], 1ull);
    atomicAdd(&h6[b6], 1ull);
}

// Wigner 3j (4 4 4; m1 m2 m3) by the Racah formula, on the host
static double hfact(int k) { double r = 1; for (int i = 2; i <= k; ++i) r *= i; return r; }
static double w3j444(int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || std::abs(m1) > 4 || std::abs(m2) > 4 || std::abs(m3) > 4) return 0.0;
    const int j = 4;
    const double tri = hfact(j) * hfact(j) * hfact(j) / hfact(3 * j + 1);
    const double pre = std::sqrt(tri * hfact(j + m1) * hfact(j - m1) * hfact(j + m2) * hfact(j - m2) * hfact(j + m3) *
                                 hfact(j - m3));
    double sum = 0;
    for (int k = 0; k <= 3 * j; ++k) {
        const int a1 = k + m1, a2 = k - m2, a3 = j - k, a4 = j - k - m1, a5 = j - k + m2;
        if (a1 < 0 || a2 < 0 || a3 < 0 || a4 < 0 || a5 < 0) continue;
        sum += ((k & 1) ? -1.0 : 1.0) / (hfact(k) * hfact(a1) * hfact(a2) * hfact(a3) * hfact(a4) * hfact(a5));
    }
    const int ph = -m3;
    return (((ph % 2) + 2) % 2 ? -1.0 : 1.0) * pre * sum;
}

static Wig4 wig4() {
    Wig4 W;
    for (int m1 = -4; m1 <= 4; ++m1)
        for (int m2 = -4; m2 <= 4; ++m2) W.w[(m1 + 4) * 9 + (m2 + 4)] = w3j444(m1, m2, -m1 - m2);
    return W;
}

int knn_hist_launch(const int64_t* d_idx, const int64_t* d_count, int64_t cap, const int ext[3], double h, int k,
                    int* d_ijk, unsigned long long* d_h4, unsigned long long* d_h6, int nbins, hipStream_t s) {
    MCS_CHECK(k >= 1 && k <= KMAX, MCS_ERR_INVALID, "k must be in [1, 32]");
    const unsigned g1 = (unsigned)std::min<int64_t>(4096, (cap + 255) / 256);
    hipLaunchKernelGGL(k_idx_to_ijk, dim3(g1 ? g1 : 1), dim3(256), 0, s, d_idx, d_count, cap, ext[0], ext[1], d_ijk);
    const Wig4 W = wig4();
    const unsigned g2 = (unsigned)((cap + 127) / 128);
    hipLaunchKernelGGL(k_bo_knn_hist, dim3(g2 ? g2 : 1), dim3(128), 0, s, d_ijk, d_count, cap, ext[0], ext[1], ext[2],
                       h, k, W, d_h4, d_h6, nbins);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

extern "C" int mcs_bond_order_cutoff(const double* xyz, int npts, double box, double rlim, int* nb, double* out9) {
    MCS_CHECK(npts >= 0, MCS_ERR_INVALID, "negative point count");
    if (npts == 0) return MCS_OK;
    MCS_CHECK(xyz && nb && out9, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    double *dx = nullptr, *dq = nullptr, *dout = nullptr;
    int *dnb = nullptr, *dlast = nullptr;
    int st = MCS_OK;
    if (hipMalloc((void**)&dx, 3L * npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dq, (long)QS * npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dout, 9L * npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dnb, npts * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&dlast, npts * sizeof(int)) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, "bond-order allocation failed");
        st = MCS_ERR_NOMEM;
    }
    if (!st) {
        const Wig4 W = wig4();
        const unsigned g = (unsigned)((npts + 127) / 128);
        if (hipMemcpyAsync(dx, xyz, 3L * npts * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) st = MCS_ERR_HIP;
        hipLaunchKernelGGL(k_bo_qlm, dim3(g), dim3(128), 0, s, dx, npts, box, rlim, dq, dnb, dlast);
        hipLaunchKernelGGL(k_bo_stats, dim3(g), dim3(128), 0, s, dx, npts, box, rlim, dq, dnb, dlast, W, dout);
        if (!st && hipGetLastError() != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hipMemcpyAsync(nb, dnb, npts * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hipMemcpyAsync(out9, dout, 9L * npts * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
            st = MCS_ERR_HIP;
        if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
        if (st) mcs_set_last_error(__FILE__, __LINE__, "HIP failure in bond order");
    }
    (void)hipFree(dx); (void)hipFree(dq); (void)hipFree(dout); (void)hipFree(dnb); (void)hipFree(dlast);
    return st;
}

extern "C" int mcs_bond_order_knn(const int* ijk, int npts, int n, double h, int k, int* nbr, double* out4,
                                  int64_t* hist_q4, int64_t* hist_q6, int nbins) {
    MCS_CHECK(npts >= 0 && n >= 1 && k >= 1 && k <= KMAX, MCS_ERR_INVALID, "bad argument (1 <= k <= 32)");
    MCS_CHECK(!(hist_q4 || hist_q6) || (hist_q4 && hist_q6 && nbins > 0), MCS_ERR_INVALID, "bad histogram args");
    if (hist_q4) for (int b = 0; b < nbins; ++b) { hist_q4[b] = 0; hist_q6[b] = 0; }
    if (npts == 0) return MCS_OK;
    MCS_CHECK(ijk && nbr && out4, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    int *dijk = nullptr, *dnbr = nullptr;
    double* dout = nullptr;
    unsigned long long* dh = nullptr;
    int st = MCS_OK;
    const int nb2 = hist_q4 ? 2 * nbins : 2;
    if (hipMalloc((void**)&dijk, 3L * npts * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&dnbr, (long)k * npts * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&dout, 4L * npts * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dh, nb2 * sizeof(unsigned long long)) != hipSuccess) {
        mcs_set_last_error(__FILE__, __LINE__, "bond-order allocation failed");
        st = MCS_ERR_NOMEM;
    }
    if (!st) {
        const Wig4 W = wig4();
        const unsigned g = (unsigned)((npts + 127) / 128);
        if (hipMemcpyAsync(dijk, ijk, 3L * npts * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hipMemsetAsync(dh, 0, nb2 * sizeof(unsigned long long), s) != hipSuccess) st = MCS_ERR_HIP;
        hipLaunchKernelGGL(k_bo_knn, dim3(g), dim3(128), 0, s, dijk, npts, n, h, k, W, dnbr, dout,
                           hist_q4 ? dh : nullptr, hist_q4 ? dh + nbins : nullptr, nbins);
        if (!st && hipGetLastError() != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hipMemcpyAsync(nbr, dnbr, (long)k * npts * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
            st = MCS_ERR_HIP;
        if (!st && hipMemcpyAsync(out4, dout, 4L * npts * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
            st = MCS_ERR_HIP;
        std::vector<unsigned long long> hh(nb2);
        if (!st && hipMemcpyAsync(hh.data(), dh, nb2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) != hipSuccess)
            st = MCS_ERR_HIP;
        if (!st && hipStreamSynchronize(s) != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hist_q4)
            for (int b = 0; b < nbins; ++b) { hist_q4[b] = (int64_t)hh[b]; hist_q6[b] = (int64_t)hh[nbins + b]; }
        if (st) mcs_set_last_error(__FILE__, __LINE__, "HIP failure in bond order");
    }
    (void)hipFree(dijk); (void)hipFree(dnbr); (void)hipFree(dout); (void)hipFree(dh);
    return st;
}
