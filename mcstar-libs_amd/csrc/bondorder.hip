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
// Cutoff mode: one thread per point visits its neighbours in ascending index
// order, so the sums run in the reference's order (distances with contraction
// off, so neighbour sets match the CPU bit for bit).  The neighbours come from
// a cell list when every point lies in the box, the box holds at least 3 cells
// of width rlim per axis and no point has more than 64 neighbours; otherwise
// from a scan of all points.
//
// k-nearest mode: integer squared distances on the periodic grid, ties broken
// by the smaller index, exact Y_lm, q4 q6 q8 w4 and histograms of q4, q6.  The
// neighbours come from a periodic cell list (below), not an all-pairs scan.
#include "knn.hpp"

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

// accumulate the 2l + 1 harmonics of degree L of one bond into q[2 OFF ..] (re, im interleaved).
// Unrolled over m, so every index into the accumulator is a constant and it stays in registers
// (with run-time indices the 78 accumulators of a thread went to scratch memory); the arithmetic
// of each term is the same as with loops.
template <int L, int OFF>
MCS_DEV void add_l(double (&q)[78], double ct, double phi, bool pole_rule) {
#pragma unroll
    for (int m = -L; m <= L; ++m) {
        if (pole_rule && ((m < 0 && fabs(1.0 + ct) < 1e-3) || (m > 0 && fabs(1.0 - ct) < 1e-3))) continue;
        double re, im;
        ylm_dev(L, m, ct, phi, &re, &im);
        q[2 * (OFF + m + L)] += re;
        q[2 * (OFF + m + L) + 1] += im;
    }
}

// accumulate the 39 (l = 4, 6, 8) harmonics of one bond into q[78] (re, im interleaved)
MCS_DEV void add_bond(double (&q)[78], double ct, double phi, bool pole_rule) {
    add_l<4, 0>(q, ct, phi, pole_rule);
    add_l<6, 9>(q, ct, phi, pole_rule);
    add_l<8, 22>(q, ct, phi, pole_rule);
}

MCS_DEV double ql_dev(const double* q, int l) {
    double s = 0;
#pragma unroll
    for (int m = 0; m <= 2 * l; ++m) s += q[2 * m] * q[2 * m] + q[2 * m + 1] * q[2 * m + 1];
    return sqrt(4.0 * M_PI / (2.0 * l + 1.0) * s);
}

MCS_DEV double w4_dev(const double* q4, const Wig4& W) {
    double acc = 0;
#pragma unroll
    for (int m1 = -4; m1 <= 4; ++m1)
#pragma unroll
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
static constexpr int CMAXN = 64;  // neighbours per point the cutoff cell list holds

// The neighbours of point i, in ascending index order: every j (all-pairs form) or the
// sorted list nl[i*CMAXN ..] of cnt[i] candidates within rlim (cell-list form); the bond
// test is repeated for both, so both forms visit the same j in the same order.
template <bool LIST>
MCS_DEV int nbr_count(int P, const int* nc, int i) { return LIST ? nc[i] : P; }
template <bool LIST>
MCS_DEV int nbr_at(const int* nl, int i, int u) { return LIST ? nl[(long)CMAXN * i + u] : u; }

template <bool LIST>
__global__ void __launch_bounds__(128) k_bo_qlm(const double* xyz, int P, double box, double rlim, double* q,
                                                int* nb, int* last, const int* nl, const int* nc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    double acc[78];
#pragma unroll
    for (int t = 0; t < 78; ++t) acc[t] = 0;
    int cnt = 0, lj = -1;
    const int nn = nbr_count<LIST>(P, nc, i);
    for (int u = 0; u < nn; ++u) {
        const int j = nbr_at<LIST>(nl, i, u);
        if (j == i) continue;
        double ct, ph;
        if (bond(xyz, i, j, box, rlim, &ct, &ph)) { add_bond(acc, ct, ph, true); ++cnt; lj = j; }
    }
    double* qi = q + (long)QS * i;
#pragma unroll
    for (int t = 0; t < 78; ++t) qi[t] = cnt ? acc[t] / (double)cnt : 0.0;
    nb[i] = cnt;
    last[i] = lj;
}

template <bool LIST>
__global__ void __launch_bounds__(128) k_bo_stats(const double* xyz, int P, double box, double rlim, const double* q,
                                                  const int* nb, const int* last, Wig4 W, double* out, const int* nl,
                                                  const int* nc) {
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
    const int nn = nbr_count<LIST>(P, nc, i);
    for (int u = 0; u < nn; ++u) {
        const int j = nbr_at<LIST>(nl, i, u);
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

// ---------------------------------------------------------------- cutoff cell list
// Points in [0, box)^3 are binned into g^3 cells of width box/g >= rlim (1 + 1e-6), g >= 3: a
// neighbour within rlim under the minimum image lies in one of the 27 cells around a point's
// own.  Each point gathers the candidates of those cells that pass the bond test, sorts
// them by index and keeps at most CMAXN (a larger count sets the overflow flag and the
// caller falls back to the all-pairs form).
MCS_DEV int cut_cell(double x, double box, int g) {
    int c = (int)(x * g / box);
    return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

__global__ void __launch_bounds__(256) k_cut_count(const double* xyz, int P, double box, int g, int* cnt) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P; i += gridDim.x * 256) {
        const int c = (cut_cell(xyz[3 * i + 2], box, g) * g + cut_cell(xyz[3 * i + 1], box, g)) * g +
                      cut_cell(xyz[3 * i], box, g);
        atomicAdd(&cnt[c], 1);
    }
}

__global__ void __launch_bounds__(256) k_cut_scatter(const double* xyz, int P, double box, int g, const int* start,
                                                     int* cnt, int* sorted) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P; i += gridDim.x * 256) {
        const int c = (cut_cell(xyz[3 * i + 2], box, g) * g + cut_cell(xyz[3 * i + 1], box, g)) * g +
                      cut_cell(xyz[3 * i], box, g);
        sorted[start[c] + atomicSub(&cnt[c], 1) - 1] = i;
    }
}

__global__ void __launch_bounds__(128) k_cut_nbrs(const double* xyz, int P, double box, double rlim, int g,
                                                  const int* start, const int* sorted, int* nl, int* nc,
                                                  int* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int cx = cut_cell(xyz[3 * i], box, g), cy = cut_cell(xyz[3 * i + 1], box, g),
              cz = cut_cell(xyz[3 * i + 2], box, g);
    int* li = nl + (long)CMAXN * i;
    int n = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int c = ((cz + dz + g) % g * g + (cy + dy + g) % g) * g + (cx + dx + g) % g;
                for (int p = start[c]; p < start[c + 1]; ++p) {
                    const int j = sorted[p];
                    double ct, ph;
                    if (j == i || !bond(xyz, i, j, box, rlim, &ct, &ph)) continue;
                    if (n == CMAXN) {
                        atomicOr(overflow, 1);
                        nc[i] = n;
                        return;
                    }
                    int pos = n++;  // insertion in ascending index order
                    while (pos > 0 && li[pos - 1] > j) {
                        li[pos] = li[pos - 1];
                        --pos;
                    }
                    li[pos] = j;
                }
            }
    nc[i] = n;
}

static constexpr int KMAX = 32;

// ---------------------------------------------------------------- k nearest neighbours on a cell list
//
// Peaks sit on the integer grid, so squared periodic distances are exact
// integers and the key (d2 << 32 | index) orders candidates strictly: the k
// smallest keys form one set whatever order candidates arrive in, the same set
// an all-pairs scan keeps (distance first, then the smaller index).
//
// Points are binned into a periodic grid of g0 x g1 x g2 cells; coordinate x
// falls in cell x*g/n, so every cell spans floor(n/g) or ceil(n/g) grid lines.
// A query visits shells of cells at Chebyshev cell distance s = 0, 1, 2, ...,
// with per-axis offsets in [-(g-1)/2, g/2] so that each cell is visited once
// also when a shell wraps around the box.  A cell not yet visited after shell
// s lies s+1 or more cells away along some axis, so every point in it is at
// least s*wmin + 1 grid lines away (wmin = the narrowest cell); the search
// stops once the k-th key's d2 is below (s*wmin + 1)^2.  Cells are sized from
// the point density so that the k-th neighbour usually lies within shell 1.

struct KnnGrid {
    int64_t P;   // points
    int n[3];    // periodic extents
    int g[3];    // cells per axis
    int wmin;    // min over axes of floor(n / g)
    int smax;    // shells that cover every cell: max over axes of g / 2
    int ncell;
};

static constexpr int KNN_MAX_EXT = 32768;  // keeps 3 (n/2)^2 below 2^32 and x*g below 2^31

MCS_DEV int knn_cell(const KnnGrid& G, int x, int y, int z) {
    const int cx = x * G.g[0] / G.n[0], cy = y * G.g[1] / G.n[1], cz = z * G.g[2] / G.n[2];
    return (cz * G.g[1] + cy) * G.g[0] + cx;
}

MCS_DEV int pmod(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

// signed minimum image of a - b for a, b in [0, n): into (-n/2, n/2]
MCS_DEV int min_image_norm(int a, int b, int n) {
    int d = a - b;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

// cell grid from the point count: cells about as wide as the radius of the ball
// that holds k points at the mean density, coarsened until at most cmax cells
__global__ void k_knn_setup(const int64_t* d_count, int64_t P_host, int64_t cap, int nx, int ny, int nz, int k,
                            int cmax, KnnGrid* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t P = d_count ? *d_count : P_host;
    P = P < 0 ? 0 : (P > cap ? cap : P);
    KnnGrid G;
    G.P = P;
    G.n[0] = nx; G.n[1] = ny; G.n[2] = nz;
    const int dims = (nx > 1) + (ny > 1) + (nz > 1);
    int w = KNN_MAX_EXT;
    if (P > 1 && dims > 0) {
        const double per = (double)nx * ny * nz * k / (double)P;  // volume that holds k points on average
        const double r = dims == 3 ? cbrt(per * 3.0 / (4.0 * M_PI)) : (dims == 2 ? sqrt(per / M_PI) : 0.5 * per);
        w = r < (double)KNN_MAX_EXT ? (int)ceil(r) : KNN_MAX_EXT;
        w = w < 1 ? 1 : w;
    }
    for (;;) {
        long cells = 1;
        for (int a = 0; a < 3; ++a) {
            const int g = G.n[a] / w;
            G.g[a] = g < 1 ? 1 : g;
            cells *= G.g[a];
        }
        if (cells <= cmax || w >= KNN_MAX_EXT) break;
        w *= 2;
    }
    G.wmin = KNN_MAX_EXT;
    G.smax = 0;
    G.ncell = 1;
    for (int a = 0; a < 3; ++a) {
        const int wa = G.n[a] / G.g[a];
        G.wmin = wa < G.wmin ? wa : G.wmin;
        G.smax = G.g[a] / 2 > G.smax ? G.g[a] / 2 : G.smax;
        G.ncell *= G.g[a];
    }
    *out = G;
}

__global__ void __launch_bounds__(256) k_knn_zero(const KnnGrid* gp, int* cnt) {
    const int nc = gp->ncell;
    for (int c = blockIdx.x * 256 + threadIdx.x; c <= nc; c += gridDim.x * 256) cnt[c] = 0;
}

// points -> (x, y, z, index) with coordinates in [0, n), and the count of every cell
__global__ void __launch_bounds__(256) k_knn_bin(const KnnGrid* gp, const int64_t* idx, const int* ijk, int4* pts,
                                                 int* cnt) {
    const KnnGrid G = *gp;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < G.P; i += (int64_t)gridDim.x * 256) {
        int x, y, z;
        if (idx) {
            const int64_t c = idx[i];
            x = (int)(c % G.n[0]);
            y = (int)((c / G.n[0]) % G.n[1]);
            z = (int)((c / ((int64_t)G.n[0] * G.n[1])) % G.n[2]);
        } else {
            x = pmod(ijk[3 * i], G.n[0]);
            y = pmod(ijk[3 * i + 1], G.n[1]);
            z = pmod(ijk[3 * i + 2], G.n[2]);
        }
        pts[i] = make_int4(x, y, z, (int)i);
        atomicAdd(&cnt[knn_cell(G, x, y, z)], 1);
    }
}

// exclusive scan of the cell counts into start[0..ncell] (start[ncell] = P), one block
__global__ void __launch_bounds__(1024) k_knn_scan(const KnnGrid* gp, const int* cnt, int* start) {
    __shared__ int wsum[16];
    const int nc = gp->ncell;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int carry = 0;
    for (int base = 0; base < nc; base += 4096) {
        const int b = base + 4 * (int)threadIdx.x;
        int v[4], s = 0;
        for (int e = 0; e < 4; ++e) {
            v[e] = b + e < nc ? cnt[b + e] : 0;
            s += v[e];
        }
        int x = s;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int q = 0; q < 16; ++q) {
            pre += q < wv ? wsum[q] : 0;
            tot += wsum[q];
        }
        int run = carry + pre + x - s;
        for (int e = 0; e < 4; ++e) {
            if (b + e < nc) start[b + e] = run;
            run += v[e];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[nc] = carry;
}

// points into cell order (the order inside a cell is arbitrary; the result does not depend on it)
__global__ void __launch_bounds__(256) k_knn_scatter(const KnnGrid* gp, const int4* pts, const int* start, int* cnt,
                                                     int4* sorted) {
    const KnnGrid G = *gp;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < G.P; i += (int64_t)gridDim.x * 256) {
        const int4 p = pts[i];
        const int c = knn_cell(G, p.x, p.y, p.z);
        sorted[start[c] + atomicSub(&cnt[c], 1) - 1] = p;
    }
}

// one thread per point, threads in cell order: the k nearest keys in registers
// (kept sorted, KC >= k slots), then the bond harmonics in key order
template <int KC, bool FULL>
__global__ void __launch_bounds__(128) k_knn_query(const KnnGrid* gp, const int4* __restrict__ pts,
                                                   const int4* __restrict__ sorted, const int* __restrict__ start,
                                                   int k, double h, Wig4 W, int* nbr, double* out,
                                                   unsigned long long* h4, unsigned long long* h6, int nbins) {
    __shared__ int js[KC][128];
    const KnnGrid G = *gp;
    const int64_t t = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (t >= G.P) return;
    const int4 me = sorted[t];
    const int c0 = me.x * G.g[0] / G.n[0], c1 = me.y * G.g[1] / G.n[1], c2 = me.z * G.g[2] / G.n[2];
    unsigned long long best[KC];
#pragma unroll
    for (int e = 0; e < KC; ++e) best[e] = ~0ull;
    unsigned long long kth = ~0ull;
    int found = 0;
    for (int s = 0; s <= G.smax; ++s) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = max(-s, -((G.g[a] - 1) / 2));
            hi[a] = min(s, G.g[a] / 2);
        }
        for (int oz = lo[2]; oz <= hi[2]; ++oz) {
            const int cz = pmod(c2 + oz, G.g[2]);
            for (int oy = lo[1]; oy <= hi[1]; ++oy) {
                const int cy = pmod(c1 + oy, G.g[1]);
                const int row = (cz * G.g[1] + cy) * G.g[0];
                // rows on the shell's faces take every x offset, inner rows only the two x ends
                const bool face = abs(oz) == s || abs(oy) == s;
                const int xb = face ? lo[0] : (-s >= lo[0] ? -s : s);
                const int xe = face ? hi[0] : (s <= hi[0] ? s : -s);
                const int xs = face ? 1 : 2 * s;
                for (int ox = xb; ox <= xe; ox += xs) {
                    const int cell = row + pmod(c0 + ox, G.g[0]);
                    const int pe = start[cell + 1];
                    for (int p = start[cell]; p < pe; ++p) {
                        const int4 q = sorted[p];
                        if (q.w == me.w) continue;
                        int dx = abs(q.x - me.x), dy = abs(q.y - me.y), dz = abs(q.z - me.z);
                        dx = min(dx, G.n[0] - dx);
                        dy = min(dy, G.n[1] - dy);
                        dz = min(dz, G.n[2] - dz);
                        const unsigned long long d2 =
                            (unsigned long long)(unsigned)(dx * dx) + (unsigned)(dy * dy) + (unsigned)(dz * dz);
                        const unsigned long long key = (d2 << 32) | (unsigned)q.w;
                        if (key >= kth) continue;
                        unsigned long long carry = key;  // insert, dropping the largest
#pragma unroll
                        for (int e = 0; e < KC; ++e) {
                            const unsigned long long lo_e = best[e] < carry ? best[e] : carry;
                            carry = best[e] < carry ? carry : best[e];
                            best[e] = lo_e;
                        }
                        if (found < k) ++found;
                        if (found == k) {
                            unsigned long long v = best[0];
#pragma unroll
                            for (int e = 1; e < KC; ++e) v = e == k - 1 ? best[e] : v;
                            kth = v;
                        }
                    }
                }
            }
        }
        const unsigned long long b = (unsigned long long)s * G.wmin + 1;
        if (found == k && (kth >> 32) < b * b) break;
    }
    const int i = me.w;
#pragma unroll
    for (int e = 0; e < KC; ++e) js[e][threadIdx.x] = e < found ? (int)(unsigned)(best[e] & 0xffffffffull) : -1;
    if (FULL)
        for (int e = 0; e < k; ++e) nbr[(int64_t)k * i + e] = js[e][threadIdx.x];
    if (found == 0) {
        if (FULL)
            for (int e = 0; e < 4; ++e) out[4L * i + e] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    double acc[78];
#pragma unroll
    for (int e = 0; e < 78; ++e) acc[e] = 0;
    for (int e = 0; e < found; ++e) {
        const int4 q = pts[js[e][threadIdx.x]];
        const double dx = h * min_image_norm(q.x, me.x, G.n[0]), dy = h * min_image_norm(q.y, me.y, G.n[1]),
                     dz = h * min_image_norm(q.z, me.z, G.n[2]);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        add_bond(acc, dz / r, atan2(dy, dx), false);
    }
#pragma unroll
    for (int e = 0; e < 78; ++e) acc[e] /= (double)found;
    const double q4 = ql_dev(acc, 4), q6 = ql_dev(acc + 18, 6);
    if (FULL) {
        double* o = out + 4L * i;
        o[0] = q4;
        o[1] = q6;
        o[2] = ql_dev(acc + 44, 8);
        o[3] = w4_dev(acc, W);
    }
    if (h4 && nbins > 0) {
        int b4 = (int)(q4 * nbins), b6 = (int)(q6 * nbins);
        b4 = b4 < 0 ? 0 : (b4 >= nbins ? nbins - 1 : b4);
        b6 = b6 < 0 ? 0 : (b6 >= nbins ? nbins - 1 : b6);
        atomicAdd(&h4[b4], 1ull);
        atomicAdd(&h6[b6], 1ull);
    }
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

struct KnnWs {
    KnnGrid* grid;
    int4* pts;
    int4* sorted;
    int* cnt;
    int* start;
    int cmax;
};

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

static int knn_cmax(int64_t cap) { return (int)(cap + 64); }

size_t knn_workspace_bytes(int64_t cap) {
    const size_t c = (size_t)(cap > 0 ? cap : 1);
    const size_t nc = (size_t)knn_cmax(cap) + 1;
    return al256(sizeof(KnnGrid)) + 2 * al256(c * sizeof(int4)) + 2 * al256(nc * sizeof(int));
}

static KnnWs knn_carve(void* ws, int64_t cap) {
    const size_t c = (size_t)(cap > 0 ? cap : 1);
    const size_t nc = (size_t)knn_cmax(cap) + 1;
    unsigned char* p = reinterpret_cast<unsigned char*>(ws);
    KnnWs w;
    w.grid = reinterpret_cast<KnnGrid*>(p); p += al256(sizeof(KnnGrid));
    w.pts = reinterpret_cast<int4*>(p); p += al256(c * sizeof(int4));
    w.sorted = reinterpret_cast<int4*>(p); p += al256(c * sizeof(int4));
    w.cnt = reinterpret_cast<int*>(p); p += al256(nc * sizeof(int));
    w.start = reinterpret_cast<int*>(p);
    w.cmax = knn_cmax(cap);
    return w;
}

template <int KC>
static void knn_query_launch(bool full, unsigned grid, hipStream_t s, const KnnWs& w, int k, double h, const Wig4& W,
                             int* nbr, double* out, unsigned long long* h4, unsigned long long* h6, int nbins) {
    if (full)
        hipLaunchKernelGGL((k_knn_query<KC, true>), dim3(grid), dim3(128), 0, s, w.grid, w.pts, w.sorted, w.start, k, h,
                           W, nbr, out, h4, h6, nbins);
    else
        hipLaunchKernelGGL((k_knn_query<KC, false>), dim3(grid), dim3(128), 0, s, w.grid, w.pts, w.sorted, w.start, k,
                           h, W, nbr, out, h4, h6, nbins);
}

int knn_launch(const int64_t* d_idx, const int* d_ijk, const int64_t* d_count, int64_t P_host, int64_t cap,
               const int ext[3], double h, int k, void* ws, int* d_nbr, double* d_out4, unsigned long long* d_h4,
               unsigned long long* d_h6, int nbins, hipStream_t s) {
    MCS_CHECK(k >= 1 && k <= KMAX, MCS_ERR_INVALID, "k must be in [1, 32]");
    MCS_CHECK(ext[0] >= 1 && ext[1] >= 1 && ext[2] >= 1 && ext[0] <= KNN_MAX_EXT && ext[1] <= KNN_MAX_EXT &&
                  ext[2] <= KNN_MAX_EXT,
              MCS_ERR_INVALID, "bond-order grid extents must be in [1, 32768]");
    MCS_CHECK(cap >= 0 && cap <= (int64_t)0x7fffff00, MCS_ERR_INVALID, "too many points for bond order");
    MCS_CHECK((d_idx != nullptr) != (d_ijk != nullptr), MCS_ERR_INVALID, "pass either indices or coordinates");
    MCS_CHECK(!(d_h4 || d_h6) || (d_h4 && d_h6 && nbins > 0), MCS_ERR_INVALID, "bad histogram args");
    const bool full = d_nbr != nullptr;
    MCS_CHECK(!full || d_out4, MCS_ERR_INVALID, "neighbour output needs the q output");
    if (cap == 0) return MCS_OK;
    const KnnWs w = knn_carve(ws, cap);
    const Wig4 W = wig4();
    const unsigned gp = (unsigned)std::min<int64_t>(4096, (cap + 255) / 256);
    const unsigned gz = (unsigned)std::min<int64_t>(1024, ((int64_t)w.cmax + 256) / 256);
    hipLaunchKernelGGL(k_knn_setup, dim3(1), dim3(64), 0, s, d_count, P_host, cap, ext[0], ext[1], ext[2], k, w.cmax,
                       w.grid);
    hipLaunchKernelGGL(k_knn_zero, dim3(gz), dim3(256), 0, s, w.grid, w.cnt);
    hipLaunchKernelGGL(k_knn_bin, dim3(gp), dim3(256), 0, s, w.grid, d_idx, d_ijk, w.pts, w.cnt);
    hipLaunchKernelGGL(k_knn_scan, dim3(1), dim3(1024), 0, s, w.grid, w.cnt, w.start);
    hipLaunchKernelGGL(k_knn_scatter, dim3(gp), dim3(256), 0, s, w.grid, w.pts, w.start, w.cnt, w.sorted);
    const unsigned gq = (unsigned)((cap + 127) / 128);
    if (k <= 8) knn_query_launch<8>(full, gq, s, w, k, h, W, d_nbr, d_out4, d_h4, d_h6, nbins);
    else if (k <= 12) knn_query_launch<12>(full, gq, s, w, k, h, W, d_nbr, d_out4, d_h4, d_h6, nbins);
    else if (k <= 16) knn_query_launch<16>(full, gq, s, w, k, h, W, d_nbr, d_out4, d_h4, d_h6, nbins);
    else knn_query_launch<32>(full, gq, s, w, k, h, W, d_nbr, d_out4, d_h4, d_h6, nbins);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

// Carves 256-byte aligned arrays out of one workspace; with a null base it only adds up the bytes.
struct Carve {
    unsigned char* base = nullptr;
    size_t off = 0;
    template <typename T> T* take(size_t n) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (n * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

// The cell list of the cutoff form applies when every point lies in [0, box)^3 and at least 3
// cells of width >= rlim fit per axis; returns the cells per axis, or 0.
static int cutoff_grid(const double* xyz, int P, double box, double rlim) {
    if (!(box > 0) || !(rlim > 0) || P < 256) return 0;
    for (long i = 0; i < 3L * P; ++i)
        if (!(xyz[i] >= 0.0 && xyz[i] < box)) return 0;
    const double gw = std::floor(box / (rlim * (1.0 + 1e-6)));
    if (!(gw >= 3.0)) return 0;
    return (int)std::min(gw, 256.0);
}

struct CutoffBufs {
    double *dx, *dq, *dout;
    int *dnb, *dlast;
    int *cnt, *start, *sorted, *ovf, *dnl, *dnc;  // cell list (g > 0)
    KnnGrid* dkg;
};

static CutoffBufs cutoff_carve(Carve& c, int P, int g) {
    CutoffBufs b{};
    b.dx = c.take<double>(3L * P);
    b.dq = c.take<double>((long)QS * P);
    b.dout = c.take<double>(9L * P);
    b.dnb = c.take<int>(P);
    b.dlast = c.take<int>(P);
    if (g > 0) {
        const long ncell = (long)g * g * g;
        b.cnt = c.take<int>(ncell + 1);
        b.start = c.take<int>(ncell + 1);
        b.sorted = c.take<int>(P);
        b.ovf = c.take<int>(1);
        b.dnl = c.take<int>((size_t)CMAXN * P);
        b.dnc = c.take<int>(P);
        b.dkg = c.take<KnnGrid>(1);
    }
    return b;
}

// Sorted neighbour lists of the cutoff form through the cell list on g^3 cells; *listed is false
// when a point has more than CMAXN neighbours (the all-pairs scan runs instead).
static int cutoff_lists(const CutoffBufs& b, int P, double box, double rlim, int g, hipStream_t s, bool* listed) {
    *listed = false;
    const long ncell = (long)g * g * g;
    KnnGrid kg{};
    kg.ncell = (int)ncell;
    const unsigned gp = (unsigned)std::min<long>(4096, (P + 255) / 256);
    const unsigned gz = (unsigned)std::min<long>(1024, (ncell + 256) / 256);
    MCS_HIP_TRY(hipMemcpyAsync(b.dkg, &kg, sizeof(KnnGrid), hipMemcpyHostToDevice, s));
    MCS_HIP_TRY(hipMemsetAsync(b.ovf, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_knn_zero, dim3(gz), dim3(256), 0, s, b.dkg, b.cnt);
    hipLaunchKernelGGL(k_cut_count, dim3(gp), dim3(256), 0, s, b.dx, P, box, g, b.cnt);
    hipLaunchKernelGGL(k_knn_scan, dim3(1), dim3(1024), 0, s, b.dkg, b.cnt, b.start);
    hipLaunchKernelGGL(k_cut_scatter, dim3(gp), dim3(256), 0, s, b.dx, P, box, g, b.start, b.cnt, b.sorted);
    hipLaunchKernelGGL(k_cut_nbrs, dim3((unsigned)((P + 127) / 128)), dim3(128), 0, s, b.dx, P, box, rlim, g, b.start,
                       b.sorted, b.dnl, b.dnc, b.ovf);
    MCS_HIP_TRY(hipGetLastError());
    int overflow = 0;
    MCS_HIP_TRY(hipMemcpyAsync(&overflow, b.ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipStreamSynchronize(s));
    *listed = !overflow;
    return MCS_OK;
}

extern "C" int mcs_bond_order_cutoff(const double* xyz, int npts, double box, double rlim, int* nb, double* out9) {
    MCS_CHECK(npts >= 0, MCS_ERR_INVALID, "negative point count");
    if (npts == 0) return MCS_OK;
    MCS_CHECK(xyz && nb && out9, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    const int g = cutoff_grid(xyz, npts, box, rlim);
    // one workspace, kept by the library between calls
    Carve c;
    cutoff_carve(c, npts, g);
    void* ws = nullptr;
    int st = mcs_workspace(MCS_WS_BOND, c.off, &ws);
    if (st) return st;
    c = Carve{reinterpret_cast<unsigned char*>(ws), 0};
    const CutoffBufs b = cutoff_carve(c, npts, g);
    const Wig4 W = wig4();
    const unsigned gq = (unsigned)((npts + 127) / 128);
    MCS_HIP_TRY(hipMemcpyAsync(b.dx, xyz, 3L * npts * sizeof(double), hipMemcpyHostToDevice, s));
    bool listed = false;
    if (g > 0) {
        st = cutoff_lists(b, npts, box, rlim, g, s, &listed);
        if (st) return st;
    }
    if (listed) {
        hipLaunchKernelGGL(k_bo_qlm<true>, dim3(gq), dim3(128), 0, s, b.dx, npts, box, rlim, b.dq, b.dnb, b.dlast, b.dnl,
                           b.dnc);
        hipLaunchKernelGGL(k_bo_stats<true>, dim3(gq), dim3(128), 0, s, b.dx, npts, box, rlim, b.dq, b.dnb, b.dlast, W,
                           b.dout, b.dnl, b.dnc);
    } else {
        hipLaunchKernelGGL(k_bo_qlm<false>, dim3(gq), dim3(128), 0, s, b.dx, npts, box, rlim, b.dq, b.dnb, b.dlast,
                           (const int*)nullptr, (const int*)nullptr);
        hipLaunchKernelGGL(k_bo_stats<false>, dim3(gq), dim3(128), 0, s, b.dx, npts, box, rlim, b.dq, b.dnb, b.dlast, W,
                           b.dout, (const int*)nullptr, (const int*)nullptr);
    }
    MCS_HIP_TRY(hipGetLastError());
    MCS_HIP_TRY(hipMemcpyAsync(nb, b.dnb, npts * sizeof(int), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipMemcpyAsync(out9, b.dout, 9L * npts * sizeof(double), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipStreamSynchronize(s));
    return MCS_OK;
}

extern "C" int mcs_bond_order_knn(const int* ijk, int npts, int n, double h, int k, int* nbr, double* out4,
                                  int64_t* hist_q4, int64_t* hist_q6, int nbins) {
    MCS_CHECK(npts >= 0 && n >= 1 && n <= KNN_MAX_EXT && k >= 1 && k <= KMAX, MCS_ERR_INVALID,
              "bad argument (1 <= k <= 32, 1 <= n <= 32768)");
    MCS_CHECK(!(hist_q4 || hist_q6) || (hist_q4 && hist_q6 && nbins > 0), MCS_ERR_INVALID, "bad histogram args");
    if (hist_q4) for (int b = 0; b < nbins; ++b) { hist_q4[b] = 0; hist_q6[b] = 0; }
    if (npts == 0) return MCS_OK;
    MCS_CHECK(ijk && nbr && out4, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    const int nb2 = hist_q4 ? 2 * nbins : 2;
    // coordinates, outputs, histogram bins and the cell list in one workspace kept by the library
    auto carve = [&](Carve& c, int** dijk, int** dnbr, double** dout, unsigned long long** dh, void** kws) {
        *dijk = c.take<int>(3L * npts);
        *dnbr = c.take<int>((long)k * npts);
        *dout = c.take<double>(4L * npts);
        *dh = c.take<unsigned long long>(nb2);
        *kws = c.take<unsigned char>(knn_workspace_bytes(npts));
    };
    int *dijk, *dnbr;
    double* dout;
    unsigned long long* dh;
    void* kws;
    Carve c;
    carve(c, &dijk, &dnbr, &dout, &dh, &kws);
    void* ws = nullptr;
    int st = mcs_workspace(MCS_WS_KNN, c.off, &ws);
    if (st) return st;
    c = Carve{reinterpret_cast<unsigned char*>(ws), 0};
    carve(c, &dijk, &dnbr, &dout, &dh, &kws);
    std::vector<unsigned long long> hh(nb2);
    const int ext[3] = {n, n, n};
    MCS_HIP_TRY(hipMemcpyAsync(dijk, ijk, 3L * npts * sizeof(int), hipMemcpyHostToDevice, s));
    MCS_HIP_TRY(hipMemsetAsync(dh, 0, nb2 * sizeof(unsigned long long), s));
    st = knn_launch(nullptr, dijk, nullptr, npts, npts, ext, h, k, kws, dnbr, dout, hist_q4 ? dh : nullptr,
                    hist_q4 ? dh + nbins : nullptr, nbins, s);
    if (st) return st;
    MCS_HIP_TRY(hipMemcpyAsync(nbr, dnbr, (long)k * npts * sizeof(int), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipMemcpyAsync(out4, dout, 4L * npts * sizeof(double), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipMemcpyAsync(hh.data(), dh, nb2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MCS_HIP_TRY(hipStreamSynchronize(s));
    if (hist_q4)
        for (int b = 0; b < nbins; ++b) { hist_q4[b] = (int64_t)hh[b]; hist_q6[b] = (int64_t)hh[nbins + b]; }
    return MCS_OK;
}
