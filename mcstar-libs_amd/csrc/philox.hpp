// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the library's
// conserved k-space noise built on it.  Replaces the cuRAND XORWOW host API
// and the Hermitian repair kernels of cuda/noise.hpp:8-150.
//
// Streams: key = (seed lo, seed hi); counter = (index lo, index hi, draw, tag).
#pragma once

#include "mcs_common.hpp"

#define MCS_TAG_KNOISE 0u
#define MCS_TAG_INIT 1u
#define MCS_TAG_REAL 2u
#define MCS_TAG_KROW 3u

struct u32x4 { uint32_t x, y, z, w; };

MCS_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// uniform in (0, 1) with 24 bits: exact in float and double
template <typename F> MCS_HD F u01_24(uint32_t x) { return ((F)(x >> 8) + F(0.5)) * F(5.9604644775390625e-08); }

MCS_DEV void mcs_sincospi(float x, float* s, float* c) { sincospif(x, s, c); }
MCS_DEV void mcs_sincospi(double x, double* s, double* c) { sincospi(x, s, c); }

// raw unit complex Gaussian (E|xi|^2 = 1) at half-spectrum index h; *z0 = first N(0,1) normal
template <typename F>
MCS_DEV C2<F> raw_xi(uint64_t h, uint32_t k0, uint32_t k1, uint32_t draw, F* z0) {
    const u32x4 o = philox4x32_10(u32x4{(uint32_t)h, (uint32_t)(h >> 32), draw, MCS_TAG_KNOISE}, k0, k1);
    const F u1 = u01_24<F>(o.x), u2 = u01_24<F>(o.y);
    const F r = sqrt(F(-2) * log(u1));
    F s, c;
    mcs_sincospi(F(2) * u2, &s, &c);
    *z0 = r * c;
    const F q = F(0.70710678118654752440);
    return mk<F>(r * c * q, r * s * q);
}

// Hermitian-consistent unit noise at half-spectrum mode (i0, i1, i2) of a real
// nx*ny*nz grid.  On the i0 = 0 and i0 = nx/2 planes, the member of a pair
// (i1,i2) ~ (-i1,-i2) with the smaller i2*ny+i1 draws and its mirror takes the
// conjugate; self-conjugate modes are real with unit variance (the same
// statistics as cuda/noise.hpp:36-44, 82-89).
// kept out of line: inlined into the fully unrolled update loops it would outgrow the unroller's budget
template <typename F>
__device__ __noinline__ C2<F> noise_xi(int i0, int i1, int i2, int nx, int ny, int nz, uint32_t k0, uint32_t k1,
                                       uint32_t draw) {
    const int nxc = nx / 2 + 1;
    F z0;
    if (i0 == 0 || ((nx & 1) == 0 && i0 == nx / 2)) {
        const int m1 = i1 == 0 ? 0 : ny - i1;
        const int m2 = i2 == 0 ? 0 : nz - i2;
        const uint64_t lin = (uint64_t)i2 * ny + i1, mlin = (uint64_t)m2 * ny + m1;
        if (lin == mlin) {
            raw_xi<F>(lin * nxc + i0, k0, k1, draw, &z0);
            return mk<F>(z0, F(0));
        }
        if (lin < mlin) return raw_xi<F>(lin * nxc + i0, k0, k1, draw, &z0);
        return cconj(raw_xi<F>(mlin * nxc + i0, k0, k1, draw, &z0));
    }
    return raw_xi<F>(((uint64_t)i2 * ny + i1) * nxc + i0, k0, k1, draw, &z0);
}

// Unit complex Gaussian from two 32-bit draws (Box-Muller); *z0 = r cos(theta), a unit
// normal for real modes.  In FP32 the radius uses the hardware base-2 logarithm and
// square root (1 ulp each); the angle keeps the library sincospi, since the hardware
// sine and cosine (about 1e-6 absolute) made FP32 and FP64 noisy runs drift apart
// ten times faster.
template <typename F> MCS_DEV C2<F> box_muller(uint32_t x, uint32_t y, F* z0) {
    const F u1 = u01_24<F>(x), u2 = u01_24<F>(y);
    F r, sn, cs;
    if constexpr (sizeof(F) == 4) {
        r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));  // -2 ln u1
    } else {
        r = sqrt(-2.0 * log(u1));
    }
    mcs_sincospi(F(2) * u2, &sn, &cs);
    *z0 = r * cs;
    const F q = F(0.70710678118654752440);
    return mk<F>(r * cs * q, r * sn * q);
}

// The fused stepper's noise layout (even nx, L = nx/2): the Philox block at counter
// (c*L + p, draw, KROW) of row c gives .a for the full x mode p of that row (from x, y)
// and .b for the full x mode p + L (from z, w).  Of a mirror pair of rows the smaller
// index draws and the other conjugates; in a self-mirrored row, 0 < p < L draws a at p
// and the modes 0 and L are real: sqrt(2) Re a and sqrt(2) Re b (= r cos theta).
template <typename F> struct XiPair {
    C2<F> a, b;
};
template <typename F> MCS_DEV XiPair<F> xi_block_inl(uint64_t c, uint32_t k0, uint32_t k1, uint32_t draw) {
    const u32x4 o = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), draw, MCS_TAG_KROW}, k0, k1);
    F z;
    XiPair<F> r;
    r.a = box_muller<F>(o.x, o.y, &z);
    r.b = box_muller<F>(o.z, o.w, &z);
    return r;
}
// the same, out of line: inlined into the largest unrolled x passes it pushes their register
// arrays to scratch
template <typename F>
__device__ __noinline__ XiPair<F> xi_block(uint64_t c, uint32_t k0, uint32_t k1, uint32_t draw) {
    return xi_block_inl<F>(c, k0, k1, draw);
}
