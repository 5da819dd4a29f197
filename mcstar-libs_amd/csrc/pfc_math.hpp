// Per-mode PFC arithmetic shared by the reference-layout kernels, the fused
// stepper and the ensemble.
#pragma once

#include "mcs_common.hpp"

// wavenumber of index i on an n-point axis of length len (cuda/kernels/common.hpp:27-28)
template <typename F> MCS_HD F kwave(int i, int n, F len) {
    return F(i < n / 2 + 1 ? i : i - n) * F(6.283185307179586232) / len;
}

// semi-implicit conserved update of one mode:
//   psi' = (psi - dt k^2 nl + amp |k| xi) / (1 + dt k^2 (lin0 + (1 - k^2)^2))
// lin0 = -eps (cubic split, nl = psi^3) or 1 - eps (reference split, nl = psi^3 - psi)
// In FP32 the denominator is inverted with the hardware reciprocal (1 ulp) and
// |k| is taken only when the noise amplitude is non-zero: the exact divide and
// square root were a fifth of the fused x pass's vector instructions.
template <typename F>
MCS_DEV C2<F> pfc_mode_update(C2<F> psi, C2<F> nl, C2<F> xi, F ksq, F dt, F lin0, F amp) {
    const F om = F(1) - ksq;
    const F a = dt * ksq;
    const F den = fmad(a, fmad(om, om, lin0), F(1));
    F inv, na;
    if constexpr (sizeof(F) == 4) {
        inv = __builtin_amdgcn_rcpf(den);
        na = amp != F(0) ? amp * __builtin_amdgcn_sqrtf(ksq) : F(0);
    } else {
        inv = F(1) / den;
        na = amp != F(0) ? amp * sqrt(ksq) : F(0);
    }
    C2<F> r;
    r.x = fmad(na, xi.x, fmad(-a, nl.x, psi.x)) * inv;
    r.y = fmad(na, xi.y, fmad(-a, nl.y, psi.y)) * inv;
    return r;
}
