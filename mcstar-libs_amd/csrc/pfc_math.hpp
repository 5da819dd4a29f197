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
// lin0 = -eps (north-star split, nl = psi^3) or 1 - eps (reference split, nl = psi^3 - psi)
template <typename F>
MCS_HD C2<F> pfc_mode_update(C2<F> psi, C2<F> nl, C2<F> xi, F ksq, F dt, F lin0, F amp) {
    const F om = F(1) - ksq;
    const F inv = F(1) / (F(1) + dt * ksq * (lin0 + om * om));
    const F a = dt * ksq;
    const F na = amp * sqrt(ksq);
    C2<F> r;
    r.x = (psi.x - a * nl.x + na * xi.x) * inv;
    r.y = (psi.y - a * nl.y + na * xi.y) * inv;
    return r;
}
