// call_kernel_update_el (cuda/kernels/el.hpp:27-45): non-conserved relaxation of the PFC
// free energy, kpsi <- (kpsi - dt knonlin) / (1 + dt (2 - eps - 2 k^2 + k^4)), returning the
// largest change of a mode, max |kpsi_new - kpsi_old| (the convergence measure).
#pragma once

#include "common.hpp"

inline Float call_kernel_update_el(GPUArray& kpsi, GPUArray& knonlin, Float dt, Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    double mx = 0;
    CUERR(mcs_update_el(kpsi.ptr_void(), knonlin.ptr_void(), kpsi.ext_ptr(), (double)dt, (double)eps, l,
                        Array::prec_of<Float>(), &mx));
    return (Float)mx;
}
