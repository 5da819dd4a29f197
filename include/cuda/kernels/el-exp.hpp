// Exponential time differencing for the non-conserved relaxation (cuda/kernels/el-exp.hpp):
// call_kernel_init_el_op fills kop with (E, S), E = exp(-L' dt), S = (E - 1) / L';
// call_kernel_update_el steps kpsi <- kpsi E + knonlin S and returns max |change|.
// A simulator includes this header instead of el.hpp (the names are the same).
#pragma once

#include "common.hpp"

inline void call_kernel_init_el_op(GPUArray& kop, Float dt, Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_init_el_op(kop.ptr_void(), kop.ext_ptr(), (double)dt, (double)eps, l, Array::prec_of<Float>()));
}

inline Float call_kernel_update_el(GPUArray& kop, GPUArray& kpsi, GPUArray& knonlin, Float dt, Float eps, Float3 lens) {
    (void)dt;
    (void)eps;
    (void)lens;  // folded into kop
    double mx = 0;
    CUERR(mcs_update_el_etd(kop.ptr_void(), kpsi.ptr_void(), knonlin.ptr_void(), kpsi.ext_ptr(),
                            Array::prec_of<Float>(), &mx));
    return (Float)mx;
}
