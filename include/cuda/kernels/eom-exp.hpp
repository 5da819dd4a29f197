// Exponential time differencing for the conserved PFC equation (cuda/kernels/eom-exp.hpp):
// call_kernel_init_eom_op fills kop with (E, S) per mode, E = exp(-k^2 L' dt),
// S = (E - 1) / (k^2 L'); call_kernel_update_eom then steps
//   kpsi <- kpsi E + k^2 knonlin S - namp sqrt(N) |k| knoise S.
// A simulator includes this header instead of eom.hpp (the names are the same).
#pragma once

#include "common.hpp"

inline void call_kernel_init_eom_op(GPUArray& kop, Float dt, Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_init_eom_op(kop.ptr_void(), kop.ext_ptr(), (double)dt, (double)eps, l, Array::prec_of<Float>()));
}

inline void call_kernel_update_eom(GPUArray& kop, GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise, Float namp,
                                   Float dt, Float eps, Float3 lens) {
    (void)dt;  // folded into kop
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_etd(kop.ptr_void(), kpsi.ptr_void(), knonlin.ptr_void(), knoise.ptr_void(), 0,
                             kpsi.ext_ptr(), (double)namp, (double)eps, l, Array::prec_of<Float>()));
}

inline void call_kernel_update_eom(GPUArray& kop, GPUArray& kpsi, GPUArray& knonlin, Float namp, Float dt, Float eps,
                                   Float3 lens) {
    (void)dt;
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_etd(kop.ptr_void(), kpsi.ptr_void(), knonlin.ptr_void(), nullptr, 0, kpsi.ext_ptr(),
                             (double)namp, (double)eps, l, Array::prec_of<Float>()));
}

inline void call_kernel_update_eom_with_fen(GPUArray& kop, GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise,
                                            Float namp, Float dt, Float eps, Float3 lens) {
    (void)dt;
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_etd(kop.ptr_void(), kpsi.ptr_void(), knonlin.ptr_void(), knoise.ptr_void(), 1,
                             kpsi.ext_ptr(), (double)namp, (double)eps, l, Array::prec_of<Float>()));
}

// eom-exp.hpp:71-100: the free-energy density, as in eom.hpp
inline float call_kernel_calc_fen(GPUArray& psi, GPUArray& lpsi, Float3 hh) {
    const double h[3] = {(double)hh.x, (double)hh.y, (double)hh.z};
    double f = 0;
    CUERR(mcs_calc_fen(psi.ptr_void(), lpsi.ptr_void(), psi.ext_ptr(), h, Array::prec_of<Float>(), &f));
    return (float)f;
}
