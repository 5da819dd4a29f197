// Semi-implicit conserved update (cuda/kernels/eom.hpp:60-86):
//   kpsi <- (kpsi - dt k^2 knonlin + dt namp sqrt(N) |k| knoise) / (1 + dt k^2 (1 - eps + (1 - k^2)^2))
// with knonlin = FFT(psi^3 - psi - mu) from call_kernel_psi3_split.  The
// overload without knoise is the noise-free update.  call_kernel_update_pfc is
// the same step with the noise drawn in registers from Philox (no knoise array).
#pragma once

#include "common.hpp"

inline void call_kernel_update_eom(GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise, Float namp, Float dt, Float eps,
                                   Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom(kpsi.ptr_void(), knonlin.ptr_void(), knoise.ptr_void(), kpsi.ext_ptr(), (double)namp,
                         (double)dt, (double)eps, l, Array::prec_of<Float>()));
}

inline void call_kernel_update_eom(GPUArray& kpsi, GPUArray& knonlin, Float namp, Float dt, Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom(kpsi.ptr_void(), knonlin.ptr_void(), nullptr, kpsi.ext_ptr(), (double)namp, (double)dt,
                         (double)eps, l, Array::prec_of<Float>()));
}

// temp = T of eta_k = sqrt(2 T dt / dV) sqrt(N) |k| xi_k; draw = step index of the noise sequence
inline void call_kernel_update_pfc(GPUArray& kpsi, GPUArray& knl, Float dt, Float eps, Float3 lens, double temp,
                                   unsigned long long seed, unsigned draw, int split = MCS_SPLIT_REFERENCE) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_pfc(kpsi.ptr_void(), knl.ptr_void(), kpsi.ext_ptr(), l, (double)dt, (double)eps, temp, seed, draw,
                         split, Array::prec_of<Float>()));
}

// call_kernel_update_eom_with_fen (cuda/kernels/eom.hpp:103-109): the update, and knoise <- lk kpsi
// (the linear operator of the free energy on the old kpsi); C2R of knoise then feeds calc_fen
inline void call_kernel_update_eom_with_fen(GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise, Float namp, Float dt,
                                            Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_with_fen(kpsi.ptr_void(), knonlin.ptr_void(), knoise.ptr_void(), kpsi.ext_ptr(),
                                  (double)namp, (double)dt, (double)eps, l, Array::prec_of<Float>()));
}

// call_kernel_calc_fen (eom.hpp:130-141): per-cell free energy written into lpsi, the total
// (times the cell volume) returned as float like the reference; the sum itself is FP64
inline float call_kernel_calc_fen(GPUArray& psi, GPUArray& lpsi, Float3 hh) {
    const double h[3] = {(double)hh.x, (double)hh.y, (double)hh.z};
    double f = 0;
    CUERR(mcs_calc_fen(psi.ptr_void(), lpsi.ptr_void(), psi.ext_ptr(), h, Array::prec_of<Float>(), &f));
    return (float)f;
}

// call_kernel_update_eom_hpfcnoise (eom.hpp:47-58): 2D update with three noise arrays
inline void call_kernel_update_eom_hpfcnoise(GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise1, GPUArray& knoise2,
                                             GPUArray& knoise3, Float namp, Float dt, Float eps, Float3 lens) {
    if (kpsi.real_vext().z != 1) {
        std::cerr << "the hpfc noise update is 2D only" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_hpfc(kpsi.ptr_void(), knonlin.ptr_void(), knoise1.ptr_void(), knoise2.ptr_void(),
                              knoise3.ptr_void(), kpsi.ext_ptr(), (double)namp, (double)dt, (double)eps, l,
                              Array::prec_of<Float>()));
}

// call_kernel_update_eom_py (eom.hpp:173-180): the polynomial linear operator
inline void call_kernel_update_eom_py(GPUArray& kpsi, GPUArray& knonlin, GPUArray& knoise, Float namp, Float dt,
                                      Float eps, Float3 lens) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    CUERR(mcs_update_eom_poly(kpsi.ptr_void(), knonlin.ptr_void(), knoise.ptr_void(), kpsi.ext_ptr(), (double)namp,
                              (double)dt, (double)eps, l, Array::prec_of<Float>()));
}

// FP64 free energy of a real field: dV sum[psi/2 (-eps + (1+lap)^2) psi + psi^4/4]
inline double calc_free_energy(GPUArray& psi, Float3 hh, Float eps) {
    const double h[3] = {(double)hh.x, (double)hh.y, (double)hh.z};
    double mean = 0, f = 0;
    CUERR(mcs_field_stats(psi.ptr_void(), psi.ext_ptr(), h, (double)eps, Array::prec_of<Float>(), &mean, &f));
    return f;
}
