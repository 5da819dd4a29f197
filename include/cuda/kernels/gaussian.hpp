// k-space Gaussian filter of the kept API (cuda/kernels/gaussian.hpp:22-67).
// call_kernel_gaussian_init fills each half-spectrum mode with (e, 1 - e),
// e = exp(-(sx^2 kx^2 + sy^2 ky^2 + sz^2 kz^2) / 2), s_i = 1.1 divs_per_sigma h_i:
// the real part low-passes, the imaginary part high-passes.  kmult_real / _imag
// multiply a spectrum in place by that part of a second one.
#pragma once

#include <cassert>

#include "common.hpp"

inline void call_kernel_gaussian_init(GPUArray& arr_kpsi, Float3 lens, Float3 hh, Float divs_per_sigma) {
    const double l[3] = {(double)lens.x, (double)lens.y, (double)lens.z};
    const double h[3] = {(double)hh.x, (double)hh.y, (double)hh.z};
    CUERR(mcs_gaussian_init(arr_kpsi.ptr_void(), arr_kpsi.ext_ptr(), l, h, (double)divs_per_sigma,
                            Array::prec_of<Float>()));
}

inline void call_kernel_kmult_real(GPUArray& arr_1, GPUArray& arr_2) {
    assert(arr_1.real_vext() == arr_2.real_vext() && "array dimensions do not match");
    CUERR(mcs_kmult(arr_1.ptr_void(), arr_2.ptr_void(), arr_1.ext_ptr(), 0, Array::prec_of<Float>()));
}

inline void call_kernel_kmult_imag(GPUArray& arr_1, GPUArray& arr_2) {
    assert(arr_1.real_vext() == arr_2.real_vext() && "array dimensions do not match");
    CUERR(mcs_kmult(arr_1.ptr_void(), arr_2.ptr_void(), arr_1.ext_ptr(), 1, Array::prec_of<Float>()));
}
