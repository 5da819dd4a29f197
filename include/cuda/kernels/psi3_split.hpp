// call_kernel_psi3_split (cuda/kernels/psi3_split.hpp:15-26): in place,
// a <- (mul a)^3 - mul a - mu [+ pot], mul = 1/N when norm is set.
#pragma once

#include "common.hpp"

inline void call_kernel_psi3_split(GPUArray& arr, GPUArray* pot = NULL, Float mu = Float(0), bool norm = false) {
    CUERR(mcs_psi3_split(arr.ptr_void(), pot ? pot->ptr_void() : nullptr, (double)mu, norm ? 1 : 0, arr.ext_ptr(),
                         Array::prec_of<Float>()));
}
