// call_kernel_norm (cuda/kernels/norm.hpp:5-10): a <- a / N after an unnormalised C2R.
#pragma once

#include "common.hpp"

inline void call_kernel_norm(GPUArray& arr) { CUERR(mcs_norm(arr.ptr_void(), arr.ext_ptr(), Array::prec_of<Float>())); }
