// Host-side index helpers of the kept API (cuda/kernels/common.hpp:10-36).
// The consumer defines FLOAT and typedefs Float, Float2, Float3, GPUArray,
// CPUArray before including the kernel wrappers, as with the original.
#pragma once

#include <cstdlib>

inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
inline unsigned int div_up(int a, int b) { return (unsigned)((a + b - 1) / b); }
// non-negative remainder of n modulo m
inline unsigned int posrem(int n, int m) { return (unsigned)(((n % m) + m) % m); }
inline int calc_idx(int i0, int i1, int i2, int3 dims) { return (i2 * dims.y + i1) * dims.x + i0; }
