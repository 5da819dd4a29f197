// Kept host API: strip-confined square-lattice potential (cuda/potentials.hpp:1-20)
// for call_kernel_psi3_split(arr, &pot, ...).  Needs the consumer's Float3 and
// CPUArray typedefs, as the original does.
#pragma once

#include <cassert>
#include <cmath>
#include <iostream>

inline double sm_square(double x, double y) { return (std::cos(x) + std::cos(y) + 2.0) / 4.0; }

// pot(x, y) = amp * sm_square(2 pi x / sigma, 2 pi y / sigma)^power * w(y), with the
// strip window w(y) = (1 + tanh(y + sigma/2)) (1 + tanh(sigma/2 - y)) / 4 centred in the box
inline void init_pot_line(CPUArray& arr, const Float3 l, const Float3 h, const int power, const double sigma,
                          const double amp) {
    const int3 s = arr.real_vext();
    assert(s.z == 1 && s.y > 1 && s.x > 1 && "array is not 2d");
    const double f = 2.0 * M_PI / sigma;
    Float* p = arr.ptr_real();
    for (int i1 = 0; i1 < s.y; ++i1) {
        const double y = i1 * h.y - l.y / 2.0;
        const double w = 0.25 * (std::tanh(y + sigma / 2) + 1.0) * (std::tanh(sigma / 2 - y) + 1.0);
        for (int i0 = 0; i0 < s.x; ++i0) {
            const double x = i0 * h.x - l.x / 2.0;
            p[i1 * s.x + i0] = (Float)(amp * std::pow(sm_square(f * x, f * y), power) * w);
        }
    }
    std::cout << "\"init_pot_line\" potential initialized" << std::endl;
}
