// call_kernel_findpeaks_2d (cuda/kernels/findpeaks_2d.hpp:29-56): strict
// periodic 8-neighbour maxima >= ths of a 2D array, "%f %f %f %f" lines.
#pragma once

#include "findpeaks.hpp"

inline void call_kernel_findpeaks_2d(GPUArray& input, std::string peaks_outfn, Float3 hh, Float3 len, Float ths) {
    (void)len;
    const int3 s = input.real_vext();
    if (!(s.x > 1 && s.y > 1 && s.z == 1)) {
        std::cerr << "input not a 2d array" << std::endl;
        return;
    }
    std::vector<Max> m = mcs_collect_peaks(input, ths);
    std::cerr << "found " << m.size() << " maxima" << std::endl;
    mcs_write_peaks(peaks_outfn, m, hh);
}
