// Launch helpers shared by the boundary FFT, the fused stepper and the ensemble.
#pragma once

#include "fft_plan.hpp"

// A column pass transforms every line along y or z of a complex array
// [nz][ny][nxc]: line (outer, x) starts at outer*os + x and has stride ls.
struct ColGeom {
    int nxc;     // x extent (columns)
    long ls;     // element stride along the transformed axis
    int nouter;  // number of outer indices
    long os;     // stride between outer indices
};

int launch_col(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, int dir, hipStream_t s);

struct mcs_fft_plan;
int fft_exec_async(mcs_fft_plan* p, const void* in, void* out, hipStream_t s);
