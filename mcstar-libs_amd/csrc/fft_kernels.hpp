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

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8,
// each XCD with its own L2).  With swz, workgroup b takes logical tile
// xcd_tile(b): every XCD gets a contiguous run of tiles, so the cache lines two
// neighbouring tiles share (rows whose length is not a multiple of the tile, as
// the nx/2+1 half spectrum) are fetched once into one L2 instead of into two.
MCS_DEV unsigned xcd_tile(unsigned b, unsigned nb, bool swz) {
    if (!swz) return b;
    const unsigned q = nb / 8, r = nb % 8, k = b % 8;
    return k * q + (k < r ? k : r) + b / 8;
}

struct mcs_fft_plan;
int fft_exec_async(mcs_fft_plan* p, const void* in, void* out, hipStream_t s);
