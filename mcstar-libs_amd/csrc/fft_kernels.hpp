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
// each XCD with its own L2).  Workgroup b takes logical tile xcd_tile(b) of a
// grid of nouter rows of tilesx tiles (tile = outer * tilesx + x tile):
//   order 0: b itself;
//   order 1: every XCD gets a contiguous run of tiles, so the cache lines two
//            neighbouring tiles share (rows whose length is not a multiple of the
//            tile, as the nx/2+1 half spectrum) are fetched once into one L2
//            instead of into two;
//   order 2: whole rows of tiles go to one XCD and eight neighbouring rows run
//            side by side on the eight XCDs (XCD k takes rows 8j + k in turn), so
//            shared lines stay in one L2 while the XCDs read nearby addresses; the
//            nouter % 8 last rows run in order 0.
MCS_DEV unsigned xcd_tile(unsigned b, unsigned nb, int order, unsigned tilesx, unsigned nouter) {
    if (order == 1) {
        const unsigned q = nb / 8, r = nb % 8, k = b % 8;
        return k * q + (k < r ? k : r) + b / 8;
    }
    if (order == 2 && b < (nouter & ~7u) * tilesx) {
        const unsigned i = b / 8;
        return ((i / tilesx) * 8 + b % 8) * tilesx + i % tilesx;
    }
    return b;
}

// the tile order of a column pass with G lines per workgroup and esz-byte elements:
// order 1 along short strides (y), order 2 along long strides (z) where neighbouring
// tiles share cache lines, order 0 along long strides where they do not
int col_tile_order(const ColGeom& g, int G, long esz);

struct mcs_fft_plan;
int fft_exec_async(mcs_fft_plan* p, const void* in, void* out, hipStream_t s);
