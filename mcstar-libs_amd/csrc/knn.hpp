// k-nearest-neighbour bond order on a periodic cell list: launch interface.
#pragma once

#include "mcs_common.hpp"

// device workspace for up to cap points (cell grid, binned points, cell offsets)
size_t knn_workspace_bytes(int64_t cap);

// k nearest neighbours of every point of a periodic ext[0] x ext[1] x ext[2]
// grid (spacing h) and their q4 q6 q8 w4.  Points come either as linear cell
// indices d_idx (from the peak finder) or as coordinates d_ijk (3 ints per
// point, any integers: they are taken modulo the extents).  The point count is
// *d_count when d_count is not null (capped at cap), else P_host.  Optional
// outputs: d_nbr (k per point, -1 past the count), d_out4 (q4 q6 q8 w4, NaN
// without neighbours), d_h4/d_h6 (histograms of q4/q6 over [0,1), added to).
// Every launch reads the count on the device, so the sequence can be captured.
int knn_launch(const int64_t* d_idx, const int* d_ijk, const int64_t* d_count, int64_t P_host, int64_t cap,
               const int ext[3], double h, int k, void* ws, int* d_nbr, double* d_out4, unsigned long long* d_h4,
               unsigned long long* d_h6, int nbins, hipStream_t s);
