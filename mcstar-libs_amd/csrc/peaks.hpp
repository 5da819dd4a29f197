// Peak finder launch interface (device outputs, caller-provided workspace).
#pragma once

#include "mcs_common.hpp"

size_t peaks_workspace_bytes(const int ext[3]);
// writes up to cap ordered linear indices (and values) and the total count, all on the device
int findpeaks_launch(const void* real, const int ext[3], double ths, int64_t* d_idx, double* d_vals, int64_t cap,
                     int64_t* d_total, void* workspace, int prec, hipStream_t s);
// argument checks shared by the peak finder's entry points
int findpeaks_check(const void* real, const int ext[3], int64_t* count, int64_t cap, int prec);
