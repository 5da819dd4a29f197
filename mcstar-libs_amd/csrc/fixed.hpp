// Fixed-plan kernels of the fused step: launch interface.
#pragma once

#include "stepper.hpp"

// A column pass of the step src -> W (kind -1 forward, +1 inverse, 0 nonlinear; src may be W)
// through a fixed-plan kernel when one matches lp; *done tells whether it launched (when false
// the caller runs the runtime-plan kernel).
int fixed_step_col(const LinePlan& lp, const void* src, void* W, const ColGeom& g, int kind, double scale, int split,
                   int64_t* d_step, hipStream_t s, bool* done);
// a column pass src -> dst (may alias) through a fixed-plan kernel when one matches lp
int fixed_col(const LinePlan& lp, const void* src, void* dst, const ColGeom& g, int dir, hipStream_t s, bool* done);
// a row pass of the boundary transform (mode 0 R2C, else C2R) with G rows per workgroup and lds
// bytes of LDS, likewise
int fixed_row(int mode, const LinePlan& lp, const void* wph, const void* src, void* dst, long nrows, int G,
              size_t lds, hipStream_t s, bool* done);
// the x pass with the update, likewise
int fixed_xupd(const StepGeom& sg, const StepPhys& ph, void* W, void* U, hipStream_t s, bool* done);
// the fixed-plan work-table column kernel of an ensemble pass (kind -1 forward, +1 inverse,
// 0 nonlinear in the given split) for lp, or null: same arguments as the runtime-plan kernel
const void* fixed_ens_col_kernel(const LinePlan& lp, int kind, int split);
// whether a fixed-plan kernel exists for lp (x pass or column passes) and its bit is set
bool fixed_available(const LinePlan& lp, bool xpass);
// the mcs_set_fixed_kernels mask
int fixed_kernel_mask();
// whether a fixed-plan kernel of the mixed form's x pass exists for lp and the x-pass bit is set
bool fixed_mixed_available(const LinePlan& lp);
// whether that kernel also has a lane-shuffle form (and shuffles are not switched off)
bool fixed_mixed_shuffle_available(const LinePlan& lp);
// whether the spectral form's fixed x kernel for lp has a lane-shuffle form (and shuffles are not off)
bool fixed_shuffle_available(const LinePlan& lp);
// the mcs_set_lane_shuffle mode
int lane_shuffle_mode();
