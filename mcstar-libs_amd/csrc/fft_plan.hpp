// Host-side 1D line plans shared by the boundary FFT, the fused stepper and
// the ensemble: radix choice, engine choice and FP64-generated twiddles.
#pragma once

#include "fft_core.hpp"

enum { ENGINE_REG = 0, ENGINE_GEN = 1 };

struct LinePlan {
    int L = 0;
    int prec = 32;
    int engine = ENGINE_GEN;
    int P = 0;             // register engine: points per thread
    FftPlanDev dev{};      // passed by value to kernels
    void* tw = nullptr;    // device twiddle table (C2<F>)
    size_t twcount = 0;
};

// Builds the plan (and uploads twiddles).  Returns MCS_OK or an error code.
// max_p caps the points per thread of power-of-two register plans (kernels holding several
// register blocks, like the stepper's x pass, run better with fewer)
int line_plan_build(LinePlan* lp, int L, int prec, bool allow_reg = true, int max_p = 32);
void line_plan_free(LinePlan* lp);
// register engine eligibility for L (without building)
int line_plan_reg_P(int L, int prec = 32);
// generic engine LDS bytes for G lines
size_t line_plan_gen_lds(const LinePlan& lp, int G);
int line_plan_gen_Lp(const LinePlan& lp);

// exp(-2 pi i k / n) for k in [0, count), generated in FP64, rounded once
int make_phase_table(void** dev, int n, int count, int prec);

// the library's maximum dynamic LDS per workgroup
#define MCS_LDS_MAX (160 * 1024)
