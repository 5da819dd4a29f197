// Host-side 1D line plans shared by the boundary FFT, the fused stepper and
// the ensemble: radix choice, engine choice and FP64-generated twiddles.
#pragma once

#include "fft_core.hpp"

#include <vector>

enum { ENGINE_REG = 0, ENGINE_GEN = 1 };

struct LinePlan {
    int L = 0;
    int prec = 32;
    int engine = ENGINE_GEN;
    int P = 0;             // register engine: points per thread
    int G = 0;             // lines (or row pairs) per workgroup chosen by the tuner; 0 = default rule
    bool nofx = false;     // the tuner found the runtime-plan kernel faster than the fixed-plan one
    bool shuf = false;     // the tuner found the fixed-plan kernel's lane-shuffle form faster
    FftPlanDev dev{};      // passed by value to kernels
    void* tw = nullptr;    // device twiddle table (C2<F>)
    size_t twcount = 0;
};

// Builds the plan (and uploads twiddles).  Returns MCS_OK or an error code.
// max_p caps the points per thread of power-of-two register plans (kernels holding several
// register blocks, like the stepper's x pass, run better with fewer)
int line_plan_build(LinePlan* lp, int L, int prec, bool allow_reg = true, int max_p = 32);
// register plan with exactly P points per thread (P a power of two; the radix-P passes
// follow a remainder pass); MCS_ERR_UNSUPPORTED if no such plan exists
int line_plan_build_reg(LinePlan* lp, int L, int prec, int P);
// the P values line_plan_build_reg accepts for L, the default first
std::vector<int> line_plan_reg_options(int L, int prec);
void line_plan_free(LinePlan* lp);
// lines per workgroup of a column pass / row pairs per workgroup of the stepper's x pass
// (lp.G when the tuner set it, else the default rule)
int col_lines_per_wg(const LinePlan& lp);
int xupd_pairs_per_wg(const LinePlan& lp);
int col_lines_per_wg_default(int L, int P, int prec);
int xupd_pairs_per_wg_default(int L, int P, int prec);
// the launch bound of every x-pass kernel (single grid, fixed plan and work table)
#define MCS_XPASS_MAX_THREADS 256
// whether G lines (column pass) or G row pairs (x pass) of a register plan with P points per
// thread fit the kernel's launch bound and 64 KB of exchange LDS (a single line or pair up to
// MCS_LDS_MAX): the rule every tile obeys, the default ones, the tuner's candidates and explicit
// plans alike
bool col_tile_ok(int L, int P, int prec, int G);
bool x_tile_ok(int L, int P, int prec, int G);
// register engine eligibility for L (without building)
int line_plan_reg_P(int L, int prec = 32);
// generic engine LDS bytes for G lines
size_t line_plan_gen_lds(const LinePlan& lp, int G);
int line_plan_gen_Lp(const LinePlan& lp);

// exp(-2 pi i k / n) for k in [0, count), generated in FP64, rounded once
int make_phase_table(void** dev, int n, int count, int prec);

// the library's maximum dynamic LDS per workgroup
#define MCS_LDS_MAX (160 * 1024)

// twiddle entries a launch of nblocks workgroups with `exchange` bytes of exchange LDS each stages
// into LDS behind that area: the plan's whole table where it fits the maximum and either adds at
// most a quarter to the workgroup's LDS (hardly fewer workgroups fit a CU) or the launch is short
// (up to two workgroups per CU, where the chain of twiddle loads, not occupancy, sets the time);
// else none (the kernel reads them from memory).  A 4096-point line with its 32 KB table behind
// 64 KB of exchange area ran 30% slower staged.
inline int twiddles_in_lds(const LinePlan& lp, size_t exchange, long nblocks) {
    const size_t esz = lp.prec == 64 ? 16 : 8;
    const size_t tw = lp.twcount * esz;
    if (exchange + tw > (size_t)MCS_LDS_MAX) return 0;
    return (4 * tw <= exchange || nblocks <= 512) ? (int)lp.twcount : 0;
}
