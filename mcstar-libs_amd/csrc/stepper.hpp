// Fused-step building blocks shared by the single-grid stepper and the ensemble.
#pragma once

#include "pass_bodies.hpp"

#include <cmath>
#include <functional>

struct StepGeom {
    int nx = 0, ny = 0, nz = 0, L = 0, prec = 32;
    bool is3d = false;
    // MCS_FORM_SPECTRAL: the state is U = FFT3(u), the x pass reads and writes U and W;
    // MCS_FORM_MIXED: the state is S = IFFT_x(U) / L, the x pass reads W and S and writes S
    // (ensembles run the spectral form)
    int form = MCS_FORM_SPECTRAL;
    LinePlan px, py, pz;    // x: nx/2 (packed), y: ny, z: nz
    void* wph = nullptr;    // exp(-2 pi i p / nx), p in [0, nx/2)
    int2* pairs = nullptr;  // canonical (row, mirror row) pairs
    int npairs = 0;
    int64_t* d_step = nullptr;  // device step counter (Philox draw number)
};

struct StepPhys {
    double h[3];
    double eps, dt, temp;
    uint64_t seed;
    int split;
    double mass;  // sum of psi over the grid (the k = 0 mode), FP64
};

int step_geom_build(StepGeom* sg, const int ext[3], int prec);
void step_geom_free(StepGeom* sg);
int step_import(const StepGeom& sg, const StepPhys& ph, const void* psi, void* W, void* U, double* mass, hipStream_t s);
int step_export(const StepGeom& sg, const void* U, void* psi, void* tmp, hipStream_t s);
int step_run(const StepGeom& sg, const StepPhys& ph, void* W, void* U, int nsteps, hipStream_t s);
// W as the last pass of a step leaves it, from the state U (after an import or a restore)
int step_prime(const StepGeom& sg, const StepPhys& ph, void* W, const void* U, hipStream_t s);
// converts the state U in place to the given form and sets sg->form
int step_convert_form(StepGeom* sg, void* U, int form, hipStream_t s);
int launch_row_packed(const LinePlan& lp, const void* src, void* dst, long nrows, int dir, hipStream_t s);

double pfc_noise_amp(const int ext[3], const double h[3], double dt, double temp);

template <typename F> inline XArgs<F> make_xargs(const StepGeom& sg, const StepPhys& ph) {
    XArgs<F> ar;
    ar.nx = sg.nx; ar.ny = sg.ny; ar.nz = sg.nz; ar.L = sg.L;
    ar.dkx = (F)(2.0 * M_PI / (sg.nx * ph.h[0]));
    ar.dky = (F)(2.0 * M_PI / (sg.ny * ph.h[1]));
    ar.dkz = (F)(2.0 * M_PI / (sg.nz * ph.h[2]));
    ar.dt = (F)ph.dt;
    ar.lin0 = (F)(ph.split == MCS_SPLIT_REFERENCE ? 1.0 - ph.eps : -ph.eps);
    const int ext[3] = {sg.nx, sg.ny, sg.nz};
    ar.amp = (F)pfc_noise_amp(ext, ph.h, ph.dt, ph.temp);
    ar.key0 = (uint32_t)ph.seed;
    ar.key1 = (uint32_t)(ph.seed >> 32);
    ar.noise = ph.temp > 0;
    ar.mass = ph.mass;
    return ar;
}

int launch_xupd(const StepGeom& sg, const StepPhys& ph, void* W, void* U, hipStream_t s);
// the x pass of the mixed form: reads W and the state S, writes S
int launch_xupd_mixed(const StepGeom& sg, const StepPhys& ph, const void* W, void* Sx, hipStream_t s);
// row pairs per workgroup of that pass (three rows of LDS per pair)
int xupd_mixed_pairs_per_wg(const LinePlan& lp);
// the x pass of the stepper's form (W, state)
inline int launch_xpass(const StepGeom& sg, const StepPhys& ph, void* W, void* state, hipStream_t s) {
    return sg.form == MCS_FORM_MIXED ? launch_xupd_mixed(sg, ph, W, state, s) : launch_xupd(sg, ph, W, state, s);
}
// the factor that takes the last inverse column transform of a step to psi: 2 / N in the
// spectral form, 2 L / N = 1 / (ny nz) in the mixed form
inline double step_nl_scale(const StepGeom& sg) {
    const double n = (double)sg.nx * sg.ny * sg.nz;
    return sg.form == MCS_FORM_MIXED ? 2.0 * sg.L / n : 2.0 / n;
}
// the y, z passes of a step after the x pass: W <- FFT_yz(N(psi)) from the x-real rows in src
// (W itself in the spectral form, the state S in the mixed form); d_step as for the column passes
int step_nonlinear_half(const StepGeom& sg, const StepPhys& ph, const void* src, void* W, int64_t* d_step,
                        hipStream_t s);
// the nonlinear column pass src -> W (src may be W)
int launch_col_nl(const LinePlan& lp, const void* src, void* W, const ColGeom& g, double scale, int split,
                  int64_t* d_step, hipStream_t s);
// a column pass of the step on W in place: kind -1 forward, +1 inverse, 0 nonlinear
int launch_step_col(const LinePlan& lp, void* W, const ColGeom& g, int kind, double scale, int split,
                    int64_t* d_step, hipStream_t s);
// the same from src into W (src may be W)
int launch_step_col_to(const LinePlan& lp, const void* src, void* W, const ColGeom& g, int kind, double scale,
                       int split, int64_t* d_step, hipStream_t s);
ColGeom step_geom_y(const StepGeom& sg);
ColGeom step_geom_z(const StepGeom& sg);

// autotuner of a single-grid stepper: picks the register plan and tile shape of every pass
// for this shape and precision by timing the stepper's own pass kernels, fixed-plan ones
// included (cached per process); a no-op when tuning is switched off.  Overwrites W and U.
int step_geom_tune(StepGeom* sg, const StepPhys& ph, void* W, void* U, hipStream_t s);
// the autotuner over caller-supplied pass launchers: each runs one pass with a candidate plan
// (run_y the y pass of a 3D step, or the nonlinear y pass of a 2D one; run_z the z pass);
// fixed (MCS_FIXED_* bits): for those pass kinds, also time the fixed-plan kernel of a candidate
// where one exists (the launcher runs it unless lp.nofx).  Choices are cached under
// (extents, precision, tag); the caller's kernels overwrite their buffers.
using PassRun = std::function<int(const LinePlan&)>;
int step_geom_tune_runs(StepGeom* sg, int tag, int fixed, const PassRun& run_x, const PassRun& run_y,
                        const PassRun& run_z, hipStream_t s, int xform = MCS_FORM_SPECTRAL);
// {P_x, pairs per workgroup, P_y, lines per workgroup, P_z, lines per workgroup}
void step_geom_plan_info(const StepGeom& sg, int info[6]);
// installs an explicit plan in that layout after checking it against the candidate lists
int step_geom_apply_plan(StepGeom* sg, const int plan[6]);
// an explicit plan against the grid's axis lengths: each P one of line_plan_reg_options, each tile
// within its kernel's launch bound and LDS (host only, before anything is allocated or launched)
int step_plan_validate(const int ext[3], int prec, const int plan[6]);
