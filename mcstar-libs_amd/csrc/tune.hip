// Autotuner of the fused step.  For one grid shape and precision it times, on
// the caller's own buffers, every candidate register plan (points per thread P,
// hence the radix sequence and the number of LDS exchanges) and tile shape
// (lines or row pairs per workgroup, hence the global segment width and the
// LDS footprint) of each pass, and keeps the fastest.  Where a fixed-plan
// kernel exists for a candidate, it is timed against the runtime-plan kernel of
// the same plan and the faster one is kept.  Choices are cached per
// (nx, ny, nz, precision, tag) for the life of the process: steppers of one
// shape share plans (tagged noisy or not), and ensemble groups of one shape and
// member count share theirs, timed on the group's own work-table kernels.  A
// stepper and an ensemble replica agree bit for bit when they run the same plan
// (the default plans with tuning off, or an explicit plan).  The pass kernels
// overwrite the buffers, so tuning runs before a field is loaded.
#include "fixed.hpp"
#include "stepper.hpp"

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace {

struct PassChoice {
    int P = 0, G = 0;
    double ms = 0;
    bool nofx = false;  // runtime-plan kernel (where a fixed-plan one exists and was slower)
    bool shuf = false;  // the fixed-plan kernel's lane-shuffle form (where it has one and was faster)
};
struct StepChoice {
    PassChoice x, y, z;
};

std::mutex g_mu;
std::map<std::tuple<int, int, int, int, int>, StepChoice> g_cache;
int g_enabled = 1;

// milliseconds of one launch: one warm-up, then three batches of enough launches for about a
// millisecond each; the fastest batch's mean (a batch slowed by the rest of the system does not
// decide the choice)
template <typename Fn> double time_launch(hipStream_t s, Fn&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
    double ms = -1;
    float el = 0;
    int reps = 1;
    if (launch() == MCS_OK && hipEventRecord(a, s) == hipSuccess && launch() == MCS_OK &&
        hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
        hipEventElapsedTime(&el, a, b) == hipSuccess) {
        reps = std::max(3, std::min(50, (int)(1.0 / std::max(el, 1e-3f))));
        bool ok = true;
        for (int batch = 0; batch < 3 && ok; ++batch) {
            ok = hipEventRecord(a, s) == hipSuccess;
            for (int i = 0; i < reps && ok; ++i) ok = launch() == MCS_OK;
            ok = ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                 hipEventElapsedTime(&el, a, b) == hipSuccess;
            if (ok && (ms < 0 || el / reps < ms)) ms = el / reps;
        }
        if (!ok) ms = -1;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

size_t esz(int prec) { return prec == 64 ? 16 : 8; }

// tile shapes of a column pass: the default tile first, then G lines of L points for G in
// {4, 8, 16, 32}, every one within the launch bound and 64 KB of LDS (col_tile_ok)
std::vector<int> col_tiles(int L, int P, int prec) {
    std::vector<int> out;
    const int def = col_lines_per_wg_default(L, P, prec);
    if (col_tile_ok(L, P, prec, def)) out.push_back(def);
    for (int G : {4, 8, 16, 32})
        if (G != def && col_tile_ok(L, P, prec, G)) out.push_back(G);
    return out;
}

// tile shapes of the x pass: the default tile first, then G row pairs (powers of two) of 64..256
// threads, every one within the launch bound and two rows of LDS per pair within 64 KB (x_tile_ok)
std::vector<int> x_tiles(int L, int P, int prec) {
    const int S = L / P;
    std::vector<int> out;
    const int def = xupd_pairs_per_wg_default(L, P, prec);
    if (x_tile_ok(L, P, prec, def)) out.push_back(def);
    for (int G = 1; G * S <= MCS_XPASS_MAX_THREADS; G *= 2)
        if (G != def && (G * S >= 64 || (G == 1 && S >= 64)) && x_tile_ok(L, P, prec, G)) out.push_back(G);
    return out;
}

int rebuild(LinePlan* lp, int L, int prec, int P) {
    if (lp->P == P && lp->engine == ENGINE_REG) return MCS_OK;
    LinePlan np;
    int st = line_plan_build_reg(&np, L, prec, P);
    if (st) return st;
    line_plan_free(lp);
    *lp = np;
    return MCS_OK;
}

// best (P, G) of one pass; `run` launches the pass with the candidate plan; with_fixed: also time
// the fixed-plan kernel of each candidate that has one
template <typename Run>
int tune_pass(int L, int prec, int maxP, bool xpass, bool with_fixed, int xform, hipStream_t s, Run&& run,
              PassChoice* best) {
    best->ms = -1;
    const std::vector<int> opts = line_plan_reg_options(L, prec);
    for (int P : opts) {
        if (P > maxP && P != opts[0]) continue;  // the default plan is always a candidate
        LinePlan cand;
        if (line_plan_build_reg(&cand, L, prec, P) != MCS_OK) continue;
        const bool fx = with_fixed && (xpass && xform == MCS_FORM_MIXED ? fixed_mixed_available(cand)
                                                                         : fixed_available(cand, xpass));
        const bool sh = fx && xpass && (xform == MCS_FORM_MIXED ? fixed_mixed_shuffle_available(cand)
                                                                 : fixed_shuffle_available(cand));
        for (int G : xpass ? x_tiles(L, P, prec) : col_tiles(L, P, prec)) {
            cand.G = G;
            // the fixed-plan kernel (when one exists; with and without lane shuffles where it has
            // them) and the runtime-plan kernel of this plan
            for (int v = fx ? 0 : 2; v <= 2; ++v) {
                if (v == 1 && !sh) continue;
                cand.nofx = v == 2;
                cand.shuf = v == 1;
                const double ms = time_launch(s, [&]() { return run(cand); });
                if (ms > 0 && (best->ms < 0 || ms < best->ms)) *best = PassChoice{P, G, ms, cand.nofx, cand.shuf};
            }
        }
        line_plan_free(&cand);
    }
    (void)hipGetLastError();
    MCS_CHECK(best->ms > 0, MCS_ERR_UNSUPPORTED, "no tunable plan ran");
    return MCS_OK;
}

int apply(StepGeom* sg, const StepChoice& c) {
    int st = rebuild(&sg->px, sg->L, sg->prec, c.x.P);
    if (!st) st = rebuild(&sg->py, sg->ny, sg->prec, c.y.P);
    if (!st && sg->is3d) st = rebuild(&sg->pz, sg->nz, sg->prec, c.z.P);
    if (st) return st;
    sg->px.G = c.x.G;
    sg->py.G = c.y.G;
    sg->px.nofx = c.x.nofx;
    sg->py.nofx = c.y.nofx;
    sg->px.shuf = c.x.shuf;
    if (sg->is3d) {
        sg->pz.G = c.z.G;
        sg->pz.nofx = c.z.nofx;
    }
    return MCS_OK;
}

}  // namespace

extern "C" int mcs_set_autotune(int on) {
    std::lock_guard<std::mutex> l(g_mu);
    g_enabled = on ? 1 : 0;
    return MCS_OK;
}

extern "C" int mcs_get_autotune(int* on) {
    MCS_CHECK(on, MCS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> l(g_mu);
    *on = g_enabled;
    return MCS_OK;
}

int step_geom_tune_runs(StepGeom* sg, int tag, int fixed, const PassRun& run_x, const PassRun& run_y,
                        const PassRun& run_z, hipStream_t s, int xform) {
    const auto key = std::make_tuple(sg->nx, sg->ny, sg->nz, sg->prec, tag);
    {
        std::lock_guard<std::mutex> l(g_mu);
        if (!g_enabled) return MCS_OK;
        auto it = g_cache.find(key);
        if (it != g_cache.end()) return apply(sg, it->second);
    }
    const int prec = sg->prec;
    const int maxX = prec == 64 ? 8 : 16, maxC = prec == 64 ? 16 : 32;
    // about 20 ms of the default plans first, so that every candidate is timed at the clocks of
    // a running step (a pass runs slower during the first steps of a run)
    const auto t0 = std::chrono::steady_clock::now();
    while (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() < 0.02) {
        if (run_x(sg->px) != MCS_OK || run_y(sg->py) != MCS_OK || (sg->is3d && run_z(sg->pz) != MCS_OK)) break;
        if (hipStreamSynchronize(s) != hipSuccess) break;
    }
    (void)hipGetLastError();
    StepChoice c;
    const bool fx = (fixed & MCS_FIXED_XPASS) != 0, fc = (fixed & MCS_FIXED_COLUMNS) != 0;
    int st = tune_pass(sg->L, prec, maxX, true, fx, xform, s, run_x, &c.x);
    if (!st) st = tune_pass(sg->ny, prec, maxC, false, fc, xform, s, run_y, &c.y);
    if (!st && sg->is3d) st = tune_pass(sg->nz, prec, maxC, false, fc, xform, s, run_z, &c.z);
    if (st) return st;
    MCS_HIP_TRY(hipStreamSynchronize(s));
    {
        std::lock_guard<std::mutex> l(g_mu);
        auto it = g_cache.find(key);  // another thread may have tuned the same shape meanwhile
        if (it != g_cache.end()) c = it->second;
        else g_cache.emplace(key, c);
    }
    return apply(sg, c);
}

int step_geom_tune(StepGeom* sg, const StepPhys& ph_in, void* W, void* U, hipStream_t s) {
    StepPhys ph = ph_in;
    ph.mass = 0;
    const double scale = step_nl_scale(*sg);
    const ColGeom gy = step_geom_y(*sg), gz = step_geom_z(*sg);
    const bool is3d = sg->is3d;
    const int form = sg->form;
    // noisy and noise-free steps run different x-pass kernels (the draw costs registers), and
    // the two forms different x passes, so each tunes on its own; the 2D nonlinear pass of the
    // mixed form reads the state and writes W
    const int tag = (ph.temp > 0 ? 1 : 0) + 2 * form;
    const void* rows = form == MCS_FORM_MIXED ? U : W;
    return step_geom_tune_runs(
        sg, tag, MCS_FIXED_COLUMNS | MCS_FIXED_XPASS,
        [&](const LinePlan& lp) {
            StepGeom g = *sg;
            g.px = lp;
            return launch_xpass(g, ph, W, U, s);
        },
        [&](const LinePlan& lp) {
            return is3d ? launch_step_col(lp, W, gy, -1, scale, ph.split, nullptr, s)
                        : launch_col_nl(lp, rows, W, gy, scale, ph.split, nullptr, s);
        },
        [&](const LinePlan& lp) { return launch_col_nl(lp, W, W, gz, scale, ph.split, nullptr, s); }, s, form);
}

int step_plan_validate(const int ext[3], int prec, const int plan[6]) {
    MCS_CHECK(plan, MCS_ERR_INVALID, "null plan");
    const bool is3d = ext[2] > 1;
    const int L[3] = {ext[0] / 2, ext[1], ext[2]};
    for (int i = 0; i < (is3d ? 3 : 2); ++i) {
        const int P = plan[2 * i], G = plan[2 * i + 1];
        const std::vector<int> opts = line_plan_reg_options(L[i], prec);
        MCS_CHECK(std::find(opts.begin(), opts.end(), P) != opts.end(), MCS_ERR_UNSUPPORTED,
                  "points per thread not available for this axis length");
        MCS_CHECK(i == 0 ? x_tile_ok(L[i], P, prec, G) : col_tile_ok(L[i], P, prec, G), MCS_ERR_UNSUPPORTED,
                  "tile shape outside the launch and LDS limits of this plan");
    }
    return MCS_OK;
}

int step_geom_apply_plan(StepGeom* sg, const int plan[6]) {
    const int ext[3] = {sg->nx, sg->ny, sg->nz};
    int st = step_plan_validate(ext, sg->prec, plan);
    if (st) return st;
    StepChoice c;
    c.x = PassChoice{plan[0], plan[1], 0};
    c.y = PassChoice{plan[2], plan[3], 0};
    c.z = PassChoice{plan[4], plan[5], 0};
    return apply(sg, c);
}

extern "C" int mcs_plan_options(int L, int prec, int* P, int cap, int* n) {
    MCS_CHECK(P && n && cap >= 0, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const std::vector<int> o = line_plan_reg_options(L, prec);
    *n = (int)o.size();
    for (int i = 0; i < (int)o.size() && i < cap; ++i) P[i] = o[i];
    return MCS_OK;
}

extern "C" int mcs_tile_options(int L, int prec, int P, int xpass, int* G, int cap, int* n) {
    MCS_CHECK(G && n && cap >= 0 && P > 0 && L >= P, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    const std::vector<int> o = xpass ? x_tiles(L, P, prec) : col_tiles(L, P, prec);
    *n = (int)o.size();
    for (int i = 0; i < (int)o.size() && i < cap; ++i) G[i] = o[i];
    return MCS_OK;
}

void step_geom_plan_info(const StepGeom& sg, int info[6]) {
    info[0] = sg.px.P;
    info[1] = sg.form == MCS_FORM_MIXED ? xupd_mixed_pairs_per_wg(sg.px) : xupd_pairs_per_wg(sg.px);
    info[2] = sg.py.P;
    info[3] = col_lines_per_wg(sg.py);
    info[4] = sg.is3d ? sg.pz.P : 0;
    info[5] = sg.is3d ? col_lines_per_wg(sg.pz) : 0;
}
