// Persistent work-table kernels of the batched ensembles: one launch per pass
// and size group, workgroups walking the (replica, work unit) table with the
// pass bodies of the single-grid stepper.  FP selects the runtime line plan
// (RunPlan) or a fixed plan.
#pragma once

#include <type_traits>

#include "stepper.hpp"

template <typename F> struct EnsRep {
    C2<F>* U;
    C2<F>* W;
    XArgs<F> xa;
};

// ------------------------------------------------------------------ kernels
template <typename F, int P, int SPLIT, bool WAVE, typename FP = RunPlan, bool NOISE = true>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS) k_ens_xupd(const EnsRep<F>* __restrict__ reps, int nr, int units,
                                                  const int2* __restrict__ pairs, int npairs, int G, FftPlanDev pl,
                                                  const C2<F>* __restrict__ tw, const C2<F>* __restrict__ wph,
                                                  const int64_t* __restrict__ d_step, int draw_off, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const uint32_t draw = NOISE ? (uint32_t)(*d_step + draw_off) : 0u;
    const long total = (long)nr * units;
    // the twiddle table behind the exchange area (twn entries; 0: read from memory), copied once
    // per workgroup
    const int L = FP::fixed ? FP::L : pl.L;
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * 2 * L);
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        const int r = (int)(it / units), u = (int)(it % units);
        const EnsRep<F>& R = reps[r];
        xupd_body<F, P, SPLIT, WAVE, FP, NOISE>(R.W, R.U, pairs, npairs, G, R.xa, pl, twl, wph, draw, lds, u,
                                                opaque_tid());
        __syncthreads();
    }
}

// Column work units are dealt in the tile order swz of xcd_tile over the grid of nr * nouter rows
// of tiles: the launch grid is a multiple of 8 (or the unit count), so unit `it` runs on XCD it % 8
// as a workgroup index would.
template <typename F, int P, int DIR, typename FP = RunPlan>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_ens_col(const EnsRep<F>* __restrict__ reps, int nr, ColGeom g, int G,
                                                  FftPlanDev pl, const C2<F>* __restrict__ tw, int swz, int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int tilesx = (g.nxc + G - 1) / G;
    const long units = (long)tilesx * g.nouter;
    const long total = (long)nr * units;
    const int L = FP::fixed ? FP::L : pl.L;
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * L);
    for (long it0 = blockIdx.x; it0 < total; it0 += gridDim.x) {
        const long it = xcd_tile((unsigned)it0, (unsigned)total, swz, (unsigned)tilesx, (unsigned)(nr * g.nouter));
        const int r = (int)(it / units);
        const long u = it % units;
        C2<F>* W = reps[r].W;
        col_body<F, P, DIR, FP>(W, W, g, G, pl, twl, lds, (int)(u % tilesx), (int)(u / tilesx), opaque_tid());
        __syncthreads();
    }
}

template <typename F, int P, int SPLIT, typename FP = RunPlan>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_ens_col_nl(const EnsRep<F>* __restrict__ reps, int nr, ColGeom g, int G,
                                                     FftPlanDev pl, const C2<F>* __restrict__ tw, F scale, int swz,
                                                     int twn) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int tilesx = (g.nxc + G - 1) / G;
    const long units = (long)tilesx * g.nouter;
    const long total = (long)nr * units;
    const int L = FP::fixed ? FP::L : pl.L;
    const C2<F>* twl = stage_twiddles<F>(tw, twn, lds + G * L);
    for (long it0 = blockIdx.x; it0 < total; it0 += gridDim.x) {
        const long it = xcd_tile((unsigned)it0, (unsigned)total, swz, (unsigned)tilesx, (unsigned)(nr * g.nouter));
        const int r = (int)(it / units);
        const long u = it % units;
        col_nl_body<F, P, SPLIT, FP>(reps[r].W, reps[r].W, g, G, pl, twl, scale, lds, (int)(u % tilesx), (int)(u / tilesx),
                                     opaque_tid());
        __syncthreads();
    }
}


// ------------------------------------------------------------------ plan-class kernels
// One launch covers every size group of one pass whose register plan falls in the same tier of
// points per thread (below): the groups sit one after another in the launch's unit space
// ([first, total) each), every group with its own runtime line plan, twiddles, geometry and
// replicas, and G lines (row pairs) per workgroup chosen so that G * (L / P) is the launch's one
// workgroup size (at most 256 threads).
template <typename F> struct EnsClassGroup {
    FftPlanDev pl;
    const C2<F>* tw;
    const EnsRep<F>* reps;
    const int2* pairs;  // x pass
    const C2<F>* wph;   // x pass
    ColGeom g;          // column passes
    F scale;            // nonlinear pass
    int npairs;
    int P;              // points per thread of this group's plan
    int G;
    int units;          // work units per replica
    long first, total;  // this group's units in the launch
};

template <typename F> MCS_DEV int cls_p(const EnsClassGroup<F>& c) { return c.P; }

// the group of unit `it`; a workgroup visits increasing units, so the search only moves forward
template <typename F> MCS_DEV int cls_group(const EnsClassGroup<F>* cg, int gi, long it) {
    while (it >= cg[gi].total) ++gi;
    return gi;
}

// The points per thread of a class come in two register tiers, so that one launch can hold groups
// of different plans without giving the small plans the register budget of the large ones: tier 0
// takes P in {4, 6, 8, 10, 12}, tier 1 P in {16, 20, 30, 32}.  The body for the group's P is
// selected at run time (every thread of a workgroup takes the same branch).
inline int cls_tier(int P) { return P <= 12 ? 0 : 1; }
template <int TIER, typename Fn> MCS_DEV void with_p(int P, Fn&& f) {
    if constexpr (TIER == 0) {
        switch (P) {
            case 4: f(std::integral_constant<int, 4>{}); break;
            case 6: f(std::integral_constant<int, 6>{}); break;
            case 8: f(std::integral_constant<int, 8>{}); break;
            case 10: f(std::integral_constant<int, 10>{}); break;
            case 12: f(std::integral_constant<int, 12>{}); break;
            default: break;
        }
    } else {
        switch (P) {
            case 16: f(std::integral_constant<int, 16>{}); break;
            case 20: f(std::integral_constant<int, 20>{}); break;
            case 30: f(std::integral_constant<int, 30>{}); break;
            case 32: f(std::integral_constant<int, 32>{}); break;
            default: break;
        }
    }
}

// the x pass's update does not depend on the split (the operator sits in the replica's arguments)
template <typename F, int TIER, bool NOISE>
__global__ void __launch_bounds__(MCS_XPASS_MAX_THREADS) k_cls_xupd(const EnsClassGroup<F>* __restrict__ cg, long total,
                                                                    const int64_t* __restrict__ d_step, int draw_off) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const uint32_t draw = NOISE ? (uint32_t)(*d_step + draw_off) : 0u;
    int gi = 0;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        gi = cls_group(cg, gi, it);
        const EnsClassGroup<F>& c = cg[gi];
        const long rel = it - c.first;
        const EnsRep<F>& R = c.reps[rel / c.units];
        const int u = (int)(rel % c.units);
        // a row pair whose S = L / P threads sit in one wavefront exchanges through LDS with
        // wavefront fences instead of workgroup barriers, as in the per-group kernels
        const bool wave = 64 % (c.pl.L / c.P) == 0;
        with_p<TIER>(cls_p(c), [&](auto PC) {
            if (wave)
                xupd_body<F, decltype(PC)::value, MCS_SPLIT_CUBIC, true, RunPlan, NOISE>(
                    R.W, R.U, c.pairs, c.npairs, c.G, R.xa, c.pl, c.tw, c.wph, draw, lds, u, opaque_tid());
            else
                xupd_body<F, decltype(PC)::value, MCS_SPLIT_CUBIC, false, RunPlan, NOISE>(
                    R.W, R.U, c.pairs, c.npairs, c.G, R.xa, c.pl, c.tw, c.wph, draw, lds, u, opaque_tid());
        });
        __syncthreads();
    }
}

template <typename F, int TIER, int DIR>
__global__ void __launch_bounds__(256) k_cls_col(const EnsClassGroup<F>* __restrict__ cg, long total) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    int gi = 0;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        gi = cls_group(cg, gi, it);
        const EnsClassGroup<F>& c = cg[gi];
        const long rel = it - c.first;
        C2<F>* W = c.reps[rel / c.units].W;
        const int u = (int)(rel % c.units);
        const int tilesx = (c.g.nxc + c.G - 1) / c.G;
        with_p<TIER>(cls_p(c), [&](auto PC) {
            col_body<F, decltype(PC)::value, DIR>(W, W, c.g, c.G, c.pl, c.tw, lds, u % tilesx, u / tilesx, opaque_tid());
        });
        __syncthreads();
    }
}

template <typename F, int TIER, int SPLIT>
__global__ void __launch_bounds__(256) k_cls_col_nl(const EnsClassGroup<F>* __restrict__ cg, long total) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    int gi = 0;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        gi = cls_group(cg, gi, it);
        const EnsClassGroup<F>& c = cg[gi];
        const long rel = it - c.first;
        const int u = (int)(rel % c.units);
        const int tilesx = (c.g.nxc + c.G - 1) / c.G;
        C2<F>* W = c.reps[rel / c.units].W;
        with_p<TIER>(cls_p(c), [&](auto PC) {
            col_nl_body<F, decltype(PC)::value, SPLIT>(W, W, c.g, c.G, c.pl, c.tw, c.scale, lds, u % tilesx, u / tilesx,
                                                       opaque_tid());
        });
        __syncthreads();
    }
}
