// Persistent work-table kernels of the batched ensembles: one launch per pass
// and size group, workgroups walking the (replica, work unit) table with the
// pass bodies of the single-grid stepper.  FP selects the runtime line plan
// (RunPlan) or a fixed plan.
#pragma once

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
                                                  const int64_t* __restrict__ d_step, int draw_off) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const uint32_t draw = NOISE ? (uint32_t)(*d_step + draw_off) : 0u;
    const long total = (long)nr * units;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        const int r = (int)(it / units), u = (int)(it % units);
        const EnsRep<F>& R = reps[r];
        xupd_body<F, P, SPLIT, WAVE, FP, NOISE>(R.W, R.U, pairs, npairs, G, R.xa, pl, tw, wph, draw, lds, u,
                                                opaque_tid());
        __syncthreads();
    }
}

template <typename F, int P, int DIR, typename FP = RunPlan>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_ens_col(const EnsRep<F>* __restrict__ reps, int nr, ColGeom g, int G,
                                                  FftPlanDev pl, const C2<F>* __restrict__ tw) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int tilesx = (g.nxc + G - 1) / G;
    const long units = (long)tilesx * g.nouter;
    const long total = (long)nr * units;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        const int r = (int)(it / units);
        const long u = it % units;
        C2<F>* W = reps[r].W;
        col_body<F, P, DIR, FP>(W, W, g, G, pl, tw, lds, (int)(u % tilesx), (int)(u / tilesx), opaque_tid());
        __syncthreads();
    }
}

template <typename F, int P, int SPLIT, typename FP = RunPlan>
__global__ void __launch_bounds__(MCS_REG_MAX_THREADS(P, F)) k_ens_col_nl(const EnsRep<F>* __restrict__ reps, int nr, ColGeom g, int G,
                                                     FftPlanDev pl, const C2<F>* __restrict__ tw, F scale) {
    extern __shared__ unsigned char smem_raw[];
    C2<F>* lds = reinterpret_cast<C2<F>*>(smem_raw);
    const int tilesx = (g.nxc + G - 1) / G;
    const long units = (long)tilesx * g.nouter;
    const long total = (long)nr * units;
    for (long it = blockIdx.x; it < total; it += gridDim.x) {
        const int r = (int)(it / units);
        const long u = it % units;
        col_nl_body<F, P, SPLIT, FP>(reps[r].W, g, G, pl, tw, scale, lds, (int)(u % tilesx), (int)(u / tilesx),
                                     opaque_tid());
        __syncthreads();
    }
}

