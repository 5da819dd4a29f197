// Kept host API: FFT plans (cuda/cufft_wrapper.hpp:40-86) on the library's own
// Stockham FFT.  FFT::CufftPlan<F, FFT::R2C|FFT::C2R>(int3 real_extent) plans
// an unnormalised rank-1/2/3 transform from the extents; execute(in, out)
// asserts the extents and returns with the transform complete (86: sync).
// The cuFFT handle types of the original (FFT::exec, CufftType) are not part
// of this API.
#pragma once

#include "array_abstraction.hpp"

namespace FFT {
struct R2C { static constexpr int dir = MCS_R2C; };
struct C2R { static constexpr int dir = MCS_C2R; };

template <typename F, typename FT> class CufftPlan {
    int3 m_shape;
    mcs_fft_plan* m_plan = nullptr;

  public:
    explicit CufftPlan(int3 shape) : m_shape(shape) {
        const int ext[3] = {shape.x, shape.y, shape.z};
        CUERR(mcs_fft_plan_create(&m_plan, ext, Array::prec_of<F>(), FT::dir));
    }
    ~CufftPlan() { CUERR(mcs_fft_plan_destroy(m_plan)); }
    CufftPlan(const CufftPlan&) = delete;
    CufftPlan& operator=(const CufftPlan&) = delete;
    void execute(Array::GPUArray<F>& in, Array::GPUArray<F>& out) {
        assert(in.real_vext() == m_shape && out.real_vext() == m_shape &&
               "array dimensions in plan and execute phase do not match up");
        CUERR(mcs_fft_exec(m_plan, in.ptr_void(), out.ptr_void()));
    }
};

template <typename F, typename FT> using Plan = CufftPlan<F, FT>;
}  // namespace FFT
