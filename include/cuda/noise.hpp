// Kept host API: NoiseHostApi<F> (cuda/noise.hpp:93-150) on counter-based
// Philox4x32-10.  fill() writes N(0, sigma^2) reals; fill_kspace() writes unit
// complex Gaussians over the half spectrum, Hermitian on the x = 0 and nx/2
// planes, self-conjugate modes real with unit variance.  Each call draws a new
// counter block (key = seed), so runs are reproducible and resumable.
#pragma once

#include "array_abstraction.hpp"

template <typename F> class NoiseHostApi {
    mcs_rng* m_rng = nullptr;

  public:
    explicit NoiseHostApi(unsigned long long seed = 1234ULL) { CUERR(mcs_rng_create(&m_rng, seed)); }
    ~NoiseHostApi() { CUERR(mcs_rng_destroy(m_rng)); }
    NoiseHostApi(const NoiseHostApi&) = delete;
    NoiseHostApi& operator=(const NoiseHostApi&) = delete;
    void fill(Array::GPUArray<F>& arr, F sigma) {
        CUERR(mcs_rng_fill_normal(m_rng, arr.ptr_void(), (size_t)arr.real_elems(), (double)sigma,
                                  Array::prec_of<F>()));
    }
    void fill_kspace(Array::GPUArray<F>& arr) {
        CUERR(mcs_rng_fill_kspace(m_rng, arr.ptr_void(), arr.ext_ptr(), Array::prec_of<F>()));
    }
    // draw counter (the step index of a k-space noise sequence)
    unsigned draw() { uint32_t d = 0; CUERR(mcs_rng_get_draw(m_rng, &d)); return d; }
    void set_draw(unsigned d) { CUERR(mcs_rng_set_draw(m_rng, d)); }
};
