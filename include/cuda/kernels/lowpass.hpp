// Low-pass amplitude of a 2D field and its gz text dump, kept API
// (cuda/kernels/lowpass.hpp:3-93).
//
// calc_amplitude_lowpass(r2c, c2r, spectrum, aux, kgauss, host, ll, fn), with
// kgauss from call_kernel_gaussian_init and spectrum = R2C(psi):
//   psi_hi = C2R(spectrum * Im kgauss) / N     high-passed field
//   A      = R2C(psi_hi^2)                     spectrum of its square
//   fn     <- text of the modes of A * Re kgauss with |.|^2 > 1e-6 (gzip)
//   host   <- C2R(A * Re kgauss) / N           the low-passed square: the local amplitude
// spectrum and aux are overwritten.  The file holds "LOWPASS\n<header bytes>\n", a header
// naming n0 n1 l0 l1 in Python's pickle protocol-0 text, then one line "i0 i1 re im"
// per kept mode, i0 over the half axis (outer loop), i1 over the full axis.
#pragma once

#include <cassert>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include <common/zstream.hpp>

#include "common.hpp"
#include "norm.hpp"

inline void _call_kernel_multassign_cmpl_with_real(GPUArray& karr1, GPUArray& karr2) {
    CUERR(mcs_kmult(karr1.ptr_void(), karr2.ptr_void(), karr1.ext_ptr(), 0, Array::prec_of<Float>()));
}

inline void _call_kernel_multassign_cmpl_with_imag(GPUArray& karr1, GPUArray& karr2) {
    CUERR(mcs_kmult(karr1.ptr_void(), karr2.ptr_void(), karr1.ext_ptr(), 1, Array::prec_of<Float>()));
}

inline void _call_kernel_square(GPUArray& karr1) {
    CUERR(mcs_square(karr1.ptr_void(), karr1.ext_ptr(), Array::prec_of<Float>()));
}

inline void write_lowpass_2d(GPUArray& spectrum, GPUArray& kgaussian, CPUArray& host, Float3 ll, std::string fn) {
    const int3 shape = spectrum.real_vext();
    const int3 cshape = spectrum.cmpl_vext();
    assert(shape == kgaussian.real_vext() && host.real_vext() == shape && "arrays have different dimensions");
    assert(shape.z == 1 && shape.y > 1 && shape.x > 1 && "array is not 2d");
    _call_kernel_multassign_cmpl_with_real(spectrum, kgaussian);
    spectrum >> host;
    char head[256];
    std::snprintf(head, sizeof head, "(S'n0'\nI%d\nS'n1'\nI%d\nS'l0'\nF%f\nS'l1'\nF%f\nd.\n", shape.x, shape.y,
                  (double)ll.x, (double)ll.y);
    std::ostringstream text;
    text << "LOWPASS\n" << std::strlen(head) << "\n" << head;
    const auto* k = host.ptr_cmpl();
    for (int i0 = 0; i0 < cshape.x; ++i0)
        for (int i1 = 0; i1 < cshape.y; ++i1) {
            const auto v = k[(size_t)i1 * cshape.x + i0];
            const Float re2 = v.x * v.x;
            const Float im2 = v.y * v.y;
            const Float p = re2 + im2;
            if (p > 1e-6) text << i0 << " " << i1 << " " << v.x << " " << v.y << "\n";
        }
    ozstream out(fn.c_str());
    out << text.str();
}

inline void calc_amplitude_lowpass(PlanR2C& r2c, PlanC2R& c2r, GPUArray& spectrum, GPUArray& aux,
                                   GPUArray& kgaussian, CPUArray& host, Float3 ll, std::string fn) {
    const int3 shape = spectrum.real_vext();
    assert(shape == aux.real_vext() && host.real_vext() == shape && shape == kgaussian.real_vext() &&
           "arrays have different dimensions");
    assert(shape.z == 1 && shape.y > 1 && shape.x > 1 && "array is not 2d");
    (void)shape;
    _call_kernel_multassign_cmpl_with_imag(spectrum, kgaussian);
    c2r.execute(spectrum, aux);
    call_kernel_norm(aux);
    _call_kernel_square(aux);
    r2c.execute(aux, spectrum);
    write_lowpass_2d(spectrum, kgaussian, host, ll, fn);
    c2r.execute(spectrum, aux);
    call_kernel_norm(aux);
    aux >> host;
}
