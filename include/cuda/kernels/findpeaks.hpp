// call_kernel_findpeaks (cuda/kernels/findpeaks.hpp:31-63): strict periodic
// 26-neighbour maxima >= ths, written as "%f %f %f %f" (i*h, field) lines,
// then bond_order_analysis on them.  The list is ordered by cell index.
#pragma once

#include <string>
#include <vector>

#include "../../bond-order-analysis/bond-order.hpp"
#include "common.hpp"

typedef Maximum<Float> Max;

inline std::vector<Max> mcs_collect_peaks(GPUArray& input, Float ths) {
    const int3 e = input.real_vext();
    const int64_t cap = std::max<int64_t>(1, (int64_t)input.real_elems() / (e.z > 1 ? 8 : 4) + 8);
    std::vector<int64_t> idx(cap);
    std::vector<double> val(cap);
    int64_t n = 0;
    CUERR(mcs_findpeaks(input.ptr_void(), input.ext_ptr(), (double)ths, idx.data(), val.data(), cap, &n,
                        Array::prec_of<Float>()));
    std::vector<Max> out((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = idx[i];
        out[i].i0 = (Float)(c % e.x);
        out[i].i1 = (Float)((c / e.x) % e.y);
        out[i].i2 = (Float)(c / ((int64_t)e.x * e.y));
        out[i].field = (float)val[i];
    }
    return out;
}

inline void mcs_write_peaks(const std::string& fn, const std::vector<Max>& m, Float3 hh) {
    FILE* f = fopen(fn.c_str(), "w");
    if (!f) { std::cout << "output file <" << fn << "> cannot be opened for writing" << std::endl; return; }
    for (const Max& p : m) {
        const float x0 = p.i0 * hh.x, x1 = p.i1 * hh.y, x2 = p.i2 * hh.z;
        fprintf(f, "%f %f %f %f\n", x0, x1, x2, p.field);
    }
    fclose(f);
}

inline void call_kernel_findpeaks(GPUArray& input, std::string peaks_outfn, std::string oorder_outfn, Float3 hh,
                                  Float3 len, Float ths, Float rlim) {
    if (len.x != len.y || len.x != len.z) {
        std::cerr << "orientational mapping needs a cube" << std::endl;
        return;
    }
    std::vector<Max> m = mcs_collect_peaks(input, ths);
    std::cerr << "found " << m.size() << " maxima" << std::endl;
    mcs_write_peaks(peaks_outfn, m, hh);
    FILE* f = fopen(oorder_outfn.c_str(), "w");
    if (f) {
        bond_order_analysis(m.data(), (int)m.size(), (double)hh.x, (double)hh.y, (double)hh.z, (double)ths,
                            (double)len.x, (double)rlim, f);
        fclose(f);
    } else {
        std::cout << "output file <" << oorder_outfn << "> cannot be opened for writing" << std::endl;
    }
}
