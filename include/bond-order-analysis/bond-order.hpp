// bond_order_analysis<T> (bond-order-analysis/bond-order.hpp:4, bond-order.cpp:289-349)
// on the GPU: keeps peaks with field > ths at positions i*h, neighbours within
// rlim under the minimum image of a cubic box, and writes the reference's table
//   # x y z #nb xi q4 q6 q8 w4 q4b q6b q8b w4b
// for every point with at least one neighbour.
#pragma once

#include <cstdio>
#include <iostream>
#include <vector>

#include "../common/cu_error.hpp"
#include "../common/maximum_struct.hpp"

template <typename T>
void bond_order_analysis(Maximum<T>* input, int input_len, double h0, double h1, double h2, double ths,
                         double boxsize, double rlim, FILE* fp) {
    std::vector<double> xyz;
    xyz.reserve(3 * (size_t)input_len);
    for (int i = 0; i < input_len; ++i) {
        const Maximum<T>& m = input[i];
        if (m.field > ths) {
            xyz.push_back(m.i0 * h0);
            xyz.push_back(m.i1 * h1);
            xyz.push_back(m.i2 * h2);
        }
    }
    const int P = (int)(xyz.size() / 3);
    std::cerr << P << " atoms above threshold" << std::endl;
    std::vector<int> nb(P > 0 ? P : 1);
    std::vector<double> out(9 * (size_t)(P > 0 ? P : 1));
    CUERR(mcs_bond_order_cutoff(xyz.data(), P, boxsize, rlim, nb.data(), out.data()));
    std::cerr << "all nbors computed" << std::endl;
    fprintf(fp, "# x y z #nb xi q4 q6 q8 w4 q4b q6b q8b w4b\n");
    for (int i = 0; i < P; ++i) {
        if (nb[i] == 0) continue;
        const double* o = &out[9 * (size_t)i];
        fprintf(fp, "%.6f\t%.6f\t%.6f\t%d\t%d\t%.6f\t%.6f\t%.6f\t%.6f\t%.6f\t%.6f\t%.6f\t%.6f\n", xyz[3 * i],
                xyz[3 * i + 1], xyz[3 * i + 2], nb[i], (int)o[8], o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7]);
    }
    fflush(fp);
}
