// Drop-in check of the kept headers a simulator includes instead of eom.hpp / el.hpp
// (cuda/kernels/eom-exp.hpp, el-exp.hpp: the same call names with exponential time
// differencing), of call_kernel_findpeaks_2d (findpeaks_2d.hpp:29-56) and of the external
// potential (potentials.hpp:5-20 into psi3_split.hpp:1-6), built with -DFLOAT=float or
// -DFLOAT=double and linked to libmcstar_hip.so.
//
//   etd     n0 n1 n2 in.npy out.npy steps eps dt     init_eom_op, then per step
//           c2r -> psi3_split(norm) -> r2c -> update_eom (ETD, conserved)
//   elexp   n0 n1 n2 in.npy out.npy steps eps dt     init_el_op, then per step
//           c2r -> psi3_split(norm) -> r2c -> update_el (ETD, non-conserved); prints max |change|
//   peaks2d n0 n1 in.npy peaks.txt h ths              call_kernel_findpeaks_2d
//   pot     n0 n1 in.npy out.npy h power sigma amp   init_pot_line -> psi3_split(arr, &pot)
#include <cuda/array_abstraction.hpp>
#include <cuda/cufft_wrapper.hpp>
#include <cuda/math.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

typedef FLOAT Float;
struct Float3 { Float x, y, z; };
typedef Array::CPUArray<Float> CPUArray;
typedef Array::GPUArray<Float> GPUArray;
typedef FFT::CufftPlan<Float, FFT::R2C> PlanR2C;
typedef FFT::CufftPlan<Float, FFT::C2R> PlanC2R;

#include <cuda/kernels/el-exp.hpp>
#include <cuda/kernels/eom-exp.hpp>
#include <cuda/kernels/findpeaks_2d.hpp>
#include <cuda/kernels/norm.hpp>
#include <cuda/kernels/psi3_split.hpp>
#include <cuda/potentials.hpp>

static int3 shape_arg(char** a) { return int3{atoi(a[0]), atoi(a[1]), atoi(a[2])}; }

int main(int argc, char* argv[]) {
    if (argc < 2) { std::cerr << "usage: " << argv[0] << " etd|elexp|peaks2d|pot ..." << std::endl; return 1; }
    CUERR(mcs_init(0));
    const std::string mode = argv[1];
    if ((mode == "etd" || mode == "elexp") && argc == 10) {
        const int3 shape = shape_arg(argv + 2);
        const int steps = atoi(argv[7]);
        const Float eps = (Float)atof(argv[8]), dt = (Float)atof(argv[9]);
        const double h = M_PI / 4;
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray psi(host), kpsi(shape), knl(shape), kop(shape);
        PlanR2C r2c(shape);
        PlanC2R c2r(shape);
        const Float3 lens{(Float)(shape.x * h), (Float)(shape.y > 1 ? shape.y * h : 1.0),
                          (Float)(shape.z > 1 ? shape.z * h : 1.0)};
        if (mode == "etd") call_kernel_init_eom_op(kop, dt, eps, lens);
        else call_kernel_init_el_op(kop, dt, eps, lens);
        r2c.execute(psi, kpsi);
        for (int s = 0; s < steps; ++s) {
            c2r.execute(kpsi, psi);
            call_kernel_psi3_split(psi, NULL, Float(0), true);
            r2c.execute(psi, knl);
            if (mode == "etd") call_kernel_update_eom(kop, kpsi, knl, Float(0), dt, eps, lens);
            else printf("%.17g\n", (double)call_kernel_update_el(kop, kpsi, knl, dt, eps, lens));
        }
        c2r.execute(kpsi, psi);
        call_kernel_norm(psi);
        psi >> host;
        host.save(argv[6]);
        return 0;
    }
    if (mode == "peaks2d" && argc == 8) {
        const int3 shape{atoi(argv[2]), atoi(argv[3]), 1};
        CPUArray host(shape);
        host.load(argv[4]);
        GPUArray dev(host);
        const Float h = (Float)atof(argv[6]);
        const Float3 hh{h, h, h}, len{shape.x * h, shape.y * h, h};
        call_kernel_findpeaks_2d(dev, argv[5], hh, len, (Float)atof(argv[7]));
        return 0;
    }
    if (mode == "pot" && argc == 10) {
        const int3 shape{atoi(argv[2]), atoi(argv[3]), 1};
        const double h = atof(argv[6]);
        CPUArray host(shape), pot_host(shape);
        host.load(argv[4]);
        const Float3 l{(Float)(shape.x * h), (Float)(shape.y * h), (Float)1}, hh{(Float)h, (Float)h, (Float)1};
        init_pot_line(pot_host, l, hh, atoi(argv[7]), atof(argv[8]), atof(argv[9]));
        GPUArray psi(host), pot(pot_host);
        call_kernel_psi3_split(psi, &pot, Float(0.1), false);
        psi >> host;
        host.save(argv[5]);
        return 0;
    }
    std::cerr << "bad arguments for mode " << mode << std::endl;
    return 1;
}
