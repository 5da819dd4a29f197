// Drop-in check of the kept C++ host API: the flows of the reference's own
// drivers and of a PFC simulator loop, written against the original header
// paths and consumer typedefs (cufft_unittest/main.cu:1-16), built with
// -DFLOAT=float or -DFLOAT=double and linked to libmcstar_hip.so.
//
//   fft   n0 n1 n2 in.npy out.npy cycle      (cufft_unittest/main.cu:24-74)
//   noise n0 n1 n2 seed out.npy              (curand_kspace_unittest/main.cu:13-34)
//   pfc   n0 n1 n2 in.npy out.npy steps eps dt namp seed h
//         c2r -> psi3_split(norm) -> r2c -> fill_kspace -> update_eom, per step
//   peaks n0 n1 n2 in.npy peaks.txt bo.txt h ths rlim   (findpeaks.hpp:31-63)
//   fen   n0 n1 n2 in.npy eps h        update_eom_with_fen -> c2r -> norm -> calc_fen
//   relax n0 n1 n2 in.npy out.npy steps eps dt   c2r -> psi3_split(norm) -> r2c -> update_el
//   lowpass n0 n1 1 in.npy out.gz amp.npy h divs   gaussian_init -> r2c -> calc_amplitude_lowpass
#include <cuda/array_abstraction.hpp>
#include <cuda/cufft_wrapper.hpp>
#include <cuda/math.hpp>
#include <cuda/noise.hpp>

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

#include <cuda/kernels/el.hpp>
#include <cuda/kernels/eom.hpp>
#include <cuda/kernels/findpeaks.hpp>
#include <cuda/kernels/gaussian.hpp>
#include <cuda/kernels/lowpass.hpp>
#include <cuda/kernels/norm.hpp>
#include <cuda/kernels/psi3_split.hpp>

static int3 shape_arg(char** a) { return int3{atoi(a[0]), atoi(a[1]), atoi(a[2])}; }

int main(int argc, char* argv[]) {
    if (argc < 2) { std::cerr << "usage: " << argv[0] << " fft|noise|pfc|fen|relax|peaks|lowpass ..." << std::endl; return 1; }
    CUERR(mcs_init(0));
    const std::string mode = argv[1];
    if (mode == "fft" && argc == 8) {
        const int3 shape = shape_arg(argv + 2);
        CPUArray host(shape);
        GPUArray d1(host), d2(host);
        PlanR2C r2c(host.real_vext());
        PlanC2R c2r(host.real_vext());
        host.load(argv[5]);
        host >> d1;
        r2c.execute(d1, d2);
        if (atoi(argv[7]) == 0) {
            d2 >> host;
            host.save<Array::AsCmpl>(argv[6]);
        } else {
            c2r.execute(d2, d1);
            d1 >> host;
            host.save(argv[6]);
        }
        return 0;
    }
    if (mode == "noise" && argc == 7) {
        const int3 shape = shape_arg(argv + 2);
        CPUArray host(shape);
        GPUArray dev(host);
        NoiseHostApi<Float> noise(strtoull(argv[5], nullptr, 10));
        noise.fill_kspace(dev);
        dev >> host;
        host.save<Array::AsCmpl>(argv[6]);
        return 0;
    }
    if (mode == "pfc" && argc == 13) {
        const int3 shape = shape_arg(argv + 2);
        const int steps = atoi(argv[7]);
        const Float eps = (Float)atof(argv[8]), dt = (Float)atof(argv[9]), namp = (Float)atof(argv[10]);
        const unsigned long long seed = strtoull(argv[11], nullptr, 10);
        const double h = atof(argv[12]);
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray psi(host), kpsi(shape), kscratch(shape), knl(shape), knoise(shape);
        PlanR2C r2c(shape);
        PlanC2R c2r(shape);
        NoiseHostApi<Float> noise(seed);
        const Float3 lens{(Float)(shape.x * h), (Float)(shape.y * h), (Float)(shape.z * h)};
        r2c.execute(psi, kpsi);
        for (int s = 0; s < steps; ++s) {
            kpsi >> kscratch;
            c2r.execute(kscratch, psi);
            call_kernel_psi3_split(psi, NULL, Float(0), true);
            r2c.execute(psi, knl);
            if (namp > 0) {
                noise.fill_kspace(knoise);
                call_kernel_update_eom(kpsi, knl, knoise, namp, dt, eps, lens);
            } else {
                call_kernel_update_eom(kpsi, knl, namp, dt, eps, lens);
            }
        }
        c2r.execute(kpsi, psi);
        call_kernel_norm(psi);
        psi >> host;
        host.save(argv[6]);
        return 0;
    }
    if (mode == "fen" && argc == 8) {
        // free energy the reference way: update_eom_with_fen leaves lk * kpsi in knoise,
        // C2R and norm give L psi, calc_fen sums f (cuda/kernels/eom.hpp:88-141)
        const int3 shape = shape_arg(argv + 2);
        const Float eps = (Float)atof(argv[6]);
        const double h = atof(argv[7]);
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray psi(host), kpsi(shape), knl(shape), knoise(shape), lpsi(shape);
        PlanR2C r2c(shape);
        PlanC2R c2r(shape);
        const Float3 lens{(Float)(shape.x * h), (Float)(shape.y * h), (Float)(shape.z * h)};
        r2c.execute(psi, kpsi);
        CUERR(mcs_memset(knl.ptr_void(), 0, knl.bytes()));
        CUERR(mcs_memset(knoise.ptr_void(), 0, knoise.bytes()));
        call_kernel_update_eom_with_fen(kpsi, knl, knoise, Float(0), Float(0), eps, lens);
        c2r.execute(knoise, lpsi);
        call_kernel_norm(lpsi);
        const Float3 hh{(Float)h, (Float)h, (Float)h};
        printf("%.17g\n", (double)call_kernel_calc_fen(psi, lpsi, hh));
        return 0;
    }
    if (mode == "relax" && argc == 10) {
        // non-conserved relaxation to a stationary state (cuda/kernels/el.hpp:27-45)
        const int3 shape = shape_arg(argv + 2);
        const int steps = atoi(argv[7]);
        const Float eps = (Float)atof(argv[8]), dt = (Float)atof(argv[9]);
        const double h = M_PI / 4;
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray psi(host), kpsi(shape), kscratch(shape), knl(shape);
        PlanR2C r2c(shape);
        PlanC2R c2r(shape);
        const Float3 lens{(Float)(shape.x * h), (Float)(shape.y * h), (Float)(shape.z * h)};
        r2c.execute(psi, kpsi);
        for (int s = 0; s < steps; ++s) {
            c2r.execute(kpsi, psi);
            call_kernel_psi3_split(psi, NULL, Float(0), true);
            r2c.execute(psi, knl);
            printf("%.17g\n", (double)call_kernel_update_el(kpsi, knl, dt, eps, lens));
        }
        c2r.execute(kpsi, psi);
        call_kernel_norm(psi);
        psi >> host;
        host.save(argv[6]);
        return 0;
    }
    if (mode == "peaks" && argc == 11) {
        const int3 shape = shape_arg(argv + 2);
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray dev(host);
        const Float h = (Float)atof(argv[8]);
        const Float3 hh{h, h, h}, len{shape.x * h, shape.y * h, shape.z * h};
        call_kernel_findpeaks(dev, argv[6], argv[7], hh, len, (Float)atof(argv[9]), (Float)atof(argv[10]));
        return 0;
    }
    if (mode == "lowpass" && argc == 10) {
        // calc_amplitude_lowpass (cuda/kernels/lowpass.hpp:79-93) of a 2D field (z extent 1)
        const int3 shape = shape_arg(argv + 2);
        const Float h = (Float)atof(argv[8]), divs = (Float)atof(argv[9]);
        CPUArray host(shape);
        host.load(argv[5]);
        GPUArray psi(host), spec(shape), aux(shape), kg(shape);
        PlanR2C r2c(shape);
        PlanC2R c2r(shape);
        const Float3 lens{shape.x * h, shape.y * h, shape.z * h}, hh{h, h, h};
        call_kernel_gaussian_init(kg, lens, hh, divs);
        r2c.execute(psi, spec);
        calc_amplitude_lowpass(r2c, c2r, spec, aux, kg, host, lens, argv[6]);
        host.save(argv[7]);
        return 0;
    }
    std::cerr << "bad arguments for mode " << mode << std::endl;
    return 1;
}
