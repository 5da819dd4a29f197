// Host-only checks of the kept API helpers (no device calls): FFT-size policy,
// layout helpers and npy round trips.
#include <cuda/math.hpp>
#include <common/npyutil.hpp>
#include <common/technicalities.hpp>

#include <cstdio>
#include <vector>

// the consumer typedefs the kept potentials header expects (cufft_unittest/main.cu:6-16 style),
// with a host-only array standing in for CPUArray (init_pot_line reads real_vext and ptr_real)
typedef double Float;
typedef double3 Float3;
struct HostField {
    int3 ext;
    std::vector<Float> v;
    int3 real_vext() const { return ext; }
    Float* ptr_real() { return v.data(); }
};
typedef HostField CPUArray;
#include <cuda/potentials.hpp>

int main(int argc, char** argv) {
    const int ns[] = {1, 5, 7, 11, 13, 14, 26, 49, 97, 120, 625, 1000, 1023};
    for (int n : ns) printf("%d %d %d\n", n, (int)factors235(n), roundup235(n));
    const int3 a = shape_tr(int3{4, 5, 6}), b = shape_tr(int3{4, 5, 1}), c = shape_tr(int3{4, 1, 1});
    printf("shape_tr %d %d %d | %d %d %d | %d %d %d\n", a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
    printf("cond %.12f\n", crystal_condition(int3{1, 0, 0}, int3{0, 1, 1}));
    auto p = calc_closest_lattice_indices(int2{0, 0}, double2{1, 0}, double2{0.5, 0.8660254037844386}, double2{0.9, 0.1});
    printf("closest %d %d %.12f\n", p.first.x, p.first.y, p.second);
    {
        // 16 x 12 grid, h = 0.5, box (8, 6), power 2, sigma 2, amp 1.5
        HostField f{int3{16, 12, 1}, std::vector<Float>(16 * 12)};
        init_pot_line(f, Float3{8.0, 6.0, 1.0}, Float3{0.5, 0.5, 1.0}, 2, 2.0, 1.5);
        printf("pot");
        for (double v : f.v) printf(" %.17g", v);
        printf("\n");
    }
    if (argc > 1) {
        const int shape[3] = {2, 3, 4};
        double d[24];
        for (int i = 0; i < 24; ++i) d[i] = 0.5 * i - 3;
        aoba::SaveArrayAsNumpy(argv[1], 3, shape, d);
        std::vector<int> s;
        std::vector<double> back;
        aoba::LoadArrayFromNumpy(argv[1], s, back);
        printf("npy %zu %d %d %d %g %g\n", back.size(), s[0], s[1], s[2], back[0], back[23]);
        try {
            std::vector<float> wrong;
            aoba::LoadArrayFromNumpy(argv[1], s, wrong);
            printf("npy dtype check missing\n");
        } catch (const std::runtime_error&) {
            printf("npy dtype mismatch rejected\n");
        }
    }
    return 0;
}
