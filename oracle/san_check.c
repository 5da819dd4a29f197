/*
 * Exercises every function of the FP64 oracle on small, odd and even shapes, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle san).  TEST INFRASTRUCTURE
 * ONLY: it checks the checker's own memory safety; the values are pinned elsewhere.
 * Exit status 0 and no sanitizer report on stderr is a pass.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double* dalloc(size_t n) {
    double* p = (double*)malloc(n * sizeof(double));
    if (!p) { fprintf(stderr, "out of memory\n"); exit(2); }
    for (size_t i = 0; i < n; ++i) p[i] = 0.01 * (double)((i * 7919u) % 101u) - 0.5;
    return p;
}

static int run_shape(const int n[3]) {
    const int nx = n[0], ny = n[1], nz = n[2], nxc = nx / 2 + 1;
    const size_t N = (size_t)nx * ny * nz, Nc = (size_t)nxc * ny * nz;
    const double h[3] = {0.7, 0.6, 0.5}, lens[3] = {nx * 0.7, ny * 0.6, nz * 0.5};
    double *psi = dalloc(N), *out = dalloc(N), *k1 = dalloc(2 * Nc), *k2 = dalloc(2 * Nc), *k3 = dalloc(2 * Nc);
    double *x1 = dalloc(2 * Nc), *x2 = dalloc(2 * Nc), *kop = dalloc(2 * Nc);
    orc_init_field(psi, n, -0.25, 0.05, 3);
    orc_r2c(psi, k1, n);
    orc_c2r(k1, out, n);
    orc_kspace_noise(x1, n, 5, 1);
    if (nx % 2 == 0) orc_stepper_noise(x2, n, 5, 1);
    memcpy(k2, k1, 2 * Nc * sizeof(double));
    orc_update(k2, k1, x1, n, lens, 0.3, 0.25, 0.1, ORC_SPLIT_CUBIC);
    orc_update(k2, k1, NULL, n, lens, 0.3, 0.25, 0.0, ORC_SPLIT_REFERENCE);
    orc_update_eom_fen(k2, k1, x1, n, lens, 0.1, 0.3, 0.25);
    (void)orc_calc_fen(psi, out, n, h);
    (void)orc_update_el(k2, k1, n, lens, 0.3, 0.25);
    orc_etd_op(kop, n, lens, 0.3, 0.25, 0);
    orc_update_eom_etd(kop, k2, k1, x1, n, lens, 0.1);
    orc_etd_op(kop, n, lens, 0.3, 0.25, 1);
    (void)orc_update_el_etd(kop, k2, k1, n);
    if (nz == 1) orc_update_eom_hpfc(k2, k1, x1, x1, x1, n, lens, 0.1, 0.3, 0.25);
    const double coef[3] = {-1.0, 2.0, -1.0};
    orc_update_eom_poly(k2, k1, x1, n, lens, 0.1, 0.3, 0.25, NULL, 0);
    orc_update_eom_poly(k2, k1, NULL, n, lens, 0.1, 0.3, 0.25, coef, 3);
    (void)orc_free_energy(psi, n, h, 0.25);
    orc_gaussian_init(k3, n, lens, h, 1.5);
    double mass[4];
    if (orc_pfc_run(psi, n, h, 0.25, 0.5, 1e-3, 7, 0, 3, ORC_SPLIT_CUBIC, mass)) return 1;
    int64_t* idx = (int64_t*)malloc(N * sizeof(int64_t));
    const int64_t np = orc_findpeaks(psi, n, -10.0, idx, (int64_t)N);
    const int64_t np2 = orc_findpeaks(psi, n, -10.0, idx, np > 0 ? np - 1 : 0);  /* capacity short by one */
    if (np2 != np) return 1;
    free(idx);
    free(psi); free(out); free(k1); free(k2); free(k3); free(x1); free(x2); free(kop);
    return 0;
}

int main(void) {
    const int shapes[][3] = {{8, 6, 4}, {9, 5, 3}, {16, 12, 1}, {7, 1, 1}, {12, 10, 6}, {2, 2, 2}};
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); ++s)
        if (run_shape(shapes[s])) { fprintf(stderr, "shape %zu failed\n", s); return 1; }
    double c[64];
    for (int i = 0; i < 32; ++i) { c[2 * i] = cos(0.3 * i); c[2 * i + 1] = sin(0.2 * i); }
    orc_fft_c2c(c, 32, -1);
    orc_fft_c2c(c, 15, +1);
    orc_fft_c2c(c, 26, -1);
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, o4[4], st[40];
    orc_philox4x32_10(ctr, key, o4);
    orc_philox_stream(9, 0xfffffffeu, 1, 2, 10, st);
    double re, im;
    for (int l = 0; l <= 8; l += 2)
        for (int m = -l; m <= l; ++m) orc_ylm(l, m, 0.3, 1.1, &re, &im);
    (void)orc_wigner3j(4, 4, 4, 2, -2, 0);
    /* points: a small cubic lattice with jitter, for both bond-order forms and the pair distances */
    const int P = 64, g = 4;
    double* xyz = dalloc(3 * P);
    int* ijk = (int*)malloc(3 * P * sizeof(int));
    for (int i = 0; i < P; ++i) {
        ijk[3 * i] = 2 * (i % g); ijk[3 * i + 1] = 2 * ((i / g) % g); ijk[3 * i + 2] = 2 * (i / (g * g));
        for (int d = 0; d < 3; ++d) xyz[3 * i + d] = ijk[3 * i + d] + 0.01 * ((i * 3 + d) % 7);
    }
    int* nb = (int*)malloc(P * sizeof(int));
    double* o9 = dalloc(9 * P);
    orc_bond_order_cutoff(xyz, P, 8.0, 2.1, nb, o9);
    int* nbr = (int*)malloc(12 * P * sizeof(int));
    double* o4q = dalloc(4 * P);
    orc_bond_order_knn(ijk, P, 8, 0.785, 12, nbr, o4q);
    const int qi[3] = {0, 5, 63};
    orc_bond_order_knn_points(ijk, P, 8, 0.785, 6, qi, 3, nbr, o4q);
    double* pd = dalloc((size_t)P * (P - 1));
    if (orc_pair_distances(xyz, P, 8.0, pd) != (int64_t)P * (P - 1)) return 1;
    double cf[7];
    orc_poly_coefficients(cf);
    free(xyz); free(ijk); free(nb); free(o9); free(nbr); free(o4q); free(pd);
    printf("oracle sanitizer check done\n");
    return 0;
}
