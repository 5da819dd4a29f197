/*
 * FP64 CPU oracle for the phase-field-crystal (PFC) spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libmcstar_hip.so) never links or calls it.
 *
 * Conventions (restated from the reference, file:line relative to the
 * mcstar-libs tree):
 *   - real arrays are unpadded, x fastest: idx = (i2*ny + i1)*nx + i0
 *     (cuda/array_abstraction.hpp:46-55, 163-164)
 *   - complex arrays are the half spectrum (nx/2+1)*ny*nz, x fastest
 *     (cuda/array_abstraction.hpp:49-53)
 *   - R2C and C2R are both unnormalised, c2r(r2c(x)) = N*x
 *     (cuda/cufft_wrapper.hpp:50-58, cufft_unittest/my_unit_test.py:95)
 *   - wavenumber k_i = (i < n/2+1 ? i : i-n) * 2*pi / L_i
 *     (cuda/kernels/common.hpp:27-28)
 */
#ifndef MCSTAR_ORACLE_H
#define MCSTAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (own mixed-radix Cooley-Tukey, any length; FP64) ---- */
/* in-place 1D complex DFT of n points (interleaved re,im), sign -1 forward, +1 inverse, unnormalised */
int orc_fft_c2c(double* data, int n, int sign);
/* n = {nx, ny, nz}; out has (nx/2+1)*ny*nz complex (interleaved) */
int orc_r2c(const double* in, double* out, const int n[3]);
/* in is not modified; numpy irfftn semantics for non-Hermitian input, unnormalised */
int orc_c2r(const double* in, double* out, const int n[3]);

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* out[4*i + j] = philox(ctr = {i_lo + c0, i_hi, c2, c3}, key = seed) for i in [0, n) */
void orc_philox_stream(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int64_t n, uint32_t* out);

/* ---- conserved k-space noise: unit complex Gaussians with Hermitian repair ---- */
/* xi over the half spectrum of a real n-grid; E|xi|^2 = 1, self-conjugate modes real with variance 1 */
void orc_kspace_noise(double* xi, const int n[3], uint64_t seed, uint32_t step);
/* the same statistics in the fused stepper's row layout (even nx): one Philox block per
   (row, packed position) gives the full x modes p and p + nx/2 of that row */
void orc_stepper_noise(double* xi, const int n[3], uint64_t seed, uint32_t step);

/* ---- initial condition psi = psi0 + amp*(2u-1), u 53-bit uniform from Philox(seed, cell) ---- */
void orc_init_field(double* psi, const int n[3], double psi0, double amp, uint64_t seed);

/* ---- PFC semi-implicit update on arrays in the half spectrum ---- */
enum { ORC_SPLIT_CUBIC = 0, ORC_SPLIT_REFERENCE = 1 };
/* kpsi <- (kpsi - dt k^2 knl + noise_amp*|k|*xi) / D(k); xi may be NULL (noise off) */
void orc_update(double* kpsi, const double* knl, const double* xi, const int n[3], const double lens[3],
                double dt, double eps, double noise_amp, int split);
/* full steps, psi real field in/out (normalised). temp = T (0 = noise off); the noise
   follows orc_stepper_noise for even nx and orc_kspace_noise for odd nx.
   mass_hist (optional, nsteps+1 entries) receives the mean density after each step. */
int orc_pfc_run(double* psi, const int n[3], const double h[3], double eps, double dt, double temp,
                uint64_t seed, int step0, int nsteps, int split, double* mass_hist);
/* update with the free-energy operator (eom.hpp:88-101): kpsi updated, knoise <- lk * old kpsi */
void orc_update_eom_fen(double* kpsi, const double* knl, double* knoise, const int n[3], const double lens[3],
                        double namp, double dt, double eps);
/* per-cell free energy 0.5 psi (lpsi - psi + 0.5 psi^3) into lpsi; returns dV * sum (eom.hpp:111-141) */
double orc_calc_fen(const double* psi, double* lpsi, const int n[3], const double h[3]);
/* non-conserved relaxation (el.hpp:3-45); returns max |kpsi_new - kpsi_old| */
double orc_update_el(double* kpsi, const double* knl, const int n[3], const double lens[3], double dt, double eps);
/* exponential time differencing tables (E, S) per mode, conserved (el = 0) or not (el = 1) */
void orc_etd_op(double* kop, const int n[3], const double lens[3], double dt, double eps, int el);
void orc_update_eom_etd(const double* kop, double* kpsi, const double* knl, const double* knoise, const int n[3],
                        const double lens[3], double namp);
double orc_update_el_etd(const double* kop, double* kpsi, const double* knl, const int n[3]);
/* 2D update with the three-array hpfc noise (eom.hpp:16-46) */
void orc_update_eom_hpfc(double* kpsi, const double* knl, const double* x1, const double* x2, const double* x3,
                         const int n[3], const double lens[3], double namp, double dt, double eps);
/* update with the polynomial linear operator (eom.hpp:143-171): denominator 1 + dt k^2 (1 - eps - P(k^2)),
   P(q) = sum_j coef[j] q^j over ncoef coefficients; coef NULL takes the reference's fitted ones */
void orc_update_eom_poly(double* kpsi, const double* knl, const double* knoise, const int n[3],
                         const double lens[3], double namp, double dt, double eps, const double* coef, int ncoef);
/* the reference's fitted coefficients c_0, c_2, ..., c_12 (eom.hpp:143-149), P(q) = sum_j c[j] q^j */
void orc_poly_coefficients(double out[7]);
/* free energy F = dV * sum[ psi/2 * (-eps + (1-k^2)^2) psi + psi^4/4 ] (cuda/kernels/eom.hpp:111-141) */
double orc_free_energy(const double* psi, const int n[3], const double h[3], double eps);

/* ---- local maxima (cuda/kernels/findpeaks.hpp:7-29, findpeaks_2d.hpp:7-27) ---- */
/* strict periodic 26/8-neighbour maxima with v >= ths, ascending linear index; returns count */
int64_t orc_findpeaks(const double* v, const int n[3], double ths, int64_t* idx_out, int64_t cap);

/* ---- Steinhardt bond order ---- */
/* cutoff mode, restating bond-order-analysis/bond-order.cpp:53-111,177-287.
   xyz[3P]; per point: nb[i], out[9*i + {0..8}] = q4 q6 q8 w4 q4b q6b q8b w4b xi (NaN if nb == 0) */
void orc_bond_order_cutoff(const double* xyz, int P, double box, double rlim, int* nb, double* out);
/* k nearest neighbours on a periodic integer grid (cube of side n grid cells, spacing h):
   nbr[k*i + j] = index of the j-th nearest neighbour (integer squared distance, ties by index);
   out[4*i + {0..3}] = q4 q6 q8 w4 */
void orc_bond_order_knn(const int* ijk, int P, int n, double h, int k, int* nbr, double* out);
/* the same for the points qi[0..nq) only (all P points are candidates): row q of
   nbr and out belongs to point qi[q]; qi NULL means every point (nq = P) */
void orc_bond_order_knn_points(const int* ijk, int P, int n, double h, int k, const int* qi, int nq, int* nbr,
                               double* out);
/* normalised spherical harmonic Y_lm(cos(theta), phi), Condon-Shortley phase */
void orc_ylm(int l, int m, double ct, double phi, double* re, double* im);
/* Wigner 3j symbol (j1 j2 j3; m1 m2 m3) for integer arguments */
double orc_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);

/* ---- post-processing ---- */
/* Gaussian filter table (cuda/kernels/gaussian.hpp:23-33): kg[mode] = (e, 1 - e) over the half spectrum */
void orc_gaussian_init(double* kg, const int n[3], const double lens[3], const double h[3], double divs);
/* minimum-image distances of all ordered pairs (particlepairs/main.cpp:8-53), i-major; returns P (P - 1) */
int64_t orc_pair_distances(const double* xyz, int P, double box, double* out);

#ifdef __cplusplus
}
#endif
#endif
