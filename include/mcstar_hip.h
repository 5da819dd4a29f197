/*
 * mcstar_hip.h -- C ABI of libmcstar_hip.so, the MI355X (gfx950) PFC time step.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * MCS_OK (0) or an MCS_ERR_* code; mcs_error_string() names the code and
 * mcs_last_error() gives "file:line message" of the most recent failure.
 * The C++ host headers under include/cuda and include/common keep the
 * reference's host API and call only these functions.
 *
 * Array conventions (cuda/array_abstraction.hpp:37-70):
 *   ext[3] = {nx, ny, nz}, real extents, x fastest (int3 real_vext()).
 *   real view:    (i2*ny + i1)*nx + i0, unpadded
 *   complex view: (i2*ny + i1)*(nx/2+1) + i0, interleaved re,im
 *   one buffer of (nx/2+1)*ny*nz*2*sizeof(F) bytes backs both views.
 * Precision: prec = 32 (float) or 64 (double).
 * Synchronisation: every entry point that hands data back to the host, or
 * that the reference made synchronous (cufft_wrapper.hpp:85, every
 * call_kernel_* wrapper), returns with its device work complete.
 */
#ifndef MCSTAR_HIP_H
#define MCSTAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define MCS_OK 0
#define MCS_ERR_INVALID 1     /* bad argument (null pointer, extent < 1, mismatch) */
#define MCS_ERR_HIP 2         /* HIP runtime failure */
#define MCS_ERR_UNSUPPORTED 3 /* size or mode outside what the kernels implement */
#define MCS_ERR_NOMEM 4       /* device or host allocation failed */
#define MCS_ERR_CAPACITY 5    /* output capacity too small; count still reported */

#define MCS_F32 32
#define MCS_F64 64

/* Replaces cudaGetErrorString / cufftGetErrorString (common/cu_error.hpp:5-11,
   cuda/cufft_wrapper.hpp:7-29). */
const char* mcs_error_string(int code);
const char* mcs_last_error(void);

/* Replaces cudaSetDevice(0) (cufft_unittest/main.cu:29). */
int mcs_init(int device);
int mcs_sync(void);

/* ---------------------------------------------------------------- memory
 * GPUArray / CPUArray storage (cuda/array_abstraction.hpp:74-131):
 * cudaMalloc / cudaFree / cudaHostAlloc / cudaFreeHost / cudaMemcpy. */
#define MCS_H2D 1
#define MCS_D2H 2
#define MCS_D2D 3
#define MCS_H2H 4
int mcs_malloc(void** ptr, size_t bytes);
int mcs_free(void* ptr);
int mcs_host_alloc(void** ptr, size_t bytes);
int mcs_host_free(void* ptr);
int mcs_memcpy(void* dst, const void* src, size_t bytes, int kind);
int mcs_memset(void* dst, int value, size_t bytes);

/* ---------------------------------------------------------------- FFT
 * FFT::CufftPlan<F, R2C|C2R> (cuda/cufft_wrapper.hpp:61-86).  Rank 1/2/3
 * from the extents (72-74), unnormalised both ways, out-of-place.  The C2R
 * input is left unmodified (the plan owns a scratch buffer).  Axis lengths
 * may have prime factors 2, 3, 5, 7, 11 and 13 (cuda/math.hpp:125-149 rounds
 * sizes to 2,3,5-smooth ones).  Lengths 2^a, 3*2^a and 5*2^a run in
 * registers up to 8192, 12288 and 10240 points in f32 (8192, 6144 and 5120
 * in f64); other lengths need two copies of the line in 160 KB of LDS (at
 * most 10239 points in f32, 5119 in f64).  Larger lengths return
 * MCS_ERR_UNSUPPORTED.
 * C2R of a non-Hermitian spectrum follows numpy.fft.irfftn: the
 * imaginary parts of the x = 0 and x = nx/2 entries are ignored.
 * in == out returns MCS_ERR_INVALID for R2C, and for C2R of a 1D transform
 * (the reference's transforms are out of place: its real view is unpadded). */
#define MCS_R2C 0
#define MCS_C2R 1
typedef struct mcs_fft_plan mcs_fft_plan;
int mcs_fft_plan_create(mcs_fft_plan** plan, const int ext[3], int prec, int dir);
int mcs_fft_exec(mcs_fft_plan* plan, const void* in, void* out);
int mcs_fft_plan_destroy(mcs_fft_plan* plan);

/* ---------------------------------------------------------------- noise
 * NoiseHostApi<F> (cuda/noise.hpp:93-150), with the cuRAND XORWOW stream
 * replaced by counter-based Philox4x32-10: key = seed, counter =
 * (element index lo, hi, draw number, stream tag).  Each fill advances the
 * draw number by one. */
typedef struct mcs_rng mcs_rng;
int mcs_rng_create(mcs_rng** rng, uint64_t seed);
int mcs_rng_destroy(mcs_rng* rng);
int mcs_rng_set_draw(mcs_rng* rng, uint32_t draw);
int mcs_rng_get_draw(mcs_rng* rng, uint32_t* draw);
/* fill(GPUArray&, F sigma): n real N(0, sigma^2) values (noise.hpp:108-109) */
int mcs_rng_fill_normal(mcs_rng* rng, void* dst, size_t n, double sigma, int prec);
/* fill_kspace(GPUArray&): unit complex Gaussians over the half spectrum with
   Hermitian repair on the x = 0, nx/2 planes (noise.hpp:110-142) */
int mcs_rng_fill_kspace(mcs_rng* rng, void* cmpl, const int ext[3], int prec);
/* raw Philox4x32-10 blocks: out[4*i + j] for counter {c0 + i (64-bit lo,hi), c2, c3} */
int mcs_philox_blocks(uint64_t seed, uint64_t c0, uint32_t c2, uint32_t c3, size_t n, uint32_t* out_dev);

/* ---------------------------------------------------------------- pointwise
 * call_kernel_psi3_split (cuda/kernels/psi3_split.hpp:15-26):
 *   a <- (mul*a)^3 - mul*a - mu [+ pot]; mul = 1/N when norm != 0. */
int mcs_psi3_split(void* real, const void* pot, double mu, int norm, const int ext[3], int prec);
/* call_kernel_norm (cuda/kernels/norm.hpp:5-10): a <- a / N */
int mcs_norm(void* real, const int ext[3], int prec);
/* call_kernel_update_eom (cuda/kernels/eom.hpp:60-86), reference split:
 *   kpsi <- (kpsi - dt k^2 knonlin + dt namp sqrt(N) |k| knoise) / (1 + dt k^2 (1 - eps + (1-k^2)^2));
 *   knoise may be NULL (noise off). */
int mcs_update_eom(void* kpsi, const void* knonlin, const void* knoise, const int ext[3], double namp,
                   double dt, double eps, const double lens[3], int prec);
/* PFC split selector for the fused paths */
#define MCS_SPLIT_CUBIC 0 /* N = psi^3,       D = 1 + dt k^2 (-eps + (1-k^2)^2) */
#define MCS_SPLIT_REFERENCE 1 /* N = psi^3 - psi, D = 1 + dt k^2 (1 - eps + (1-k^2)^2) */
/* the same update with conserved noise drawn in registers:
 *   eta_k = sqrt(2 T dt / dV) sqrt(N) |k| xi_k(seed, draw), xi_k = mcs_rng_fill_kspace's values. */
int mcs_update_pfc(void* kpsi, const void* knl, const int ext[3], const double lens[3], double dt, double eps,
                   double temp, uint64_t seed, uint32_t draw, int split, int prec);
/* psi = psi0 + amp (2u - 1), u a 53-bit uniform from Philox(seed, cell index) (computed in FP64) */
int mcs_init_field(void* real, const int ext[3], double psi0, double amp, uint64_t seed, int prec);
/* FP64 reductions of a real field: mean density and free energy
 *   F = dV sum[psi/2 (-eps + (1+lap)^2) psi + psi^4/4]  (cuda/kernels/eom.hpp:111-141) */
int mcs_field_stats(const void* real, const int ext[3], const double h[3], double eps, int prec,
                    double* mean, double* free_energy);

/* call_kernel_update_eom_with_fen (cuda/kernels/eom.hpp:88-109): the update_eom step in
 * that kernel's operation order, and knoise <- lk * kpsi (the kpsi before the update) with
 * lk = 2 - eps - 2 k^2 + k^4, the linear operator of the free energy.  knoise is read (noise)
 * and overwritten.  namp is scaled by sqrtf(N) as in update_eom. */
int mcs_update_eom_with_fen(void* kpsi, const void* knonlin, void* knoise, const int ext[3], double namp,
                            double dt, double eps, const double lens[3], int prec);
/* call_kernel_calc_fen (eom.hpp:111-141): per cell f = psi/2 (lpsi - psi + psi^3/2), written into
 * lpsi; *fen = h0 h1 h2 * sum f, summed in FP64 in a fixed order (the reference used FP32 atomics). */
int mcs_calc_fen(const void* psi, void* lpsi, const int ext[3], const double h[3], int prec, double* fen);
/* call_kernel_update_el (cuda/kernels/el.hpp:3-45): non-conserved relaxation
 * kpsi <- (kpsi - dt knl) / (1 + dt (2 - eps - 2 k^2 + k^4)); *maxdiff = max over the half
 * spectrum of |kpsi_new - kpsi_old| (the convergence measure). */
int mcs_update_el(void* kpsi, const void* knonlin, const int ext[3], double dt, double eps, const double lens[3],
                  int prec, double* maxdiff);

/* ---------------------------------------------------------------- alternative integrators
 * Exponential time differencing (cuda/kernels/eom-exp.hpp, el-exp.hpp): the init calls fill
 * a half-spectrum table kop of (E, S) pairs; conserved: E = exp(-k^2 L' dt),
 * S = (E - 1) / (k^2 L') (-dt where the operator vanishes); non-conserved: E = exp(-L' dt),
 * S = (E - 1) / L', with L' = 2 - eps - 2 k^2 + k^4.  expf in both precisions, as there.
 *   conserved:     kpsi <- kpsi E - dop.x knl S - namp sqrtf(N) |k| knoise S   (eom-exp.hpp:18-69)
 *                  (knoise may be null; with_fen also writes lk * old kpsi into knoise)
 *   non-conserved: kpsi <- kpsi E + knl S, *maxdiff = max |change|            (el-exp.hpp:17-61) */
int mcs_init_eom_op(void* kop, const int ext[3], double dt, double eps, const double lens[3], int prec);
int mcs_init_el_op(void* kop, const int ext[3], double dt, double eps, const double lens[3], int prec);
int mcs_update_eom_etd(const void* kop, void* kpsi, const void* knonlin, void* knoise, int with_fen,
                       const int ext[3], double namp, double eps, const double lens[3], int prec);
int mcs_update_el_etd(const void* kop, void* kpsi, const void* knonlin, const int ext[3], int prec, double* maxdiff);
/* call_kernel_update_eom_hpfcnoise (eom.hpp:16-58), 2D only: the semi-implicit update with the
 * noise (c1 kx^2 + c2 ky^2) xi1 + (c2 kx^2 + c1 ky^2) xi2 + 2 kx ky xi3 for |k| <= 1/4 */
int mcs_update_eom_hpfc(void* kpsi, const void* knonlin, const void* knoise1, const void* knoise2,
                        const void* knoise3, const int ext[3], double namp, double dt, double eps,
                        const double lens[3], int prec);
/* call_kernel_update_eom_py (eom.hpp:143-180): the operator 1 - eps - P(k^2) with the reference's
 * degree-6 polynomial P; knoise may be null */
int mcs_update_eom_poly(void* kpsi, const void* knonlin, const void* knoise, const int ext[3], double namp,
                        double dt, double eps, const double lens[3], int prec);

/* ---------------------------------------------------------------- stepper
 * The fused PFC step: the state lives in a packed spectrum (x pairs packed
 * into one complex value), each step is four passes in 3D (two in 2D):
 *   x: FFT, semi-implicit update + Philox noise, inverse FFT
 *   y: inverse FFT | z: inverse FFT, nonlinearity, FFT | y: FFT
 * nx even; axis lengths nx/2, ny, nz of the form 2^a, 3*2^a or 5*2^a (>= 4). */
typedef struct {
    int ext[3];     /* nx, ny, nz (nz = 1 for 2D) */
    double h[3];    /* grid spacing */
    double eps, dt;
    double temp;    /* T, 0 = noise off */
    uint64_t seed;  /* noise key */
    int split;      /* MCS_SPLIT_* */
    int prec;       /* MCS_F32 / MCS_F64 */
} mcs_pfc_params;
typedef struct mcs_stepper mcs_stepper;
int mcs_stepper_create(mcs_stepper** st, const mcs_pfc_params* p);
/* load the real field (device, reference layout, precision p->prec) */
int mcs_stepper_set_psi(mcs_stepper* st, const void* psi_dev);
int mcs_stepper_get_psi(mcs_stepper* st, void* psi_dev);
int mcs_stepper_step(mcs_stepper* st, int nsteps);
/* steps taken so far (the Philox draw number of the next step) */
int mcs_stepper_get_step(mcs_stepper* st, int64_t* step);
int mcs_stepper_set_step(mcs_stepper* st, int64_t step);
/* launch nsteps on the library stream without waiting (for timing); mcs_sync() to finish */
int mcs_stepper_step_async(mcs_stepper* st, int nsteps);
/* mean density from the FP64 k=0 bookkeeping */
int mcs_stepper_mass(mcs_stepper* st, double* mean);
int mcs_stepper_destroy(mcs_stepper* st);
/* Instrumentation (no counterpart in the reference): run nsteps with a HIP
 * event between consecutive passes on the library stream and return the mean
 * milliseconds of each pass (x_update, y_inverse, z_nonlinear, y_forward in
 * 3D; x_update, y_nonlinear in 2D) in ms[0..*npass). */
int mcs_stepper_pass_times(mcs_stepper* st, int nsteps, double* ms, int cap, int* npass);

/* Autotuning of the fused step (no reference counterpart).  When on (the
 * default), mcs_stepper_create and mcs_ensemble_create time every candidate
 * register plan (points per thread, hence radix sequence) and tile shape
 * (lines or row pairs per workgroup) of each pass on the new buffers, and for
 * steppers the fixed-plan kernel of each candidate that has one against the
 * runtime-plan kernel, and keep the fastest; choices are cached per
 * (extents, precision) for the process, separately for steppers (with and
 * without noise) and for ensemble groups, which time their own work-table
 * kernels over all the group's members (cached per member count too).  Off:
 * fixed default plans (reproducible bit for bit across processes, and a
 * stepper and an ensemble replica of one shape agree bit for bit). */
int mcs_set_autotune(int on);
int mcs_get_autotune(int* on);
/* Fixed-plan kernels of the fused step and the boundary FFT (no reference
 * counterpart).  For the passes whose bit is set (both by default), a pass whose
 * plan is one the library holds compiled with its whole radix sequence as
 * constants (axis lengths 128, 256, 512; FP32 and FP64) runs that kernel; it
 * computes the same operations in the same order as the runtime-plan kernel, so
 * results are identical either way.  A cleared bit runs the runtime-plan
 * kernels. */
#define MCS_FIXED_COLUMNS 1 /* column passes: the step's y and z, the boundary FFT's y and z, state import/export */
#define MCS_FIXED_XPASS 2   /* x passes: the step's x pass with the update, the boundary FFT's row passes */
int mcs_set_fixed_kernels(int mask);
int mcs_get_fixed_kernels(int* mask);
/* Lane shuffles in the fixed-plan x passes of the step (no reference counterpart; replaces the
 * LDS exchanges inside cuFFT, cufft_wrapper.hpp:50-58).  Where a line's threads are lanes of
 * one wavefront and its plan ends in two passes of radix P (points per thread), the exchange
 * before the last pass runs as a register/lane transpose: DPP quad permutations and row
 * rotations and gfx950's v_permlane16_swap / v_permlane32_swap, no LDS and no barrier.  The
 * same values land in the same registers, so results are identical either way.  mode 0: never;
 * 1 (default): where the autotuner timed that form faster; 2: wherever a plan has it. */
int mcs_set_lane_shuffle(int mode);
int mcs_get_lane_shuffle(int* mode);
/* Exact checkpoint of the stepper (no reference counterpart; the simulators saved psi
 * with CPUArray::save, array_abstraction.hpp:155-180).  In the spectral form the state is
 * the packed spectrum U = FFT(u) of u(m,y,z) = psi(2m,y,z) + i psi(2m+1,y,z) over
 * (nx/2) x ny x nz, complex, unnormalised; in the mixed form it is S = IFFT_x(U) / (nx/2)
 * (see MCS_FORM_MIXED below); nx*ny*nz reals of the stepper's precision either way, plus
 * the FP64 mass (sum of psi) and the step number (the noise draw).  spec may be host or
 * device memory.  Restoring into a stepper with the same plans and form continues bit for
 * bit. */
int mcs_stepper_get_state(mcs_stepper* st, void* spec, double* mass, int64_t* step);
int mcs_stepper_set_state(mcs_stepper* st, const void* spec, double mass, int64_t step);
/* The form of the stepper's state (no reference counterpart; the reference keeps the
 * half spectrum between steps and runs c2r/r2c around the nonlinearity).
 *   MCS_FORM_SPECTRAL: the state is the packed spectrum U above.  A 3D step moves 10 sweeps
 *     of N/2 complex values (the x pass reads and writes U and the work array), a 2D step 6.
 *   MCS_FORM_MIXED: the state is S = IFFT_x(U) / (nx/2), real along x and spectral along y
 *     and z, same size.  The x pass reads the work array and S and writes S only: 9 sweeps
 *     per 3D step, 5 per 2D step; the x pass does three pairs of line transforms instead of
 *     two.  The same update on the same modes; the results differ from the spectral form's
 *     by rounding only, but the state itself then passes through an x transform and back
 *     every step, which in FP32 roughly doubles the error against an exact run (512^3,
 *     500 steps: rel-L2 about 1.0e-4 against 4e-5), for about 5% less time per step.
 * mcs_set_step_form sets the form of steppers created afterwards (MCS_FORM_SPECTRAL by
 * default).
 * mcs_stepper_set_form converts a stepper's state in place (one x transform of the state;
 * the work array stays valid), so the next steps run in the new form; on a stepper without a
 * field and with tuned plans it also tunes the x pass for the new form.  get_state/set_state
 * carry the state of the stepper's current form. */
#define MCS_FORM_SPECTRAL 0
#define MCS_FORM_MIXED 1
int mcs_set_step_form(int form);
int mcs_get_step_form(int* form);
int mcs_stepper_set_form(mcs_stepper* st, int form);
int mcs_stepper_get_form(mcs_stepper* st, int* form);
/* the stepper's plans: {P_x, row pairs per workgroup, P_y, lines per workgroup,
 * P_z, lines per workgroup} (P_z and its tile are 0 in 2D) */
int mcs_stepper_plan_info(mcs_stepper* st, int info[6]);
/* a stepper with explicit plans in the layout above instead of tuned ones (for
 * pinned, reproducible runs and plan studies); each (P, tile) must come from
 * mcs_plan_options / mcs_tile_options, else MCS_ERR_UNSUPPORTED */
int mcs_stepper_create_planned(mcs_stepper** st, const mcs_pfc_params* p, const int plan[6]);
/* register-plan points per thread for an axis of length L (default first; *n is the count) */
int mcs_plan_options(int L, int prec, int* P, int cap, int* n);
/* tile shapes for that plan: row pairs per workgroup (xpass = 1, the packed x axis of
 * length L = nx/2) or lines per workgroup (xpass = 0, the y and z axes) */
int mcs_tile_options(int L, int prec, int P, int xpass, int* G, int cap, int* n);

/* HIP-event timer on the library stream (instrumentation, no reference counterpart) */
int mcs_timer_start(void);
int mcs_timer_stop(double* ms);

/* ---------------------------------------------------------------- ensembles
 * Independent replicas (own grid size, eps, dt, temperature, initial field and
 * noise seed) advanced together.  Replicas of equal size share plans; the size
 * groups whose plans for a pass have the same points per thread are stepped by
 * one launch (persistent workgroups over a (group, replica, work unit) table);
 * all replicas take the same number of steps and share the Philox draw number.  Same size rules as the stepper.  No counterpart in the
 * reference (one replica per process there). */
typedef struct {
    int ext[3];
    double h[3];
    double eps, dt, temp;
    double psi0, amp;     /* initial field psi0 + amp (2u - 1) */
    uint64_t init_seed;   /* Philox key of the initial field */
    uint64_t noise_seed;  /* Philox key of the noise */
} mcs_replica;
typedef struct mcs_ensemble mcs_ensemble;
#define MCS_ENS_LOOP 0    /* replica by replica, single-grid launches (the looping baseline) */
#define MCS_ENS_BATCHED 1 /* eager: one launch per pass and plan class (the size groups whose plans for that
                             pass fall in one register tier share one persistent work-table launch) or one
                             launch per pass and size group, whichever ran faster when created */
#define MCS_ENS_GRAPH 2   /* replayed from a captured HIP graph of 8 steps: either the plan-class launches
                             (the classes of one pass on parallel branches) or the size groups on parallel
                             branches (one per group, or packed onto fewer), whichever replayed faster when
                             the ensemble was created (all compute the same bits); remainder steps run
                             batched */
int mcs_ensemble_create(mcs_ensemble** e, const mcs_replica* reps, int nrep, int split, int prec);
/* seeded initial fields for every replica */
int mcs_ensemble_init(mcs_ensemble* e);
int mcs_ensemble_set_psi(mcs_ensemble* e, int r, const void* psi_dev);
int mcs_ensemble_get_psi(mcs_ensemble* e, int r, void* psi_dev);
int mcs_ensemble_step(mcs_ensemble* e, int nsteps, int mode);
/* enqueue without waiting (mcs_sync() to finish) */
int mcs_ensemble_step_async(mcs_ensemble* e, int nsteps, int mode);
/* per replica (3D replicas): peaks >= ths, k-nearest-neighbour q4/q6 histograms over [0,1),
 * all on the device; counts[nrep], hist_q4/hist_q6[nrep * nbins] on the host */
int mcs_ensemble_analyze(mcs_ensemble* e, double ths, int k, int nbins, int64_t* counts, int64_t* hist_q4,
                         int64_t* hist_q6);
/* nsteps (>= 8) steps replayed from captured graphs, the last graph also holding every
 * (3D) replica's peak finding and k-nearest-neighbour q4/q6 histograms after its final
 * step; outputs as mcs_ensemble_analyze.  Analysis buffers and the graph are kept for
 * later calls with the same (ths, k, nbins). */
int mcs_ensemble_run(mcs_ensemble* e, int nsteps, double ths, int k, int nbins, int64_t* counts, int64_t* hist_q4,
                     int64_t* hist_q6);
int mcs_ensemble_destroy(mcs_ensemble* e);
/* Instrumentation (no reference counterpart): the number of size groups and of plan-class launches
 * per step; whether batched steps launch the classes (1) or per size group (0), and whether the step
 * graph replays the class launches (1) or one branch per size group (0); and the milliseconds per
 * step each form took when the ensemble was created: {batched per group, batched classes, graph per
 * group, graph classes} (-1 where not timed). */
int mcs_ensemble_launch_info(mcs_ensemble* e, int* groups, int* classes_per_step, int* batched_classes,
                             int* graph_classes, double ms_per_step[4]);
/* Pins the launch forms instead of the timed choice (for pinned runs and tests): batched steps and
 * the step graph through the plan classes (1) or per size group (0); the results are the same bits
 * either way.  MCS_ERR_UNSUPPORTED when the ensemble has no classes. */
int mcs_ensemble_set_launch_form(mcs_ensemble* e, int batched_classes, int graph_classes);
/* Instrumentation (no reference counterpart): the branches of the per-group step graph.  The size
 * groups are dealt onto them longest first by their step time alone at creation, and each branch
 * runs its groups one after the other; a process has 4 hardware queues by default, so this chooses
 * which groups share one.  At creation one branch per group and packings onto 4, 3 and 2 branches
 * are timed and the fastest kept.  set: 0 = one branch per group, n > 0 = packed onto n branches
 * (the same bits either way; the graphs are captured again on their next use). */
int mcs_ensemble_graph_branches(mcs_ensemble* e, int* branches);
int mcs_ensemble_set_graph_branches(mcs_ensemble* e, int branches);

/* ---------------------------------------------------------------- peaks
 * call_kernel_findpeaks / _2d (cuda/kernels/findpeaks.hpp:7-63,
 * findpeaks_2d.hpp:7-56): strict periodic local maxima (26 neighbours in 3D,
 * 8 in 2D) with value >= ths.  Output ordered by linear cell index.
 * idx_out (host) receives up to cap linear indices, vals_out (host, may be
 * NULL) the field values as double; *count the number found (MCS_ERR_CAPACITY
 * if count > cap). */
int mcs_findpeaks(const void* real, const int ext[3], double ths, int64_t* idx_out, double* vals_out,
                  int64_t cap, int64_t* count, int prec);
/* No counterpart in the reference: the allocation-free form of mcs_findpeaks, for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same peaks in the same order, written to
 * DEVICE memory: idx_dev[cap] and vals_dev[cap] (vals_dev may be NULL); *count
 * (host) gets the number found, MCS_ERR_CAPACITY when it exceeds cap.  The
 * workspace (peak mask, counts, offsets) is a static 34 MiB device array that each
 * GPU holds until the process ends; grids whose mask does not fit (more than
 * about 2^28 cells when nx is a multiple of 64) return MCS_ERR_UNSUPPORTED before
 * any launch.  Calls on one GPU share that array: calls that overlap on one GPU
 * (several host threads) race and must be serialised by the caller. */
int mcs_findpeaks_noalloc(const void* real, const int ext[3], double ths, int64_t* idx_dev, double* vals_dev,
                          int64_t cap, int64_t* count, int prec);

/* ---------------------------------------------------------------- bond order
 * bond_order_analysis<T> (bond-order-analysis/bond-order.cpp:289-349):
 * neighbours within rlim under the minimum image of a cubic box, q4 q6 q8 w4
 * q4b q6b q8b w4b xi per point (out9[9*i + ...], NaN when nb[i] == 0).
 * Host arrays in and out; computed on the GPU in FP64. */
int mcs_bond_order_cutoff(const double* xyz, int npts, double box, double rlim, int* nb, double* out9);
/* k nearest neighbours on a periodic cubic grid of side n cells (spacing h):
 * integer squared distances, ties broken by the smaller index.  nbr[k*i+j],
 * out4[4*i + {q4,q6,q8,w4}], optional 64-bin histograms of q4 and q6 over [0,1). */
int mcs_bond_order_knn(const int* ijk, int npts, int n, double h, int k, int* nbr, double* out4,
                       int64_t* hist_q4, int64_t* hist_q6, int nbins);

/* ---------------------------------------------------------------- post-processing
 * call_kernel_gaussian_init (cuda/kernels/gaussian.hpp:23-40): kgauss[mode] =
 * (e, 1 - e), e = exp(-(sx^2 kx^2 + sy^2 ky^2 + sz^2 kz^2) / 2), s_i = 1.1 divs h_i,
 * over the half spectrum of the real extent ext (device array). */
int mcs_gaussian_init(void* kgauss, const int ext[3], const double lens[3], const double h[3],
                      double divs_per_sigma, int prec);
/* call_kernel_kmult_real / _imag (gaussian.hpp:42-67) and
 * _call_kernel_multassign_cmpl_with_real / _imag (lowpass.hpp:3-35):
 * karr1[mode] *= part == 0 ? Re karr2[mode] : Im karr2[mode] (device arrays). */
int mcs_kmult(void* karr1, const void* karr2, const int ext[3], int part, int prec);
/* _call_kernel_square (lowpass.hpp:37-53): a <- a^2 over the real array (device). */
int mcs_square(void* real, const int ext[3], int prec);
/* particlepairs/main.cpp:8-53: the distance of every ordered pair (i, j), j != i,
 * under the minimum image of a cubic periodic box, in the order i-major, j
 * ascending (npts*(npts-1) values).  Host arrays; FP64 on the GPU.  *count gets
 * npts*(npts-1); MCS_ERR_CAPACITY when cap is smaller. */
int mcs_pair_distances(const double* xyz, int npts, double box, double* out, int64_t cap, int64_t* count);
/* histogram of the same distances over unordered pairs i < j: nbins (<= 16384)
 * equal bins over [0, rmax); pairs at rmax or beyond are not counted. */
int mcs_pair_histogram(const double* xyz, int npts, double box, double rmax, int nbins, int64_t* hist);

#ifdef __cplusplus
}
#endif
#endif /* MCSTAR_HIP_H */
