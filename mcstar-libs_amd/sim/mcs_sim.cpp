// mcs_sim: a PFC simulator on the kept host API and the fused stepper.
//
// It runs the loop the reference's simulators ran on top of cuda/*.hpp: a
// seeded (or loaded) initial field, semi-implicit conserved steps with k-space
// noise, periodic npy snapshots of psi, mean density and free energy, local
// maxima written in the findpeaks format (cuda/kernels/findpeaks.hpp:45-53) and
// the 13-column bond-order table (bond-order-analysis/bond-order.cpp:333-338),
// the low-pass amplitude of a 2D field with its gz mode dump
// (cuda/kernels/lowpass.hpp:54-93), and an exact checkpoint to resume from.  Flags are --name value pairs; grid
// sizes are written NxNxN or NxN, x first, as common/tclap_defs.hpp:3-30 parses
// them.  Built with -DFLOAT=float or -DFLOAT=double.
//
//   --dims NxNxN       grid (nz = 1 for NxN)          --h 0.7853981633974483
//   --eps 0.3 --dt 0.5 --psi0 -0.25 --amp 0.05 --seed 11
//   --temp 0 --noise-seed 0                          --split cubic|reference
//   --steps 100 --every 0 --out DIR                  --load psi.npy | --resume DIR
//   --ths 0 --rlim 5.5   (peaks >= ths and bond order within rlim at every output)
//   --lowpass DIVS       (2D: amplitude_<step>.npy and lowpass_<step>.gz at every output,
//                         Gaussian width 1.1 DIVS h)
//   --plan Px,Gx,Py,Gy,Pz,Gz  (explicit plans; default: tuned)   --no-tune
//   --form spectral|mixed     (the state's form; a resumed run takes the checkpoint's)
#include <cuda/array_abstraction.hpp>
#include <cuda/cufft_wrapper.hpp>
#include <cuda/math.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

typedef FLOAT Float;
struct Float3 { Float x, y, z; };
typedef Array::CPUArray<Float> CPUArray;
typedef Array::GPUArray<Float> GPUArray;
typedef FFT::CufftPlan<Float, FFT::R2C> PlanR2C;
typedef FFT::CufftPlan<Float, FFT::C2R> PlanC2R;

#include <cuda/kernels/findpeaks.hpp>
#include <cuda/kernels/gaussian.hpp>
#include <cuda/kernels/lowpass.hpp>

namespace {

struct Opts {
    int3 dims{64, 64, 64};
    double h = M_PI / 4, eps = 0.3, dt = 0.5, psi0 = -0.25, amp = 0.05, temp = 0, ths = 0, rlim = 5.5, lowpass = 0;
    int form = -1;  // -1: the library's default
    uint64_t seed = 11, noise_seed = 0;
    int split = MCS_SPLIT_CUBIC, steps = 100, every = 0;
    std::string out = ".", load, resume;
    std::vector<int> plan;
    bool tune = true;
};

[[noreturn]] void usage(const std::string& why) {
    std::cerr << "mcs_sim: " << why << "\n(see the head of mcs_sim.cpp for the flags)" << std::endl;
    std::exit(2);
}

int3 parse_dims(const std::string& s) {
    std::istringstream is(s);
    int3 d{1, 1, 1};
    char x1 = 0, x2 = 0;
    if (!(is >> d.x >> x1 >> d.y) || (x1 != 'x' && x1 != 'X')) usage(s + " is not of form %dx%dx%d or %dx%d");
    if (is >> x2) {
        if ((x2 != 'x' && x2 != 'X') || !(is >> d.z)) usage(s + " is not of form %dx%dx%d or %dx%d");
    }
    return d;
}

std::vector<int> parse_list(const std::string& s) {
    std::vector<int> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::atoi(tok.c_str()));
    return v;
}

Opts parse(int argc, char** argv) {
    Opts o;
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        if (k == "--no-tune") { o.tune = false; continue; }
        if (i + 1 >= argc) usage("missing value for " + k);
        const std::string v = argv[++i];
        if (k == "--dims") o.dims = parse_dims(v);
        else if (k == "--h") o.h = std::atof(v.c_str());
        else if (k == "--eps") o.eps = std::atof(v.c_str());
        else if (k == "--dt") o.dt = std::atof(v.c_str());
        else if (k == "--psi0") o.psi0 = std::atof(v.c_str());
        else if (k == "--amp") o.amp = std::atof(v.c_str());
        else if (k == "--seed") o.seed = std::strtoull(v.c_str(), nullptr, 10);
        else if (k == "--temp") o.temp = std::atof(v.c_str());
        else if (k == "--noise-seed") o.noise_seed = std::strtoull(v.c_str(), nullptr, 10);
        else if (k == "--split") {
            if (v == "cubic") o.split = MCS_SPLIT_CUBIC;
            else if (v == "reference") o.split = MCS_SPLIT_REFERENCE;
            else usage("--split takes cubic or reference");
        } else if (k == "--steps") o.steps = std::atoi(v.c_str());
        else if (k == "--every") o.every = std::atoi(v.c_str());
        else if (k == "--out") o.out = v;
        else if (k == "--load") o.load = v;
        else if (k == "--resume") o.resume = v;
        else if (k == "--ths") o.ths = std::atof(v.c_str());
        else if (k == "--rlim") o.rlim = std::atof(v.c_str());
        else if (k == "--lowpass") o.lowpass = std::atof(v.c_str());
        else if (k == "--plan") o.plan = parse_list(v);
        else if (k == "--form") {
            if (v == "spectral") o.form = MCS_FORM_SPECTRAL;
            else if (v == "mixed") o.form = MCS_FORM_MIXED;
            else usage("--form takes spectral or mixed");
        }
        else usage("unknown flag " + k);
    }
    if (o.steps < 0 || o.every < 0) usage("--steps and --every must be >= 0");
    if (!o.plan.empty() && o.plan.size() != 6) usage("--plan takes six comma-separated integers");
    return o;
}

// ------------------------------------------------------------------ checkpoint
// DIR/checkpoint.txt holds the parameters, the step, the FP64 mass (printed
// exactly), the plans and the state's form; DIR/state.npy the stepper's state
// (complex, numpy shape (nz, ny, nx/2): the packed spectrum in the spectral
// form, its inverse x transform over nx/2 in the mixed form).
typedef std::complex<Float> Cplx;

void save_checkpoint(const Opts& o, mcs_stepper* st) {
    const int3 d = o.dims;
    std::vector<Cplx> spec((size_t)d.x / 2 * d.y * d.z);
    double mass = 0;
    int64_t step = 0;
    CUERR(mcs_stepper_get_state(st, spec.data(), &mass, &step));
    int plan[6], form = 0;
    CUERR(mcs_stepper_plan_info(st, plan));
    CUERR(mcs_stepper_get_form(st, &form));
    const int shp[3] = {d.z, d.y, d.x / 2};
    aoba::SaveArrayAsNumpy(o.out + "/state.npy", 3, shp, spec.data());
    FILE* f = std::fopen((o.out + "/checkpoint.txt").c_str(), "w");
    if (!f) { std::cerr << "cannot write " << o.out << "/checkpoint.txt" << std::endl; std::exit(1); }
    std::fprintf(f, "dims %dx%dx%d\nprecision %d\nsplit %d\nh %.17g\neps %.17g\ndt %.17g\ntemp %.17g\n", d.x, d.y,
                 d.z, (int)(8 * sizeof(Float)), o.split, o.h, o.eps, o.dt, o.temp);
    std::fprintf(f, "noise_seed %llu\nstep %lld\nmass %.17g\nplan %d,%d,%d,%d,%d,%d\nform %d\n",
                 (unsigned long long)o.noise_seed, (long long)step, mass, plan[0], plan[1], plan[2], plan[3],
                 plan[4], plan[5], form);
    std::fclose(f);
}

struct Resume {
    int64_t step = 0;
    double mass = 0;
    int form = MCS_FORM_SPECTRAL;  // checkpoints without a form line hold the packed spectrum
    std::vector<Cplx> spec;
};

Resume read_checkpoint(Opts& o) {
    std::ifstream in(o.resume + "/checkpoint.txt");
    if (!in) usage("cannot read " + o.resume + "/checkpoint.txt");
    std::map<std::string, std::string> kv;
    std::string k, v;
    while (in >> k >> v) kv[k] = v;
    for (const char* key : {"dims", "precision", "split", "h", "eps", "dt", "temp", "noise_seed", "step", "mass", "plan"})
        if (!kv.count(key)) usage(std::string("checkpoint lacks ") + key);
    if (std::atoi(kv["precision"].c_str()) != (int)(8 * sizeof(Float))) usage("checkpoint precision differs");
    o.dims = parse_dims(kv["dims"]);
    o.split = std::atoi(kv["split"].c_str());
    o.h = std::strtod(kv["h"].c_str(), nullptr);
    o.eps = std::strtod(kv["eps"].c_str(), nullptr);
    o.dt = std::strtod(kv["dt"].c_str(), nullptr);
    o.temp = std::strtod(kv["temp"].c_str(), nullptr);
    o.noise_seed = std::strtoull(kv["noise_seed"].c_str(), nullptr, 10);
    if (o.plan.empty()) o.plan = parse_list(kv["plan"]);  // the same plans continue bit for bit
    Resume r;
    r.step = std::strtoll(kv["step"].c_str(), nullptr, 10);
    r.mass = std::strtod(kv["mass"].c_str(), nullptr);
    if (kv.count("form")) r.form = std::atoi(kv["form"].c_str());
    std::vector<int> shape;
    aoba::LoadArrayFromNumpy(o.resume + "/state.npy", shape, r.spec);
    if (shape.size() != 3 || shape[0] != o.dims.z || shape[1] != o.dims.y || shape[2] != o.dims.x / 2)
        usage("state.npy does not match the checkpoint's dims");
    return r;
}

// ------------------------------------------------------------------ outputs
// plans, buffers and filter table of the 2D low-pass amplitude, made once
struct Lowpass {
    PlanR2C r2c;
    PlanC2R c2r;
    GPUArray spec, aux, kg;
    Float3 lens;
    Lowpass(const Opts& o)
        : r2c(o.dims), c2r(o.dims), spec(o.dims), aux(o.dims), kg(o.dims),
          lens{(Float)(o.h * o.dims.x), (Float)(o.h * o.dims.y), (Float)o.h} {
        const Float3 hh{(Float)o.h, (Float)o.h, (Float)o.h};
        call_kernel_gaussian_init(kg, lens, hh, (Float)o.lowpass);
    }
};

void write_outputs(const Opts& o, mcs_stepper* st, GPUArray& psi, CPUArray& host, int64_t step, double ms_per_step,
                   Lowpass* lp) {
    CUERR(mcs_stepper_get_psi(st, psi.ptr_void()));
    char tag[64];
    std::snprintf(tag, sizeof tag, "%08lld", (long long)step);
    psi >> host;
    host.save(o.out + "/psi_" + tag + ".npy");
    const double hh[3] = {o.h, o.h, o.h};
    double mean = 0, fen = 0;
    CUERR(mcs_field_stats(psi.ptr_void(), psi.ext_ptr(), hh, o.eps, Array::prec_of<Float>(), &mean, &fen));
    double mass = 0;
    CUERR(mcs_stepper_mass(st, &mass));
    const std::string logfn = o.out + "/log.txt";
    FILE* lg = std::fopen(logfn.c_str(), "r");
    const bool fresh = lg == nullptr;
    if (lg) std::fclose(lg);
    lg = std::fopen(logfn.c_str(), "a");
    if (lg) {
        if (fresh) std::fprintf(lg, "# step mean(k=0, FP64) mean(field) free_energy ms_per_step\n");
        std::fprintf(lg, "%lld %.17g %.17g %.17g %.6f\n", (long long)step, mass, mean, fen, ms_per_step);
        std::fclose(lg);
    }
    if (o.dims.z > 1 && o.dims.x == o.dims.y && o.dims.x == o.dims.z) {
        const Float3 h3{(Float)o.h, (Float)o.h, (Float)o.h};
        const Float3 len{(Float)(o.h * o.dims.x), (Float)(o.h * o.dims.y), (Float)(o.h * o.dims.z)};
        call_kernel_findpeaks(psi, o.out + "/peaks_" + tag + ".txt", o.out + "/bondorder_" + tag + ".txt", h3, len,
                              (Float)o.ths, (Float)o.rlim);
    }
    if (lp) {
        lp->r2c.execute(psi, lp->spec);
        calc_amplitude_lowpass(lp->r2c, lp->c2r, lp->spec, lp->aux, lp->kg, host, lp->lens,
                               o.out + "/lowpass_" + tag + ".gz");
        host.save(o.out + "/amplitude_" + tag + ".npy");
    }
}

}  // namespace

int main(int argc, char** argv) {
    Opts o = parse(argc, argv);
    CUERR(mcs_init(0));
    Resume res;
    const bool resuming = !o.resume.empty();
    if (resuming) res = read_checkpoint(o);
    if (!o.tune) CUERR(mcs_set_autotune(0));
    mcs_pfc_params p;
    std::memset(&p, 0, sizeof p);
    p.ext[0] = o.dims.x; p.ext[1] = o.dims.y; p.ext[2] = o.dims.z;
    p.h[0] = p.h[1] = p.h[2] = o.h;
    p.eps = o.eps; p.dt = o.dt; p.temp = o.temp; p.seed = o.noise_seed; p.split = o.split;
    p.prec = Array::prec_of<Float>();
    mcs_stepper* st = nullptr;
    if (o.plan.empty()) CUERR(mcs_stepper_create(&st, &p));
    else CUERR(mcs_stepper_create_planned(&st, &p, o.plan.data()));

    if (o.lowpass > 0 && (o.dims.z != 1 || o.dims.y < 2 || o.dims.x < 2)) usage("--lowpass needs a 2D grid");
    CPUArray host(o.dims);
    GPUArray psi(host);
    std::unique_ptr<Lowpass> lp(o.lowpass > 0 ? new Lowpass(o) : nullptr);
    int64_t step0 = 0;
    if (!resuming && o.form >= 0) CUERR(mcs_stepper_set_form(st, o.form));
    if (resuming) {
        CUERR(mcs_stepper_set_form(st, res.form));
        CUERR(mcs_stepper_set_state(st, res.spec.data(), res.mass, res.step));
        step0 = res.step;
    } else {
        if (!o.load.empty()) {
            host.load(o.load);
            host >> psi;
        } else {
            CUERR(mcs_init_field(psi.ptr_void(), psi.ext_ptr(), o.psi0, o.amp, o.seed, Array::prec_of<Float>()));
        }
        CUERR(mcs_stepper_set_psi(st, psi.ptr_void()));
    }
    std::cerr << "mcs_sim: " << o.dims.x << "x" << o.dims.y << "x" << o.dims.z << " from step " << step0 << ", "
              << o.steps << " steps" << std::endl;
    int done = 0;
    double ms = 0;
    while (done < o.steps) {
        const int chunk = o.every > 0 ? std::min(o.every, o.steps - done) : o.steps - done;
        const auto t0 = std::chrono::steady_clock::now();
        CUERR(mcs_stepper_step(st, chunk));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / chunk;
        done += chunk;
        if (o.every > 0 || done == o.steps) write_outputs(o, st, psi, host, step0 + done, ms, lp.get());
    }
    if (o.steps == 0) write_outputs(o, st, psi, host, step0, 0.0, lp.get());
    save_checkpoint(o, st);
    CUERR(mcs_stepper_destroy(st));
    return 0;
}
