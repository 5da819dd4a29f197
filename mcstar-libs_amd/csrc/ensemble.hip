// Batched ensembles of independent PFC replicas.
//
// Replicas are grouped by grid size; a group shares the FFT plans, twiddles
// and row-pair table, and each replica has a descriptor (state pointers and
// its own eps, dt, noise amplitude, seed and FP64 mass).  One step of the
// ensemble is one launch per pass and group: persistent workgroups walk the
// (replica, work unit) table with the same pass bodies as the single-grid
// stepper.  The x pass of step j of a call draws noise number d_step + j,
// and one single-thread kernel advances the device counter d_step at the end
// of the call, so the passes of different groups need no ordering among
// themselves.  MCS_ENS_LOOP issues the same passes replica by replica (the
// looping baseline); MCS_ENS_GRAPH replays a captured graph of several steps
// in which every size group is an independent branch, so that the groups'
// kernels overlap on the GPU.
#include "ens_kernels.hpp"
#include "fixed.hpp"
#include "knn.hpp"
#include "peaks.hpp"
#include "stepper.hpp"

#include <algorithm>
#include <map>
#include <vector>

__global__ void k_ens_advance(int64_t* d_step, int n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_step += n;
}

// clears histogram bins inside a captured graph (a kernel node, not a memset node)
__global__ void k_zero_u64(unsigned long long* p, long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0ull;
}

// ------------------------------------------------------------------ host side
struct EnsGroup {
    StepGeom sg;
    std::vector<int> members;
    void* d_reps = nullptr;  // EnsRep<F>[members]
};

// A plan class: one launch of one pass slot (0 x; 1 y inverse, or the 2D nonlinear y pass; 2 z
// nonlinear; 3 y forward) over the size groups whose plan for it falls in one register tier of
// points per thread and whose tiles make T-thread workgroups.
enum { CLS_X = 0, CLS_COL_INV = 1, CLS_COL_NL = 2, CLS_COL_FWD = 3 };
struct EnsClass {
    int slot = 0, kind = 0, tier = 0, T = 0;
    size_t lds = 0;
    long total = 0;
    bool noise = false;
    void* d_desc = nullptr;  // EnsClassGroup<F>[groups]
    std::vector<int> groups;
};

struct mcs_ensemble {
    std::vector<mcs_replica> reps;
    int split = 0, prec = 32;
    std::vector<void*> U, W;
    std::vector<double> mass;
    std::vector<EnsGroup*> groups;
    std::vector<int> group_of, slot_of;
    int64_t* d_step = nullptr;
    int64_t step = 0;
    bool loaded = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int gsteps = 0;
    std::vector<hipStream_t> side;     // one capture branch per size group
    std::vector<hipEvent_t> side_ev;   // [0] fork, [1 + gi] join of group gi
    int cus = 0;
    // in-graph analysis (mcs_ensemble_run): the last GRAPH_STEPS steps followed by peaks and
    // k-nearest-neighbour q4/q6 histograms of every replica, on per-group device buffers
    hipGraphExec_t aexec = nullptr;
    hipGraph_t agraph = nullptr;
    double a_ths = 0;
    int a_k = 0, a_nbins = 0;
    std::vector<void*> a_bufs;               // every allocation below, freed together
    std::vector<void*> a_psi, a_tmp, a_ws;   // per group: one replica's field, export scratch, peak workspace
    std::vector<int64_t*> a_idx;             // per group: peak indices (capacity n/8 + 8)
    std::vector<void*> a_knn;                // per group: bond-order cell-list workspace
    int64_t* a_counts = nullptr;             // [nrep]
    unsigned long long* a_hist = nullptr;    // [nrep][2][nbins]: q4 then q6
    // plan classes: per pass slot, the launches that cover every size group sharing a plan
    std::vector<EnsClass> cls;
    bool graph_classes = false;              // the step graph replays class launches (else group branches)
    bool batched_classes = false;            // eager batched steps launch the classes (else per group)
    double strat_ms[4] = {-1, -1, -1, -1};   // per step: eager per group, eager classes, graph per group, graph classes
    std::vector<hipStream_t> cside;          // fork streams of the class launches of one pass
    std::vector<hipEvent_t> cev;             // [0] fork, [1 + i] join of class i
    // branches of the per-group step graph: the size groups each branch runs one after the other
    // (empty: one branch per group)
    std::vector<std::vector<int>> packs;
    std::vector<double> gcost;               // per size group: ms per step launched alone (at creation)
};


static size_t elem_size(int prec) { return prec == 64 ? 16 : 8; }

static StepPhys phys_of(const mcs_replica& r, int split, double mass) {
    StepPhys ph;
    ph.h[0] = r.h[0]; ph.h[1] = r.h[1]; ph.h[2] = r.h[2];
    ph.eps = r.eps; ph.dt = r.dt; ph.temp = r.temp; ph.seed = r.noise_seed; ph.split = split; ph.mass = mass;
    return ph;
}

template <typename F> static int upload_reps_t(mcs_ensemble* e, EnsGroup* g) {
    std::vector<EnsRep<F>> h(g->members.size());
    for (size_t i = 0; i < g->members.size(); ++i) {
        const int r = g->members[i];
        h[i].U = (C2<F>*)e->U[r];
        h[i].W = (C2<F>*)e->W[r];
        h[i].xa = make_xargs<F>(g->sg, phys_of(e->reps[r], e->split, e->mass[r]));
    }
    if (!g->d_reps) MCS_HIP_TRY(hipMalloc(&g->d_reps, h.size() * sizeof(EnsRep<F>)));
    MCS_HIP_TRY(hipMemcpy(g->d_reps, h.data(), h.size() * sizeof(EnsRep<F>), hipMemcpyHostToDevice));
    return MCS_OK;
}

static int upload_reps(mcs_ensemble* e) {
    for (EnsGroup* g : e->groups) {
        int st = e->prec == 64 ? upload_reps_t<double>(e, g) : upload_reps_t<float>(e, g);
        if (st) return st;
    }
    return MCS_OK;
}

// resident workgroups per CU from the thread (2048 per CU) and LDS (160 KB) budgets, times 2 for slack;
// a multiple of 8 below the unit count (so that a unit's XCD follows from its index)
static unsigned persistent_grid(mcs_ensemble* e, int threads, size_t lds, long total) {
    long per = 2048 / std::max(threads, 64);
    if (lds > 0) per = std::min<long>(per, (long)(MCS_LDS_MAX / lds));
    per = std::max<long>(per, 1) * 2;
    long grid = per * e->cus;
    if (grid < total) grid = std::max<long>(8, grid & ~7L);
    return (unsigned)std::max<long>(1, std::min<long>(total, grid));
}

static int ens_lds_attr(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

// the tile order of a work-table column pass: where neighbouring tiles share cache lines (tiles
// narrower than a 128-byte line, or rows not aligned to one), the XCD-aware order of
// col_tile_order keeps a shared line in one L2; elsewhere plain order
static int ens_tile_order(const ColGeom& g, int G, long esz) {
    const bool shared = (G * esz) % 128 != 0 || (g.nxc * esz) % 128 != 0 || (g.os * esz) % 128 != 0;
    return shared ? col_tile_order(g, G, esz) : 0;
}

// one x pass over members [first, first + nr) of group g
template <typename F, int P, int SPLIT>
static int ens_xupd_t(mcs_ensemble* e, EnsGroup* g, int first, int nr, int draw_off, hipStream_t s) {
    const LinePlan& lp = g->sg.px;
    const int G = xupd_pairs_per_wg(lp);
    const int S = lp.L / P;
    const size_t exch = (size_t)G * 2 * lp.L * sizeof(C2<F>);
    const int units = (g->sg.npairs + G - 1) / G;
    const int twn = twiddles_in_lds(lp, exch, (long)nr * units);
    const size_t lds = exch + (size_t)twn * sizeof(C2<F>);  // exchange area, twiddles
    bool noise = false;  // the noise-free kernels carry no noise draw
    for (int i = first; i < first + nr; ++i) noise = noise || e->reps[g->members[i]].temp > 0;
    auto fn = noise ? ((64 % S == 0) ? k_ens_xupd<F, P, SPLIT, true, RunPlan, true>
                                     : k_ens_xupd<F, P, SPLIT, false, RunPlan, true>)
                    : ((64 % S == 0) ? k_ens_xupd<F, P, SPLIT, true, RunPlan, false>
                                     : k_ens_xupd<F, P, SPLIT, false, RunPlan, false>);
    const unsigned grid = persistent_grid(e, G * S, lds, (long)nr * units);
    int st = ens_lds_attr((const void*)fn, lds);
    if (st) return st;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(G * S), lds, s, (const EnsRep<F>*)g->d_reps + first, nr, units,
                       g->sg.pairs, g->sg.npairs, G, lp.dev, (const C2<F>*)lp.tw, (const C2<F>*)g->sg.wph,
                       (const int64_t*)e->d_step, draw_off, twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int P, int DIR>
static int ens_col_t(mcs_ensemble* e, EnsGroup* g, const LinePlan& lp, const ColGeom& cg, int first, int nr,
                     hipStream_t s) {
    const int G = col_lines_per_wg(lp);
    const int S = lp.L / P;
    const size_t exch = (size_t)G * lp.L * sizeof(C2<F>);
    const long total = (long)nr * ((cg.nxc + G - 1) / G) * cg.nouter;
    int twn = twiddles_in_lds(lp, exch, total);
    const size_t lds = exch + (size_t)twn * sizeof(C2<F>);  // exchange area, twiddles
    const void* fx = fixed_ens_col_kernel(lp, DIR, 0);
    const void* fn = fx ? fx : (const void*)k_ens_col<F, P, DIR>;
    const unsigned grid = persistent_grid(e, G * S, lds, total);
    const EnsRep<F>* reps = (const EnsRep<F>*)g->d_reps + first;
    FftPlanDev pl = lp.dev;
    const C2<F>* tw = (const C2<F>*)lp.tw;
    ColGeom geo = cg;
    int gg = G, n = nr;
    int swz = ens_tile_order(cg, G, (long)sizeof(C2<F>));
    MCS_CHECK(total < (1L << 32), MCS_ERR_UNSUPPORTED, "too many column units in one launch");
    int st = ens_lds_attr(fn, lds);
    if (st) return st;
    // the fixed-plan kernels have the runtime-plan kernel's signature: launched through its type,
    // so that the arguments are checked against the parameters when this compiles
    const auto kfn = reinterpret_cast<decltype(&k_ens_col<F, P, DIR>)>(const_cast<void*>(fn));
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(G * S), lds, s, reps, n, geo, gg, pl, tw, swz, twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F, int P, int SPLIT>
static int ens_col_nl_t(mcs_ensemble* e, EnsGroup* g, const LinePlan& lp, const ColGeom& cg, int first, int nr,
                        hipStream_t s) {
    const int G = col_lines_per_wg(lp);
    const int S = lp.L / P;
    const size_t exch = (size_t)G * lp.L * sizeof(C2<F>);
    const long total = (long)nr * ((cg.nxc + G - 1) / G) * cg.nouter;
    int twn = twiddles_in_lds(lp, exch, total);
    const size_t lds = exch + (size_t)twn * sizeof(C2<F>);  // exchange area, twiddles
    const void* fx = fixed_ens_col_kernel(lp, 0, SPLIT);
    const void* fn = fx ? fx : (const void*)k_ens_col_nl<F, P, SPLIT>;
    const unsigned grid = persistent_grid(e, G * S, lds, total);
    const EnsRep<F>* reps = (const EnsRep<F>*)g->d_reps + first;
    FftPlanDev pl = lp.dev;
    const C2<F>* tw = (const C2<F>*)lp.tw;
    ColGeom geo = cg;
    int gg = G, n = nr;
    F scale = (F)(2.0 / ((double)g->sg.nx * g->sg.ny * g->sg.nz));
    int swz = ens_tile_order(cg, G, (long)sizeof(C2<F>));
    MCS_CHECK(total < (1L << 32), MCS_ERR_UNSUPPORTED, "too many column units in one launch");
    int st = ens_lds_attr(fn, lds);
    if (st) return st;
    const auto kfn = reinterpret_cast<decltype(&k_ens_col_nl<F, P, SPLIT>)>(const_cast<void*>(fn));
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(G * S), lds, s, reps, n, geo, gg, pl, tw, scale, swz, twn);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

#define MCS_P_DISPATCH(Pv, CALL)                                                            \
    switch (Pv) {                                                                           \
        case 4: { constexpr int PP = 4; return CALL; }                                      \
        case 8: { constexpr int PP = 8; return CALL; }                                      \
        case 6: { constexpr int PP = 6; return CALL; }                                      \
        case 10: { constexpr int PP = 10; return CALL; }                                    \
        case 12: { constexpr int PP = 12; return CALL; }                                    \
        case 20: { constexpr int PP = 20; return CALL; }                                    \
        case 16: { constexpr int PP = 16; return CALL; }                                    \
        case 32: { constexpr int PP = 32; return CALL; }                                    \
        case 30: { constexpr int PP = 30; return CALL; }                                    \
        default: mcs_set_last_error(__FILE__, __LINE__, "no register plan"); return MCS_ERR_UNSUPPORTED; \
    }

template <typename F>
static int ens_xupd(mcs_ensemble* e, EnsGroup* g, int first, int nr, int draw_off, hipStream_t s) {
    if (e->split == MCS_SPLIT_REFERENCE) { MCS_P_DISPATCH(g->sg.px.P, (ens_xupd_t<F, PP, MCS_SPLIT_REFERENCE>(e, g, first, nr, draw_off, s))) }
    MCS_P_DISPATCH(g->sg.px.P, (ens_xupd_t<F, PP, MCS_SPLIT_CUBIC>(e, g, first, nr, draw_off, s)))
}

template <typename F>
static int ens_col(mcs_ensemble* e, EnsGroup* g, const LinePlan& lp, const ColGeom& cg, int dir, int first, int nr,
                   hipStream_t s) {
    if (dir < 0) { MCS_P_DISPATCH(lp.P, (ens_col_t<F, PP, -1>(e, g, lp, cg, first, nr, s))) }
    MCS_P_DISPATCH(lp.P, (ens_col_t<F, PP, +1>(e, g, lp, cg, first, nr, s)))
}

template <typename F>
static int ens_col_nl(mcs_ensemble* e, EnsGroup* g, const LinePlan& lp, const ColGeom& cg, int first, int nr,
                      hipStream_t s) {
    if (e->split == MCS_SPLIT_REFERENCE) { MCS_P_DISPATCH(lp.P, (ens_col_nl_t<F, PP, MCS_SPLIT_REFERENCE>(e, g, lp, cg, first, nr, s))) }
    MCS_P_DISPATCH(lp.P, (ens_col_nl_t<F, PP, MCS_SPLIT_CUBIC>(e, g, lp, cg, first, nr, s)))
}

// one step of members [first, first+nr) of g, drawing noise number d_step + draw_off
template <typename F>
static int ens_step_group(mcs_ensemble* e, EnsGroup* g, int first, int nr, int draw_off, hipStream_t s) {
    const StepGeom& sg = g->sg;
    int st = ens_xupd<F>(e, g, first, nr, draw_off, s);
    if (st) return st;
    if (sg.is3d) {
        st = ens_col<F>(e, g, sg.py, step_geom_y(sg), +1, first, nr, s);
        if (!st) st = ens_col_nl<F>(e, g, sg.pz, step_geom_z(sg), first, nr, s);
        if (!st) st = ens_col<F>(e, g, sg.py, step_geom_y(sg), -1, first, nr, s);
    } else {
        st = ens_col_nl<F>(e, g, sg.py, step_geom_y(sg), first, nr, s);
    }
    return st;
}

// ------------------------------------------------------------------ plan classes
static void ens_classes_free(mcs_ensemble* e) {
    for (EnsClass& c : e->cls)
        if (c.d_desc) (void)hipFree(c.d_desc);
    e->cls.clear();
}

// the classes of every pass slot: a group joins the class of its (slot, kind, register tier) and
// workgroup size T = G * S, with G the most lines (row pairs) of S = L / P threads that fit 256
// threads, the launch bound and the LDS rule (T = 256 whenever S divides 256).  A line of more
// than 256 threads fits no class: then the ensemble keeps per-group launches.
template <typename F> static int ens_build_classes_t(mcs_ensemble* e) {
    ens_classes_free(e);
    struct Item { int gi, slot, kind; const LinePlan* lp; ColGeom g; };
    std::vector<Item> items;
    for (int gi = 0; gi < (int)e->groups.size(); ++gi) {
        const StepGeom& sg = e->groups[gi]->sg;
        items.push_back({gi, 0, CLS_X, &sg.px, ColGeom{}});
        if (sg.is3d) {
            items.push_back({gi, 1, CLS_COL_INV, &sg.py, step_geom_y(sg)});
            items.push_back({gi, 2, CLS_COL_NL, &sg.pz, step_geom_z(sg)});
            items.push_back({gi, 3, CLS_COL_FWD, &sg.py, step_geom_y(sg)});
        } else {
            items.push_back({gi, 1, CLS_COL_NL, &sg.py, step_geom_y(sg)});
        }
    }
    std::map<std::tuple<int, int, int, int>, std::vector<std::pair<int, int>>> by;  // key -> (item, G)
    for (int i = 0; i < (int)items.size(); ++i) {
        const Item& it = items[i];
        const LinePlan& lp = *it.lp;
        if (lp.engine != ENGINE_REG || lp.P <= 0) return MCS_OK;
        const int S = lp.L / lp.P;
        int G = 256 / S;
        while (G >= 1 && !(it.kind == CLS_X ? x_tile_ok(lp.L, lp.P, e->prec, G) : col_tile_ok(lp.L, lp.P, e->prec, G)))
            --G;
        if (G < 1) return MCS_OK;
        by[std::make_tuple(it.slot, it.kind, cls_tier(lp.P), G * S)].push_back({i, G});
    }
    for (auto& kv : by) {
        EnsClass c;
        c.slot = std::get<0>(kv.first);
        c.kind = std::get<1>(kv.first);
        c.tier = std::get<2>(kv.first);
        c.T = std::get<3>(kv.first);
        std::vector<EnsClassGroup<F>> h;
        long first = 0;
        for (const auto& ig : kv.second) {
            const Item& it = items[ig.first];
            const int G = ig.second;
            EnsGroup* g = e->groups[it.gi];
            const LinePlan& lp = *it.lp;
            EnsClassGroup<F> d{};
            d.pl = lp.dev;
            d.tw = (const C2<F>*)lp.tw;
            d.reps = (const EnsRep<F>*)g->d_reps;
            d.pairs = g->sg.pairs;
            d.wph = (const C2<F>*)g->sg.wph;
            d.g = it.g;
            d.scale = (F)(2.0 / ((double)g->sg.nx * g->sg.ny * g->sg.nz));
            d.npairs = g->sg.npairs;
            d.P = lp.P;
            d.G = G;
            d.units = it.kind == CLS_X ? (g->sg.npairs + G - 1) / G : ((it.g.nxc + G - 1) / G) * it.g.nouter;
            d.first = first;
            first += (long)d.units * (long)g->members.size();
            d.total = first;
            h.push_back(d);
            const size_t lds = (size_t)G * (it.kind == CLS_X ? 2 : 1) * lp.L * sizeof(C2<F>);
            c.lds = std::max(c.lds, lds);
            c.groups.push_back(it.gi);
            for (int r : g->members) c.noise = c.noise || (it.kind == CLS_X && e->reps[r].temp > 0);
        }
        c.total = first;
        e->cls.push_back(c);
        MCS_HIP_TRY(hipMalloc(&e->cls.back().d_desc, h.size() * sizeof(EnsClassGroup<F>)));
        MCS_HIP_TRY(hipMemcpy(e->cls.back().d_desc, h.data(), h.size() * sizeof(EnsClassGroup<F>),
                              hipMemcpyHostToDevice));
    }
    return MCS_OK;
}

static int cls_lds_attr(const void* fn, size_t bytes) {
    if (bytes > 65536) MCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return MCS_OK;
}

template <typename F, int TIER>
static int cls_launch_t(mcs_ensemble* e, const EnsClass& c, int draw_off, hipStream_t s) {
    const EnsClassGroup<F>* d = (const EnsClassGroup<F>*)c.d_desc;
    const unsigned grid = persistent_grid(e, c.T, c.lds, c.total);
    const void* fn = nullptr;
    const bool ref = e->split == MCS_SPLIT_REFERENCE;
    long total = c.total;
    if (c.kind == CLS_X) {
        const auto kfn = c.noise ? k_cls_xupd<F, TIER, true> : k_cls_xupd<F, TIER, false>;
        fn = (const void*)kfn;
        int st = cls_lds_attr(fn, c.lds);
        if (st) return st;
        hipLaunchKernelGGL(kfn, dim3(grid), dim3(c.T), c.lds, s, d, total, (const int64_t*)e->d_step, draw_off);
        MCS_HIP_TRY(hipGetLastError());
        return MCS_OK;
    }
    const auto kfn = c.kind == CLS_COL_NL ? (ref ? k_cls_col_nl<F, TIER, MCS_SPLIT_REFERENCE>
                                                 : k_cls_col_nl<F, TIER, MCS_SPLIT_CUBIC>)
                                          : (c.kind == CLS_COL_INV ? k_cls_col<F, TIER, +1> : k_cls_col<F, TIER, -1>);
    fn = (const void*)kfn;
    int st = cls_lds_attr(fn, c.lds);
    if (st) return st;
    hipLaunchKernelGGL(kfn, dim3(grid), dim3(c.T), c.lds, s, d, total);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

template <typename F> static int cls_launch(mcs_ensemble* e, const EnsClass& c, int draw_off, hipStream_t s) {
    return c.tier == 0 ? cls_launch_t<F, 0>(e, c, draw_off, s) : cls_launch_t<F, 1>(e, c, draw_off, s);
}

static int ens_class_streams(mcs_ensemble* e, size_t n) {
    while (e->cside.size() < n) {
        hipStream_t b = nullptr;
        MCS_HIP_TRY(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        e->cside.push_back(b);
    }
    while (e->cev.size() < n + 1) {
        hipEvent_t ev = nullptr;
        MCS_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->cev.push_back(ev);
    }
    return MCS_OK;
}

// one step as class launches, pass slot after pass slot; with fork, the classes of a slot run on
// their own streams between a fork and a join on s
template <typename F> static int ens_class_step(mcs_ensemble* e, int draw_off, hipStream_t s, bool fork) {
    for (int slot = 0; slot < 4; ++slot) {
        std::vector<const EnsClass*> cs;
        for (const EnsClass& c : e->cls)
            if (c.slot == slot) cs.push_back(&c);
        if (cs.empty()) continue;
        if (!fork || cs.size() == 1) {
            for (const EnsClass* c : cs) {
                int st = cls_launch<F>(e, *c, draw_off, s);
                if (st) return st;
            }
            continue;
        }
        int st = ens_class_streams(e, cs.size());
        if (st) return st;
        MCS_HIP_TRY(hipEventRecord(e->cev[0], s));
        for (size_t i = 0; i < cs.size(); ++i) {
            MCS_HIP_TRY(hipStreamWaitEvent(e->cside[i], e->cev[0], 0));
            st = cls_launch<F>(e, *cs[i], draw_off, e->cside[i]);
            if (st) return st;
            MCS_HIP_TRY(hipEventRecord(e->cev[1 + i], e->cside[i]));
        }
        for (size_t i = 0; i < cs.size(); ++i) MCS_HIP_TRY(hipStreamWaitEvent(s, e->cev[1 + i], 0));
    }
    return MCS_OK;
}

// plans and tile shapes of a size group, timed on the group's own work-table kernels over all
// its members (the trial launches overwrite their states, which are loaded afterwards); cached
// per shape, member count and noise
template <typename F> static int ens_tune_group(mcs_ensemble* e, EnsGroup* g, hipStream_t s) {
    const int n = (int)g->members.size();
    bool noisy = false;
    for (int r : g->members) noisy = noisy || e->reps[r].temp > 0;
    const ColGeom gy = step_geom_y(g->sg), gz = step_geom_z(g->sg);
    const bool is3d = g->sg.is3d;
    return step_geom_tune_runs(
        &g->sg, 2 + (noisy ? 1 : 0) + 4 * n, MCS_FIXED_COLUMNS,
        [&](const LinePlan& lp) {
            EnsGroup t = *g;
            t.sg.px = lp;
            return ens_xupd<F>(e, &t, 0, n, 0, s);
        },
        [&](const LinePlan& lp) { return is3d ? ens_col<F>(e, g, lp, gy, -1, 0, n, s) : ens_col_nl<F>(e, g, lp, gy, 0, n, s); },
        [&](const LinePlan& lp) { return ens_col_nl<F>(e, g, lp, gz, 0, n, s); }, s);
}

static int ens_advance(mcs_ensemble* e, int nsteps, hipStream_t s) {
    hipLaunchKernelGGL(k_ens_advance, dim3(1), dim3(64), 0, s, e->d_step, nsteps);
    MCS_HIP_TRY(hipGetLastError());
    return MCS_OK;
}

// nsteps steps in one stream: replica by replica (LOOP) or one launch per pass and group (BATCHED)
template <typename F> static int ens_steps_t(mcs_ensemble* e, int nsteps, int mode, hipStream_t s) {
    const int ng = (int)e->groups.size();
    int st = MCS_OK;
    for (int it = 0; it < nsteps && !st; ++it) {
        if (mode == MCS_ENS_LOOP) {
            for (int r = 0; r < (int)e->reps.size() && !st; ++r)
                st = ens_step_group<F>(e, e->groups[e->group_of[r]], e->slot_of[r], 1, it, s);
        } else if (e->batched_classes) {
            st = ens_class_step<F>(e, it, s, false);
        } else {
            for (int gi = 0; gi < ng && !st; ++gi)
                st = ens_step_group<F>(e, e->groups[gi], 0, (int)e->groups[gi]->members.size(), it, s);
        }
    }
    if (!st && nsteps > 0) st = ens_advance(e, nsteps, s);
    return st;
}

static int ens_steps(mcs_ensemble* e, int nsteps, int mode, hipStream_t s) {
    return e->prec == 64 ? ens_steps_t<double>(e, nsteps, mode, s) : ens_steps_t<float>(e, nsteps, mode, s);
}

static int64_t peak_cap(const mcs_replica& p) { return (int64_t)p.ext[0] * p.ext[1] * p.ext[2] / 8 + 8; }

// analysis of the members of group gi on stream b: export, peaks >= ths, kNN q4/q6 histograms
static int ens_analysis_group(mcs_ensemble* e, int gi, hipStream_t b) {
    EnsGroup* g = e->groups[gi];
    const int nb = e->a_nbins;
    for (int r : g->members) {
        const mcs_replica& p = e->reps[r];
        unsigned long long* h = e->a_hist + 2L * nb * r;
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, b, h, 2L * nb);
        MCS_HIP_TRY(hipGetLastError());
        int st = step_export(g->sg, e->U[r], e->a_psi[gi], e->a_tmp[gi], b);
        if (!st) st = findpeaks_launch(e->a_psi[gi], p.ext, e->a_ths, e->a_idx[gi], nullptr, peak_cap(p),
                                       e->a_counts + r, e->a_ws[gi], e->prec, b);
        if (!st) st = knn_launch(e->a_idx[gi], nullptr, e->a_counts + r, 0, peak_cap(p), p.ext, p.h[0], e->a_k,
                                 e->a_knn[gi], nullptr, nullptr, h, h + nb, nb, b);
        if (st) return st;
    }
    return MCS_OK;
}

// GRAPH_STEPS steps with every size group on its own branch: fork from s, each
// group runs its steps in order (then, with `analysis`, its replicas' analysis),
// join, then advance the draw counter
template <typename F> static int ens_branches_t(mcs_ensemble* e, int nsteps, hipStream_t s, bool analysis) {
    const int ng = (int)e->groups.size();
    int st = MCS_OK;
    if (e->graph_classes) {
        // class launches pass after pass, then (with analysis) every group's analysis on its branch
        for (int it = 0; it < nsteps && !st; ++it) st = ens_class_step<F>(e, it, s, true);
        if (!st && analysis) {
            if (hipEventRecord(e->side_ev[0], s) != hipSuccess) return MCS_ERR_HIP;
            for (int gi = 0; gi < ng && !st; ++gi) {
                hipStream_t b = e->side[gi];
                if (hipStreamWaitEvent(b, e->side_ev[0], 0) != hipSuccess) return MCS_ERR_HIP;
                st = ens_analysis_group(e, gi, b);
                if (!st && hipEventRecord(e->side_ev[1 + gi], b) != hipSuccess) st = MCS_ERR_HIP;
                if (!st && hipStreamWaitEvent(s, e->side_ev[1 + gi], 0) != hipSuccess) st = MCS_ERR_HIP;
            }
        }
        if (!st) st = ens_advance(e, nsteps, s);
        return st;
    }
    std::vector<std::vector<int>> packs = e->packs;
    if (packs.empty())
        for (int gi = 0; gi < ng; ++gi) packs.push_back({gi});
    if (hipEventRecord(e->side_ev[0], s) != hipSuccess) return MCS_ERR_HIP;
    for (size_t bi = 0; bi < packs.size() && !st; ++bi) {
        hipStream_t b = e->side[bi];
        if (hipStreamWaitEvent(b, e->side_ev[0], 0) != hipSuccess) return MCS_ERR_HIP;
        for (int gi : packs[bi]) {
            EnsGroup* g = e->groups[gi];
            for (int it = 0; it < nsteps && !st; ++it) st = ens_step_group<F>(e, g, 0, (int)g->members.size(), it, b);
            if (!st && analysis) st = ens_analysis_group(e, gi, b);
        }
        if (!st && hipEventRecord(e->side_ev[1 + bi], b) != hipSuccess) st = MCS_ERR_HIP;
        if (!st && hipStreamWaitEvent(s, e->side_ev[1 + bi], 0) != hipSuccess) st = MCS_ERR_HIP;
    }
    if (!st) st = ens_advance(e, nsteps, s);
    return st;
}

// The size groups dealt onto nb branches, longest first, each to the branch with the least work
// so far (cost: a group's milliseconds per step launched alone).  A process has few hardware
// queues (4 by default) and branches beyond them share one in order, so with one branch per
// group which groups wait on each other is left to the runtime; packed branches choose it.
static std::vector<std::vector<int>> ens_pack(const std::vector<double>& cost, int nb) {
    std::vector<int> order(cost.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<std::vector<int>> packs(nb);
    std::vector<double> load(nb, 0.0);
    for (int gi : order) {
        const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        packs[b].push_back(gi);
        load[b] += cost[gi];
    }
    packs.erase(std::remove_if(packs.begin(), packs.end(), [](const std::vector<int>& v) { return v.empty(); }),
                packs.end());
    return packs;
}

static constexpr int GRAPH_STEPS = 8;

static int ens_capture(mcs_ensemble* e, hipStream_t s, bool analysis, hipGraph_t* graph, hipGraphExec_t* exec) {
    const int ng = (int)e->groups.size();
    while ((int)e->side.size() < ng) {
        hipStream_t b = nullptr;
        MCS_HIP_TRY(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        e->side.push_back(b);
    }
    while ((int)e->side_ev.size() < ng + 1) {
        hipEvent_t ev = nullptr;
        MCS_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->side_ev.push_back(ev);
    }
    size_t per_slot[4] = {0, 0, 0, 0};
    for (const EnsClass& c : e->cls) ++per_slot[c.slot];
    int cst = ens_class_streams(e, std::max(std::max(per_slot[0], per_slot[1]), std::max(per_slot[2], per_slot[3])));
    if (cst) return cst;
    MCS_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = e->prec == 64 ? ens_branches_t<double>(e, GRAPH_STEPS, s, analysis)
                           : ens_branches_t<float>(e, GRAPH_STEPS, s, analysis);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (st) { if (g) (void)hipGraphDestroy(g); return st; }
    MCS_HIP_TRY(ce);
    MCS_HIP_TRY(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    *graph = g;
    return MCS_OK;
}

static int ens_graph_build(mcs_ensemble* e, hipStream_t s) {
    if (e->gexec) return MCS_OK;
    int st = ens_capture(e, s, false, &e->graph, &e->gexec);
    if (!st) e->gsteps = GRAPH_STEPS;
    return st;
}

// The step graph either replays every size group on its own branch or the plan-class launches
// pass by pass, and eager batched steps launch either per group or per class; every form computes
// the same bits (the same plans), so for each mode the faster is kept: all are timed on the new
// ensemble's buffers (before any field is loaded; the draw counter is reset afterwards).
static int ens_choose_graph(mcs_ensemble* e, hipStream_t s) {
    hipEvent_t a = nullptr, b = nullptr;
    MCS_HIP_TRY(hipEventCreate(&a));
    MCS_HIP_TRY(hipEventCreate(&b));
    double best = -1;
    int st = MCS_OK;
    // each size group's steps launched alone, the cost the branch packing balances
    const int ng = (int)e->groups.size();
    std::vector<double> cost(ng, 1.0);
    for (int gi = 0; gi < ng && !st; ++gi) {
        EnsGroup* g = e->groups[gi];
        const int nm = (int)g->members.size();
        float el = 0;
        st = e->prec == 64 ? ens_step_group<double>(e, g, 0, nm, 0, s) : ens_step_group<float>(e, g, 0, nm, 0, s);
        if (!st && hipEventRecord(a, s) != hipSuccess) st = MCS_ERR_HIP;
        for (int it = 0; it < 4 && !st; ++it)
            st = e->prec == 64 ? ens_step_group<double>(e, g, 0, nm, it, s) : ens_step_group<float>(e, g, 0, nm, it, s);
        if (!st && (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&el, a, b) != hipSuccess))
            st = MCS_ERR_HIP;
        if (!st && el > 0) cost[gi] = el / 4.0;
    }
    e->gcost = cost;
    // the step graph's candidates: one branch per group, the groups packed onto 4, 3 and 2
    // branches, and the class launches
    std::vector<std::vector<std::vector<int>>> forms = {{}};
    for (int nb : {4, 3, 2})
        if (nb < ng) forms.push_back(ens_pack(cost, nb));
    const size_t nforms = forms.size();
    // every candidate captured, then timed in interleaved rounds (3 replays each), keeping each
    // one's fastest round: single timings of one candidate after another varied by about 10%
    std::vector<hipGraph_t> gs;
    std::vector<hipGraphExec_t> xs;
    std::vector<double> ms;
    for (size_t use = 0; use <= nforms && !st; ++use) {
        if (use == nforms && e->cls.empty()) break;
        e->graph_classes = use == nforms;
        if (use < nforms) e->packs = forms[use];
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        st = ens_capture(e, s, false, &g, &x);
        e->graph_classes = false;
        if (st) break;
        gs.push_back(g);
        xs.push_back(x);
        ms.push_back(-1);
        if (hipGraphLaunch(x, s) != hipSuccess) st = MCS_ERR_HIP;
    }
    for (int round = 0; round < 3 && !st; ++round)
        for (size_t i = 0; i < xs.size() && !st; ++i) {
            float el = 0;
            bool ok = hipEventRecord(a, s) == hipSuccess;
            for (int k = 0; k < 3 && ok; ++k) ok = hipGraphLaunch(xs[i], s) == hipSuccess;
            if (ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                hipEventElapsedTime(&el, a, b) == hipSuccess && el > 0) {
                if (ms[i] < 0 || el < ms[i]) ms[i] = el;
            }
        }
    size_t besti = xs.size(), best_pg = 0;
    for (size_t i = 0; i < xs.size(); ++i) {
        if (ms[i] <= 0) continue;
        const int slot = i == nforms ? 3 : 2;
        const double per = ms[i] / (3.0 * GRAPH_STEPS);
        if (e->strat_ms[slot] < 0 || per < e->strat_ms[slot]) e->strat_ms[slot] = per;
        if (besti == xs.size() || ms[i] < ms[besti]) besti = i;
        if (i < nforms && (ms[best_pg] <= 0 || ms[i] < ms[best_pg])) best_pg = i;
    }
    // the per-group form keeps its fastest packing also when the class launches won
    if (!st && !forms.empty()) e->packs = forms[best_pg];
    for (size_t i = 0; i < xs.size(); ++i) {
        if (i == besti) continue;
        (void)hipGraphExecDestroy(xs[i]);
        (void)hipGraphDestroy(gs[i]);
    }
    if (!st && besti < xs.size()) {
        best = ms[besti];
        if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->gexec = xs[besti];
        e->graph = gs[besti];
        e->gsteps = GRAPH_STEPS;
        e->graph_classes = besti == nforms;
    } else if (besti < xs.size()) {
        (void)hipGraphExecDestroy(xs[besti]);
        (void)hipGraphDestroy(gs[besti]);
    }
    // eager batched steps: per-group launches against class launches, the same way
    double best_eager = -1;
    for (int use = 0; use <= 1 && !st; ++use) {
        if (use && e->cls.empty()) break;
        e->batched_classes = use != 0;
        float el = 0;
        if (ens_steps(e, 2, MCS_ENS_BATCHED, s) != MCS_OK || hipEventRecord(a, s) != hipSuccess ||
            ens_steps(e, 6, MCS_ENS_BATCHED, s) != MCS_OK || hipEventRecord(b, s) != hipSuccess ||
            hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&el, a, b) != hipSuccess) {
            (void)hipGetLastError();
            if (use) e->batched_classes = false;
            continue;
        }
        e->strat_ms[use] = el / 6.0;
        if (best_eager < 0 || el < best_eager) best_eager = el;
        else e->batched_classes = false;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!st && best < 0) { mcs_set_last_error(__FILE__, __LINE__, "no step graph ran"); st = MCS_ERR_HIP; }
    if (!st) MCS_HIP_TRY(hipMemsetAsync(e->d_step, 0, sizeof(int64_t), s));
    if (!st) MCS_HIP_TRY(hipStreamSynchronize(s));
    return st;
}

static void ens_analysis_free(mcs_ensemble* e) {
    if (e->aexec) (void)hipGraphExecDestroy(e->aexec);
    if (e->agraph) (void)hipGraphDestroy(e->agraph);
    e->aexec = nullptr;
    e->agraph = nullptr;
    for (void* p : e->a_bufs) (void)hipFree(p);
    e->a_bufs.clear();
    e->a_psi.clear(); e->a_tmp.clear(); e->a_ws.clear(); e->a_idx.clear(); e->a_knn.clear();
    e->a_counts = nullptr;
    e->a_hist = nullptr;
    e->a_k = e->a_nbins = 0;
}

// buffers and the steps+analysis graph for (ths, k, nbins); rebuilt when they change
static int ens_analysis_build(mcs_ensemble* e, double ths, int k, int nbins, hipStream_t s) {
    if (e->aexec && e->a_ths == ths && e->a_k == k && e->a_nbins == nbins) return MCS_OK;
    ens_analysis_free(e);
    auto dalloc = [&](size_t b) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, b ? b : 1) != hipSuccess) return nullptr;
        e->a_bufs.push_back(p);
        return p;
    };
    const size_t fb = e->prec == 64 ? 8 : 4;
    bool ok = true;
    for (EnsGroup* g : e->groups) {
        const mcs_replica& p = e->reps[g->members[0]];
        const size_t n = (size_t)p.ext[0] * p.ext[1] * p.ext[2];
        e->a_psi.push_back(dalloc(n * fb));
        e->a_tmp.push_back(dalloc(n * fb));
        e->a_ws.push_back(dalloc(peaks_workspace_bytes(p.ext)));
        e->a_idx.push_back((int64_t*)dalloc(peak_cap(p) * sizeof(int64_t)));
        e->a_knn.push_back(dalloc(knn_workspace_bytes(peak_cap(p))));
        ok = ok && e->a_psi.back() && e->a_tmp.back() && e->a_ws.back() && e->a_idx.back() && e->a_knn.back();
    }
    e->a_counts = (int64_t*)dalloc(e->reps.size() * sizeof(int64_t));
    e->a_hist = (unsigned long long*)dalloc(2 * (size_t)nbins * e->reps.size() * sizeof(unsigned long long));
    if (!ok || !e->a_counts || !e->a_hist) {
        ens_analysis_free(e);
        mcs_set_last_error(__FILE__, __LINE__, "analysis allocation failed");
        return MCS_ERR_NOMEM;
    }
    e->a_ths = ths;
    e->a_k = k;
    e->a_nbins = nbins;
    int st = ens_capture(e, s, true, &e->agraph, &e->aexec);
    if (st) ens_analysis_free(e);
    return st;
}

// ------------------------------------------------------------------ C ABI
extern "C" int mcs_ensemble_destroy(mcs_ensemble* e) {
    if (!e) return MCS_OK;
    ens_analysis_free(e);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (hipStream_t b : e->side) (void)hipStreamDestroy(b);
    for (hipEvent_t ev : e->side_ev) (void)hipEventDestroy(ev);
    ens_classes_free(e);
    for (hipStream_t b : e->cside) (void)hipStreamDestroy(b);
    for (hipEvent_t ev : e->cev) (void)hipEventDestroy(ev);
    for (EnsGroup* g : e->groups) {
        step_geom_free(&g->sg);
        if (g->d_reps) (void)hipFree(g->d_reps);
        delete g;
    }
    for (void* p : e->U) if (p) (void)hipFree(p);
    for (void* p : e->W) if (p) (void)hipFree(p);
    if (e->d_step) (void)hipFree(e->d_step);
    delete e;
    return MCS_OK;
}

extern "C" int mcs_ensemble_create(mcs_ensemble** out, const mcs_replica* reps, int nrep, int split, int prec) {
    MCS_CHECK(out && reps && nrep > 0, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(prec == 32 || prec == 64, MCS_ERR_INVALID, "precision must be 32 or 64");
    MCS_CHECK(split == MCS_SPLIT_CUBIC || split == MCS_SPLIT_REFERENCE, MCS_ERR_INVALID, "bad split");
    for (int r = 0; r < nrep; ++r) {
        const mcs_replica& p = reps[r];
        MCS_CHECK(p.ext[0] >= 1 && p.ext[1] >= 1 && p.ext[2] >= 1, MCS_ERR_INVALID, "extent must be >=1");
        MCS_CHECK(p.dt > 0 && p.h[0] > 0 && p.h[1] > 0 && p.h[2] > 0, MCS_ERR_INVALID, "dt and h must be positive");
    }
    (void)mcs_stream();
    auto* e = new mcs_ensemble();
    e->reps.assign(reps, reps + nrep);
    e->split = split;
    e->prec = prec;
    e->U.assign(nrep, nullptr);
    e->W.assign(nrep, nullptr);
    e->mass.assign(nrep, 0.0);
    e->group_of.assign(nrep, -1);
    e->slot_of.assign(nrep, -1);
    int rc = MCS_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        e->cus = 256;
    std::map<std::vector<int>, int> by_size;
    for (int r = 0; r < nrep && !rc; ++r) {
        const std::vector<int> key = {reps[r].ext[0], reps[r].ext[1], reps[r].ext[2]};
        auto itg = by_size.find(key);
        if (itg == by_size.end()) {
            auto* g = new EnsGroup();
            e->groups.push_back(g);
            rc = step_geom_build(&g->sg, reps[r].ext, prec);
            itg = by_size.emplace(key, (int)e->groups.size() - 1).first;
        }
        EnsGroup* g = e->groups[itg->second];
        e->group_of[r] = itg->second;
        e->slot_of[r] = (int)g->members.size();
        g->members.push_back(r);
        const size_t bytes = (size_t)reps[r].ext[0] * reps[r].ext[1] * reps[r].ext[2] * (prec == 64 ? 8 : 4);
        if (!rc && (hipMalloc(&e->U[r], bytes) != hipSuccess || hipMalloc(&e->W[r], bytes) != hipSuccess)) {
            mcs_set_last_error(__FILE__, __LINE__, "ensemble state allocation failed");
            rc = MCS_ERR_NOMEM;
        }
    }
    if (!rc && hipMalloc((void**)&e->d_step, sizeof(int64_t)) != hipSuccess) rc = MCS_ERR_NOMEM;
    if (!rc && hipMemset(e->d_step, 0, sizeof(int64_t)) != hipSuccess) rc = MCS_ERR_HIP;
    if (!rc) rc = upload_reps(e);
    for (size_t gi = 0; gi < e->groups.size() && !rc; ++gi)
        rc = e->prec == 64 ? ens_tune_group<double>(e, e->groups[gi], mcs_stream())
                           : ens_tune_group<float>(e, e->groups[gi], mcs_stream());
    if (!rc) rc = e->prec == 64 ? ens_build_classes_t<double>(e) : ens_build_classes_t<float>(e);
    if (!rc) rc = ens_choose_graph(e, mcs_stream());
    if (rc) { mcs_ensemble_destroy(e); return rc; }
    *out = e;
    return MCS_OK;
}

static int ens_load(mcs_ensemble* e, int r, const void* psi) {
    EnsGroup* g = e->groups[e->group_of[r]];
    return step_import(g->sg, phys_of(e->reps[r], e->split, 0.0), psi, e->W[r], e->U[r], &e->mass[r], mcs_stream());
}

extern "C" int mcs_ensemble_init(mcs_ensemble* e) {
    MCS_CHECK(e, MCS_ERR_INVALID, "null argument");
    hipStream_t s = mcs_stream();
    size_t maxb = 0;
    for (const mcs_replica& p : e->reps)
        maxb = std::max(maxb, (size_t)p.ext[0] * p.ext[1] * p.ext[2] * (e->prec == 64 ? 8 : 4));
    void* tmp = nullptr;
    MCS_HIP_TRY(hipMalloc(&tmp, maxb));
    int rc = MCS_OK;
    for (size_t r = 0; r < e->reps.size() && !rc; ++r) {
        const mcs_replica& p = e->reps[r];
        rc = mcs_init_field(tmp, p.ext, p.psi0, p.amp, p.init_seed, e->prec);
        if (!rc) rc = ens_load(e, (int)r, tmp);
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MCS_ERR_HIP;
    (void)hipFree(tmp);
    if (!rc) rc = upload_reps(e);
    if (!rc) e->loaded = true;
    return rc;
}

extern "C" int mcs_ensemble_set_psi(mcs_ensemble* e, int r, const void* psi) {
    MCS_CHECK(e && psi && r >= 0 && r < (int)e->reps.size(), MCS_ERR_INVALID, "bad argument");
    int rc = ens_load(e, r, psi);
    if (!rc) MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    if (!rc) rc = upload_reps(e);
    if (!rc) e->loaded = true;
    return rc;
}

extern "C" int mcs_ensemble_get_psi(mcs_ensemble* e, int r, void* psi) {
    MCS_CHECK(e && psi && r >= 0 && r < (int)e->reps.size(), MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(e->loaded, MCS_ERR_INVALID, "ensemble has no fields (call mcs_ensemble_init)");
    const mcs_replica& p = e->reps[r];
    void* tmp = nullptr;
    MCS_HIP_TRY(hipMalloc(&tmp, (size_t)p.ext[0] * p.ext[1] * p.ext[2] * (e->prec == 64 ? 8 : 4)));
    int rc = step_export(e->groups[e->group_of[r]]->sg, e->U[r], psi, tmp, mcs_stream());
    if (!rc && hipStreamSynchronize(mcs_stream()) != hipSuccess) rc = MCS_ERR_HIP;
    (void)hipFree(tmp);
    return rc;
}

extern "C" int mcs_ensemble_step_async(mcs_ensemble* e, int nsteps, int mode) {
    MCS_CHECK(e && nsteps >= 0, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(mode == MCS_ENS_LOOP || mode == MCS_ENS_BATCHED || mode == MCS_ENS_GRAPH, MCS_ERR_INVALID, "bad mode");
    MCS_CHECK(e->loaded, MCS_ERR_INVALID, "ensemble has no fields (call mcs_ensemble_init)");
    hipStream_t s = mcs_stream();
    int done = 0;
    if (mode == MCS_ENS_GRAPH && nsteps >= GRAPH_STEPS) {
        int rc = ens_graph_build(e, s);
        if (rc) return rc;
        for (; done + e->gsteps <= nsteps; done += e->gsteps) MCS_HIP_TRY(hipGraphLaunch(e->gexec, s));
    }
    int rc = ens_steps(e, nsteps - done, mode == MCS_ENS_LOOP ? MCS_ENS_LOOP : MCS_ENS_BATCHED, s);
    if (rc) return rc;
    e->step += nsteps;
    return MCS_OK;
}

extern "C" int mcs_ensemble_launch_info(mcs_ensemble* e, int* groups, int* classes_per_step, int* batched_classes,
                                        int* graph_classes, double ms_per_step[4]) {
    MCS_CHECK(e && groups && classes_per_step && batched_classes && graph_classes && ms_per_step, MCS_ERR_INVALID,
              "null argument");
    *groups = (int)e->groups.size();
    *classes_per_step = (int)e->cls.size();
    *batched_classes = e->batched_classes ? 1 : 0;
    *graph_classes = e->graph_classes ? 1 : 0;
    for (int i = 0; i < 4; ++i) ms_per_step[i] = e->strat_ms[i];
    return MCS_OK;
}

extern "C" int mcs_ensemble_set_launch_form(mcs_ensemble* e, int batched_classes, int graph_classes) {
    MCS_CHECK(e, MCS_ERR_INVALID, "null argument");
    MCS_CHECK(!(batched_classes || graph_classes) || !e->cls.empty(), MCS_ERR_UNSUPPORTED,
              "this ensemble has no plan classes (a line longer than 256 threads)");
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    e->batched_classes = batched_classes != 0;
    if (e->graph_classes != (graph_classes != 0)) {
        e->graph_classes = graph_classes != 0;
        // the graphs are captured again on their next use
        if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->gexec = nullptr;
        e->graph = nullptr;
        if (e->aexec) (void)hipGraphExecDestroy(e->aexec);
        if (e->agraph) (void)hipGraphDestroy(e->agraph);
        e->aexec = nullptr;
        e->agraph = nullptr;
    }
    return MCS_OK;
}

extern "C" int mcs_ensemble_graph_branches(mcs_ensemble* e, int* branches) {
    MCS_CHECK(e && branches, MCS_ERR_INVALID, "null argument");
    *branches = e->packs.empty() ? (int)e->groups.size() : (int)e->packs.size();
    return MCS_OK;
}

extern "C" int mcs_ensemble_set_graph_branches(mcs_ensemble* e, int branches) {
    MCS_CHECK(e && branches >= 0, MCS_ERR_INVALID, "branches must be >= 0");
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    const int ng = (int)e->groups.size();
    std::vector<double> cost = e->gcost;
    cost.resize(ng, 1.0);
    e->packs = (branches == 0 || branches >= ng) ? std::vector<std::vector<int>>{} : ens_pack(cost, branches);
    // the graphs are captured again on their next use
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    if (e->aexec) (void)hipGraphExecDestroy(e->aexec);
    if (e->agraph) (void)hipGraphDestroy(e->agraph);
    e->aexec = nullptr;
    e->agraph = nullptr;
    return MCS_OK;
}

extern "C" int mcs_ensemble_step(mcs_ensemble* e, int nsteps, int mode) {
    int rc = mcs_ensemble_step_async(e, nsteps, mode);
    if (rc) return rc;
    MCS_HIP_TRY(hipStreamSynchronize(mcs_stream()));
    return MCS_OK;
}

extern "C" int mcs_ensemble_run(mcs_ensemble* e, int nsteps, double ths, int k, int nbins, int64_t* counts,
                                int64_t* hist_q4, int64_t* hist_q6) {
    MCS_CHECK(e && counts && hist_q4 && hist_q6 && nbins > 0 && k >= 1 && k <= 32, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(nsteps >= GRAPH_STEPS, MCS_ERR_INVALID, "mcs_ensemble_run takes at least 8 steps");
    MCS_CHECK(e->loaded, MCS_ERR_INVALID, "ensemble has no fields (call mcs_ensemble_init)");
    for (const mcs_replica& p : e->reps)
        MCS_CHECK(p.ext[2] > 1, MCS_ERR_UNSUPPORTED, "bond-order analysis needs 3D replicas");
    hipStream_t s = mcs_stream();
    int rc = ens_graph_build(e, s);
    if (!rc) rc = ens_analysis_build(e, ths, k, nbins, s);
    if (rc) return rc;
    // the remainder eagerly first, then graph replays; the last replay carries the analysis
    const int rem = nsteps % GRAPH_STEPS;
    if (rem) rc = ens_steps(e, rem, MCS_ENS_BATCHED, s);
    for (int i = 0; !rc && i < nsteps / GRAPH_STEPS - 1; ++i)
        if (hipGraphLaunch(e->gexec, s) != hipSuccess) rc = MCS_ERR_HIP;
    if (!rc && hipGraphLaunch(e->aexec, s) != hipSuccess) rc = MCS_ERR_HIP;
    const int nrep = (int)e->reps.size();
    std::vector<unsigned long long> hh(2 * (size_t)nbins * nrep);
    if (!rc && hipMemcpyAsync(hh.data(), e->a_hist, hh.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              s) != hipSuccess)
        rc = MCS_ERR_HIP;
    if (!rc && hipMemcpyAsync(counts, e->a_counts, nrep * sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = MCS_ERR_HIP;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MCS_ERR_HIP;
    if (rc) {
        mcs_set_last_error(__FILE__, __LINE__, "HIP failure in mcs_ensemble_run");
        return rc;
    }
    e->step += nsteps;
    for (int r = 0; r < nrep; ++r)
        for (int b = 0; b < nbins; ++b) {
            hist_q4[(long)r * nbins + b] = (int64_t)hh[2L * nbins * r + b];
            hist_q6[(long)r * nbins + b] = (int64_t)hh[2L * nbins * r + nbins + b];
        }
    return MCS_OK;
}

extern "C" int mcs_ensemble_analyze(mcs_ensemble* e, double ths, int k, int nbins, int64_t* counts, int64_t* hist_q4,
                                    int64_t* hist_q6) {
    MCS_CHECK(e && counts && hist_q4 && hist_q6 && nbins > 0 && k >= 1 && k <= 32, MCS_ERR_INVALID, "bad argument");
    MCS_CHECK(e->loaded, MCS_ERR_INVALID, "ensemble has no fields (call mcs_ensemble_init)");
    hipStream_t s = mcs_stream();
    const int nrep = (int)e->reps.size();
    std::vector<void*> bufs;
    auto dalloc = [&](size_t b) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, b ? b : 1) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    int rc = MCS_OK;
    unsigned long long* dh = (unsigned long long*)dalloc(sizeof(unsigned long long) * 2 * nbins * nrep);
    int64_t* dcnt = (int64_t*)dalloc(sizeof(int64_t) * nrep);
    if (!dh || !dcnt) rc = MCS_ERR_NOMEM;
    if (!rc) MCS_HIP_TRY(hipMemsetAsync(dh, 0, sizeof(unsigned long long) * 2 * nbins * nrep, s));
    for (int r = 0; r < nrep && !rc; ++r) {
        const mcs_replica& p = e->reps[r];
        MCS_CHECK(p.ext[2] > 1, MCS_ERR_UNSUPPORTED, "bond-order analysis needs 3D replicas");
        const long n = (long)p.ext[0] * p.ext[1] * p.ext[2];
        const int64_t cap = n / 8 + 8;
        void* psi = dalloc(n * (e->prec == 64 ? 8 : 4));
        void* tmp = dalloc(n * (e->prec == 64 ? 8 : 4));
        int64_t* idx = (int64_t*)dalloc(cap * sizeof(int64_t));
        void* kws = dalloc(knn_workspace_bytes(cap));
        void* ws = dalloc(peaks_workspace_bytes(p.ext));
        if (!psi || !tmp || !idx || !kws || !ws) { rc = MCS_ERR_NOMEM; break; }
        rc = step_export(e->groups[e->group_of[r]]->sg, e->U[r], psi, tmp, s);
        if (!rc) rc = findpeaks_launch(psi, p.ext, ths, idx, nullptr, cap, dcnt + r, ws, e->prec, s);
        if (!rc) rc = knn_launch(idx, nullptr, dcnt + r, 0, cap, p.ext, p.h[0], k, kws, nullptr, nullptr,
                                 dh + 2L * nbins * r, dh + 2L * nbins * r + nbins, nbins, s);
    }
    std::vector<unsigned long long> hh(2 * (size_t)nbins * nrep);
    if (!rc && hipMemcpyAsync(hh.data(), dh, hh.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = MCS_ERR_HIP;
    if (!rc && hipMemcpyAsync(counts, dcnt, nrep * sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = MCS_ERR_HIP;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MCS_ERR_HIP;
    for (void* p : bufs) (void)hipFree(p);
    if (rc == MCS_ERR_NOMEM) mcs_set_last_error(__FILE__, __LINE__, "analysis allocation failed");
    if (rc) return rc;
    for (int r = 0; r < nrep; ++r)
        for (int b = 0; b < nbins; ++b) {
            hist_q4[(long)r * nbins + b] = (int64_t)hh[2L * nbins * r + b];
            hist_q6[(long)r * nbins + b] = (int64_t)hh[2L * nbins * r + nbins + b];
        }
    return MCS_OK;
}
