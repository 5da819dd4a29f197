#!/usr/bin/env python3
"""Benchmark of the fused PFC time step on one MI355X.

Workload: 3D PFC 512^3 FP32 (eps=0.3, psi0=-0.25 + U[-0.05,0.05] from Philox seed
11, dx=pi/4, dt=0.5, noise off), the configuration the cell-updates/s metric
is quoted on.  Prints one JSON line:
  value          cell-updates/s = 512^3 * steps / t (HIP events around the step loop)
  roofline       the dominant pass (x_update) against the 8 TB/s HBM peak
  fp64           the same kernels built with F = double (the all-FP64 form)
  cpu_baseline   the FP64 CPU oracle on the host cores, a bounded sample of steps
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s, 3D 512^3 FP32 vs FP64-GPU at stated L2 error; % HBM roofline"


def load_pkg():
    spec = importlib.util.spec_from_file_location("mcstar_libs_amd", os.path.join(REPO, "mcstar-libs_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mcstar_libs_amd"] = mod
    spec.loader.exec_module(mod)
    mod.lib()
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_stepper(mcs, n, dtype, cfg, plan=None, form=None):
    shp = (n, n, n)
    buf = mcs.DeviceBuffer(n ** 3 * np.dtype(dtype).itemsize)
    prec = 64 if dtype == np.float64 else 32
    mcs.check(mcs.lib().mcs_init_field(buf.ptr, mcs._ext(shp), cfg["psi0"], cfg["amp"], cfg["seed"], prec))
    st = mcs.Stepper(shp, [cfg["dx"]] * 3, cfg["eps"], cfg["dt"], temp=0.0, seed=0, dtype=dtype, plan=plan,
                     form=form)
    st.set_psi_device(buf)
    return st, buf


def x_sweeps(mcs, st):
    """sweeps of the state-sized arrays in the x pass: read W and the state, write the state (mixed form);
    read and write both (spectral form)"""
    return 3 if st.form() == mcs.FORM_MIXED else 4


def form_name(mcs, st):
    return "mixed" if st.form() == mcs.FORM_MIXED else "spectral"


def timed_steps(mcs, st, k):
    mcs.sync()
    with mcs.Timer() as t:
        st.step_async(k)
    return t.ms


def cpu_baseline(n, cfg, steps):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    oracle.lib()
    psi = oracle.init_field((n, n, n), cfg["psi0"], cfg["amp"], cfg["seed"])
    h = [cfg["dx"]] * 3
    t0 = time.perf_counter()
    oracle.pfc_run(psi, h, cfg["eps"], cfg["dt"], 0.0, 0, 0, 0, 0)  # transforms in and out only
    t1 = time.perf_counter()
    oracle.pfc_run(psi, h, cfg["eps"], cfg["dt"], 0.0, 0, 0, steps, 0)
    t2 = time.perf_counter()
    per_step = max((t2 - t1) - (t1 - t0), 1e-9) / steps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n ** 3 / per_step, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} steps of the FP64 oracle (own mixed-radix FFT, OpenMP) at {n}^3, "
                      f"minus a 0-step run for the transforms in and out; {per_step:.2f} s/step"}


def measured_traffic(pass_name, n):
    """HBM bytes per launch of a pass from the committed rocprofv3 --pmc summary (512^3 FP32 only)"""
    if n != 512:
        return None
    path = os.path.join(REPO, "profiles", "pfc_step_512cubed_fp32_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    col = None
    for line in open(path):
        cols = [c.strip() for c in line.split("|")]
        if "HBM bytes (scaled)" in cols:
            col = cols.index("HBM bytes (scaled)")
        elif col is not None and len(cols) > col and cols[1] == pass_name:
            return float(cols[col].split()[0]) * 1e9
    return None


def ensemble_bench(mcs, kind, steps, pin=None):
    """replica-steps/s of the batched, graph-replayed ensemble against looping single-replica launches;
    pin (profiling): the step graph replays per-group launches on that many packed branches (0: one
    per size group) instead of the form the ensemble timed fastest at creation"""
    reps = mcs.ensemble_config(kind)
    e = mcs.Ensemble(reps, dtype=np.float32)
    if pin is not None:
        e.set_launch_form(e.launch_info()["batched_uses_classes"], False)
        e.set_graph_branches(pin)
    e.init()
    e.step(8, mcs.ENS_GRAPH)
    cells = sum(int(np.prod(r["shape"])) for r in reps)
    out = {"workload": ("64 replicas of {128,192,320,384,640,768,1024}^2, T=5e-4, dt=0.5" if kind == "2d" else
                        "32 replicas of {64,96,128,160,192}^3, T=1e-3, dt=0.25") + ", FP32, seeded sizes",
           "replicas": len(reps), "cells": cells, "steps": steps}
    for name, mode in (("loop", mcs.ENS_LOOP), ("batched", mcs.ENS_BATCHED), ("graph", mcs.ENS_GRAPH)):
        # each mode warmed by 8 untimed steps of its own: the GPU's clocks follow the load, and the
        # first replays after lighter work (the loop's launch-bound steps) ran up to 8% slower
        e.step(8, mode)
        mcs.sync()
        with mcs.Timer() as t:
            e.step_async(steps, mode)
        out[name] = {"ms_per_step": t.ms / steps, "replica_steps_per_s": len(reps) * steps / (t.ms / 1e3),
                     "cell_updates_per_s": cells * steps / (t.ms / 1e3)}
    out["speedup_graph_over_loop"] = out["graph"]["replica_steps_per_s"] / out["loop"]["replica_steps_per_s"]
    out["launches"] = e.launch_info()
    # algorithmic HBM bytes per step: 40 B per cell in 3D (10 sweeps of 4 B), 24 B in 2D (6 sweeps)
    bpc = 40 if kind == "3d" else 24
    out["graph"]["algorithmic_GB_per_s"] = round(cells * bpc / (out["graph"]["ms_per_step"] / 1e3) / 1e9, 1)
    if kind == "3d":
        # mcs_ensemble_run: the graph-replayed steps with every replica's peaks (all local maxima) and
        # 12-NN q4/q6 64-bin histograms captured after the last step; wall clock including the downloads
        e.run(8, -1.0)  # captures the steps-plus-analysis graph
        mcs.sync()
        t0 = time.perf_counter()
        counts, _, _ = e.run(steps, -1.0)
        ms = 1e3 * (time.perf_counter() - t0)
        out["run_with_analysis"] = {"steps": steps, "ms": round(ms, 3), "peaks": int(counts.sum()),
                                    "analysis_ms": round(ms - out["graph"]["ms_per_step"] * steps, 3)}
    return out


def noise_config_bench(mcs, steps):
    """3D 256^3 with conserved k-space noise (eps=0.3, dt=0.25, psi0=-0.2+U[-0.1,0.1] seed 7, T=1e-3):
    ms per step and algorithmic GB/s of every pass, FP32 and FP64"""
    n, cfg = 256, dict(eps=0.3, psi0=-0.2, amp=0.1, seed=7, dx=np.pi / 4, dt=0.25)
    out = {"workload": "3D PFC 256^3, eps=0.3, dt=0.25, psi0=-0.2+U[-0.1,0.1] seed 7, conserved noise T=1e-3",
           "steps": steps, "note": "the state (2 x 67 MB in FP32) fits the 256 MB last-level cache: effective GB/s"}
    for dtype, name in ((np.float32, "fp32"), (np.float64, "fp64")):
        shp = (n, n, n)
        buf = mcs.DeviceBuffer(n ** 3 * np.dtype(dtype).itemsize)
        prec = 64 if dtype == np.float64 else 32
        mcs.check(mcs.lib().mcs_init_field(buf.ptr, mcs._ext(shp), cfg["psi0"], cfg["amp"], cfg["seed"], prec))
        st = mcs.Stepper(shp, [cfg["dx"]] * 3, cfg["eps"], cfg["dt"], temp=1e-3, seed=7, dtype=dtype)
        st.set_psi_device(buf)
        st.step(10)
        ms = timed_steps(mcs, st, steps) / steps
        passes = st.pass_times(10)
        sweep = (n ** 3 // 2) * 2 * np.dtype(dtype).itemsize
        nbytes = {"x_update": x_sweeps(mcs, st) * sweep, "y_inverse": 2 * sweep, "z_nonlinear": 2 * sweep,
                  "y_forward": 2 * sweep}
        out[name] = {"ms_per_step": ms, "cell_updates_per_s": n ** 3 / (ms / 1e3), "plans": st.plan_info(),
                     "pass_GB_per_s": {k: round(nbytes[k] / (v / 1e3) / 1e9, 1) for k, v in passes.items()},
                     "pass_ms": {k: round(v, 5) for k, v in passes.items()}}
        del st, buf
    return out


def analysis_bench(mcs, st, field, h):
    """local maxima above the mean of the stepper's n^3 field (device-resident: the stepper's export
    buffer), then 12-NN q4/q6/q8/w4 and 64-bin histograms of every peak"""
    import ctypes
    n = field.shape[0]
    thr = float(field.mean())
    cap = field.size // 8 + 8
    idx = np.empty(cap, dtype=np.int64)
    cnt = ctypes.c_int64()
    mcs.check(mcs.lib().mcs_stepper_get_psi(st.ptr, st.buf.ptr), "mcs_stepper_get_psi")
    mcs.sync()

    def peaks():
        mcs.check(mcs.lib().mcs_findpeaks(st.buf.ptr, mcs._ext(field.shape), thr, idx.ctypes.data_as(mcs._i64p),
                                          None, cap, ctypes.byref(cnt), 32), "mcs_findpeaks")

    tf = time.perf_counter()
    peaks()  # the first call sizes the library's workspace
    t0 = time.perf_counter()
    peaks()
    t1 = time.perf_counter()
    pk = idx[:cnt.value]
    ijk = np.column_stack([pk % n, (pk // n) % n, pk // (n * n)]).astype(np.int32)
    t2 = time.perf_counter()
    nbr, q, h4, h6 = mcs.bond_order_knn(ijk, n, h, 12)
    t3 = time.perf_counter()
    # the reference's form: neighbours within rlim of the peak positions i*h in a box of n*h; rlim 1.05 a for
    # a bcc lattice of the peaks' density (first and second shells, 14 neighbours)
    xyz = ijk.astype(np.float64) * h
    box = n * h
    rlim = 1.05 * (2.0 * box ** 3 / max(pk.size, 1)) ** (1.0 / 3.0)
    t4 = time.perf_counter()
    nb, out9 = mcs.bond_order_cutoff(xyz, box, rlim)
    t5 = time.perf_counter()
    return {"workload": f"peaks >= mean of the {n}^3 FP32 field after 500 steps, then 12 nearest neighbours, "
                        f"q4 q6 q8 w4 and 64-bin q4/q6 histograms of every peak (cell list, FP64)",
            "peaks": int(pk.size), "findpeaks_ms": round(1e3 * (t1 - t0), 2),
            "findpeaks_first_call_ms": round(1e3 * (t0 - tf), 2), "bond_order_ms": round(1e3 * (t3 - t2), 2),
            "timing": "wall clock of each call: the field is on the device and the peak indices are downloaded; "
                      "the first peak call also allocates the library's kept workspace; the bond order uploads the "
                      "peak coordinates and downloads neighbours, q values and histograms",
            "q4_mean": float(q[:, 0].mean()), "q6_mean": float(q[:, 1].mean()),
            "q6_mode_bin": int(np.argmax(h6)),
            "cutoff_form": {"rlim": round(rlim, 4), "bond_order_ms": round(1e3 * (t5 - t4), 2),
                            "mean_neighbours": float(nb.mean()),
                            "note": "q4 q6 q8 w4, neighbour-averaged forms and xi of every peak (the reference's "
                                    "bond_order_analysis), cell list"}}


def fft_bench(mcs, n, reps, shp=None):
    """boundary FFT (the FFT::CufftPlan replacement) R2C and C2R on an n^3 (or shp) FP32 field"""
    shp = tuple(shp) if shp else (n, n, n)
    nb = mcs.buffer_bytes(shp, np.float32)
    x = mcs.DeviceBuffer(nb)
    mcs.check(mcs.lib().mcs_init_field(x.ptr, mcs._ext(shp), 0.0, 1.0, 3, 32), "mcs_init_field")
    k, y = mcs.DeviceBuffer(nb), mcs.DeviceBuffer(nb)
    fwd, inv = mcs.FFTPlan(shp, np.float32, mcs.R2C), mcs.FFTPlan(shp, np.float32, mcs.C2R)
    fwd.execute(x, k)
    inv.execute(k, y)
    out = {}
    # algorithmic bytes: x pass reads/writes the real field and the half spectrum, y and z read/write the spectrum
    ncell = int(np.prod(shp))
    half = (shp[-1] // 2 + 1) * (ncell // shp[-1]) * 8
    byts = (ncell * 4 + half) + 2 * (len(shp) - 1) * half
    for name, plan, a, b in (("r2c", fwd, x, k), ("c2r", inv, k, y)):
        mcs.sync()
        with mcs.Timer() as t:
            for _ in range(reps):
                mcs.check(mcs.lib().mcs_fft_exec(plan.ptr, a.ptr, b.ptr), "mcs_fft_exec")
        ms = t.ms / reps
        out[name] = {"ms": round(ms, 4), "GB_per_s": round(byts / (ms / 1e3) / 1e9, 1)}
    out["bytes_per_transform"] = byts
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-noise-config", action="store_true", help="skip the 256^3 noise-on line")
    ap.add_argument("--only-ensemble", action="store_true", help="run only the ensemble lines (profiling)")
    ap.add_argument("--only-fft", action="store_true", help="run only the boundary FFT lines (profiling)")
    ap.add_argument("--ens-pin", default=None,
                    help="with --only-ensemble (profiling): pin each step graph to per-group launches on N packed "
                         "branches, e.g. 2d=0,3d=2 (0: one branch per size group)")
    ap.add_argument("--only-analysis", action="store_true",
                    help="run only the 500-step FP32 field and its peak / bond-order analysis (profiling)")
    ap.add_argument("--no-other-form", action="store_true", help="skip the other state form's comparison line")
    ap.add_argument("--form", choices=["mixed", "spectral"], default=None,
                    help="state form of the timed steppers (default: the library's default, mixed)")
    ap.add_argument("--no-fft", action="store_true",
                    help="skip the boundary FFT line (profiling the step alone: its column passes share kernels)")
    ap.add_argument("--plan", default=None,
                    help="explicit FP32 plans Px,Gx,Py,Gy,Pz,Gz instead of the tuned ones (profiling without "
                         "the tuner's trial launches)")
    ap.add_argument("--fixed", type=int, default=None,
                    help="fixed-plan kernel mask (1 column passes, 2 x pass; default both)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference dispatch for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the PFC step runs on a single "
              f"MI355X (sharding the grid in z slabs with an all-to-all transpose per step is not built)",
              file=sys.stderr)
        sys.exit(2)

    mcs = load_pkg()
    if a.form is not None:
        mcs.set_step_form(mcs.FORM_MIXED if a.form == "mixed" else mcs.FORM_SPECTRAL)
    if a.fixed is not None:
        mcs.set_fixed_kernels(a.fixed)
    if a.copy_calibration:
        nb = 1 << 30
        src, dst = mcs.DeviceBuffer(nb).zero(), mcs.DeviceBuffer(nb)
        mcs.check(mcs.lib().mcs_memcpy(dst.ptr, src.ptr, nb, mcs.D2D), "mcs_memcpy")
        del src, dst
    if a.only_ensemble:
        pins = dict(kv.split("=") for kv in a.ens_pin.split(",")) if a.ens_pin else {}
        pin = {k: int(v) for k, v in pins.items()}
        print(json.dumps({"ensemble_2d": ensemble_bench(mcs, "2d", 40, pin.get("2d")),
                          "ensemble_3d": ensemble_bench(mcs, "3d", 32, pin.get("3d"))}))
        return
    if a.only_fft:
        print(json.dumps({"boundary_fft_512": fft_bench(mcs, a.n, 5),
                          "boundary_fft_1200sq": fft_bench(mcs, 1200, 20, shp=(1200, 1200))}))
        return
    n = a.n
    cfg = dict(eps=0.3, psi0=-0.25, amp=0.05, seed=11, dx=np.pi / 4, dt=0.5)
    N = n ** 3
    if a.only_analysis:
        st, _ = make_stepper(mcs, n, np.float32, cfg)
        st.step(500)
        f32 = st.get_psi()
        runs = [analysis_bench(mcs, st, f32, cfg["dx"]) for _ in range(3)]
        print(json.dumps({"analysis": runs[-1], "findpeaks_ms_runs": [r["findpeaks_ms"] for r in runs]}))
        return
    log(f"[bench] fp32 stepper {n}^3: warmup {a.warmup}")
    plan = [int(v) for v in a.plan.split(",")] if a.plan else None
    st, _ = make_stepper(mcs, n, np.float32, cfg, plan)
    st.step(a.warmup)
    log(f"[bench] fp32 timed {a.steps} steps")
    ms = timed_steps(mcs, st, a.steps)
    ms_step = ms / a.steps
    value = N * a.steps / (ms / 1e3)
    passes = st.pass_times(10)
    log(f"[bench] fp32 {ms_step:.4f} ms/step, passes {passes}")
    steps32 = a.warmup + a.steps + 10

    # algorithmic bytes per pass: every pass streams an array of the state's size (N/2 complex = 4N bytes in
    # FP32) once in and once out; x_update also reads the state (mixed form: 3 sweeps) or reads and writes it
    # (spectral form: 4 sweeps)
    sweep = 4 * N
    pass_bytes = {"x_update": x_sweeps(mcs, st) * sweep, "y_inverse": 2 * sweep, "z_nonlinear": 2 * sweep,
                  "y_forward": 2 * sweep}
    dom = max(passes, key=passes.get)
    achieved = pass_bytes[dom] / (passes[dom] / 1e3) / 1e9
    traffic = measured_traffic(dom, n)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/pfc_step_512cubed_fp32_hbm_traffic.md (2 x FETCH_SIZE + WRITE_SIZE)"
                if traffic else None,
                "bytes_per_launch": pass_bytes[dom], "launch_ms": round(passes[dom], 5)}
    # step-level model: 11 Cb + 2 Rb bytes per 3D step (half-spectrum FFT passes with fused pointwise work)
    cb = (n // 2 + 1) * n * n * 8
    rb = N * 4
    b_step = 11 * cb + 2 * rb
    step_frac = b_step / (ms_step / 1e3) / (HBM_PEAK_GBS * 1e9)
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (Philox-seeded initial field)",
           "config": {"workload": f"3D PFC {n}^3 FP32, eps=0.3, psi0=-0.25+U[-0.05,0.05] seed 11, dx=pi/4, "
                                  f"dt=0.5, noise off", "n": n, "precision": "fp32", "parallelism": "1 GPU"},
           "steps_per_s": 1e3 / ms_step, "hbm_roofline_frac_step": round(step_frac, 4),
           "model_bytes_per_step": b_step, "pass_ms": {k: round(v, 5) for k, v in passes.items()},
           "plans": st.plan_info(), "form": form_name(mcs, st), "sweeps_per_step": x_sweeps(mcs, st) + 6,
           "autotune": mcs.get_autotune(), "fixed_kernels": mcs.get_fixed_kernels(), "roofline": roofline}

    if not a.no_other_form:
        # the same step in the other state form (spectral: 10 sweeps per step; mixed: 9, with the state
        # passing through an x transform and back every step)
        other = mcs.FORM_MIXED if st.form() == mcs.FORM_SPECTRAL else mcs.FORM_SPECTRAL
        sp, _ = make_stepper(mcs, n, np.float32, cfg, plan, form=other)
        log(f"[bench] fp32 {form_name(mcs, sp)} form")
        sp.step(a.warmup)
        ksp = min(a.steps, 100)
        msp = timed_steps(mcs, sp, ksp) / ksp
        out["other_form"] = {"form": form_name(mcs, sp), "ms_per_step": msp, "cell_updates_per_s": N / (msp / 1e3),
                             "pass_ms": {k: round(v, 5) for k, v in sp.pass_times(10).items()},
                             "plans": sp.plan_info(), "sweeps_per_step": x_sweeps(mcs, sp) + 6}
        del sp

    if not a.no_fp64:
        log("[bench] fp64 form")
        st64, _ = make_stepper(mcs, n, np.float64, cfg)
        st64.step(a.warmup)
        k64 = min(a.steps, 50)
        ms64 = timed_steps(mcs, st64, k64)
        # precision study at 500 steps (or more, if the timed run went further)
        n_cmp = max(steps32, 500)
        st.step(n_cmp - steps32)
        st64.step(n_cmp - a.warmup - k64)
        mcs.sync()
        assert st.steps_taken() == st64.steps_taken() == n_cmp, (st.steps_taken(), st64.steps_taken(), n_cmp)
        v64 = N * k64 / (ms64 / 1e3)
        f32 = st.get_psi()
        p32 = f32.astype(np.float64)
        p64 = st64.get_psi()
        rel = float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64))
        out["fp64"] = {"value": v64, "ms_per_step": ms64 / k64, "speedup_fp32_over_fp64": value / v64,
                       "rel_l2_fp32_vs_fp64_gpu": rel, "max_abs_fp32_vs_fp64_gpu": float(np.abs(p32 - p64).max()),
                       "steps_compared": n_cmp}
        del st64, p32, p64
        log("[bench] peaks and 12-NN bond order of the 500-step field")
        out["analysis"] = analysis_bench(mcs, st, f32, cfg["dx"])
        del f32
    if not a.no_noise_config:
        log("[bench] 256^3 with noise")
        out["noise_256"] = noise_config_bench(mcs, 100)
    log("[bench] boundary FFT")
    if not a.no_fft:
        out["boundary_fft"] = fft_bench(mcs, n, 5)
        # a length of the reference's own FFT tests (cufft_unittest/my_unit_test.py:161-163) that the
        # size policy (cuda/math.hpp:139-149) yields: 1200 = 3 5^2 2^4 (30-point column plan)
        out["boundary_fft_1200sq"] = dict(fft_bench(mcs, 1200, 20, shp=(1200, 1200)), shape=[1200, 1200])
    if not a.no_ensemble:
        log("[bench] ensembles")
        out["ensemble_2d"] = ensemble_bench(mcs, "2d", 40)
        out["ensemble_3d"] = ensemble_bench(mcs, "3d", 32)
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(n, cfg, a.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
