"""Register plans and tile shapes stay inside the launch bounds of the kernels that run them
(host-side planner only, no GPU).

The register kernels are compiled with __launch_bounds__ of 256 threads when a thread holds more
than 64 bytes of line data or 16+ floats (MCS_REG_MAX_THREADS in csrc/fft_core.hpp), else 1024; the
x-pass kernels with 256 (MCS_XPASS_MAX_THREADS).  A launch beyond its kernel's bound fails on the
device, so every plan the planner can emit, every default tile, every tuner candidate and every
explicit plan must be checked on the host first.
"""
import ctypes

import numpy as np
import pytest


def reg_bound(P, prec):
    return 256 if (P * (8 if prec == 64 else 4) > 64 or (prec == 32 and P >= 16)) else 1024


def lds_ok(nbytes, G):
    # 64 KB of exchange LDS per tile, or one line (pair) up to the 160 KB a workgroup can take
    return nbytes <= 65536 or (G == 1 and nbytes <= 160 * 1024)


def smooth_lengths(limit=32768):
    out = []
    for a in range(0, 16):
        for b in range(0, 4):
            for c in range(0, 4):
                n = (1 << a) * 3 ** b * 5 ** c
                if 4 <= n <= limit:
                    out.append(n)
    return sorted(out)


@pytest.mark.parametrize("prec", [32, 64])
def test_every_plan_and_tile_within_launch_bound(mcs, prec):
    dt = np.float32 if prec == 32 else np.float64
    esz = 8 if prec == 32 else 16
    nreg = 0
    for L in smooth_lengths():
        opts = mcs.plan_options(L, dt)
        for P in opts:
            assert L % P == 0, (L, P)
            S = L // P
            assert S <= reg_bound(P, prec), (L, P, prec)
            col = mcs.tile_options(L, dt, P, False)
            assert col, (L, P)
            for G in col:
                assert G * S <= reg_bound(P, prec) and lds_ok(G * L * esz, G), (L, P, G)
            for G in mcs.tile_options(L, dt, P, True):
                assert G * S <= 256 and lds_ok(2 * G * L * esz, G), (L, P, G)
            nreg += 1
    assert nreg > 60


@pytest.mark.parametrize("prec,L,P", [(32, 8192, 32), (32, 10240, 10), (32, 12288, 12), (64, 6144, 6), (64, 5120, 20),
                                      (64, 8192, 8)])
def test_long_lines_take_a_plan_within_the_bound(mcs, prec, L, P):
    dt = np.float32 if prec == 32 else np.float64
    assert mcs.plan_options(L, dt)[0] == P


@pytest.mark.parametrize("prec,L", [(32, 16384), (64, 10240), (64, 12288)])
def test_lines_past_every_bound_have_no_register_plan(mcs, prec, L):
    assert mcs.plan_options(L, np.float32 if prec == 32 else np.float64) == []


def test_default_tile_is_the_first_candidate(mcs):
    # defaults that fall outside the tuner's power-of-two tiles: FP32 ny = 768 (12 points, 64
    # threads a line, 10 lines by the 64 KB LDS rule), ny = 2048 (16 points, 128 threads, 2 lines by
    # the 256-thread bound), the 1280-point x pass (nx = 2560, 20 points, 3 pairs by LDS); FP64
    # x pass of nx = 768 (5 pairs) and ny = 1280 (3 lines).  An explicit plan (a checkpoint's) that
    # names them must be accepted as the stepper reports it.
    assert mcs.tile_options(768, np.float32, 12, False)[0] == 10
    assert mcs.tile_options(2048, np.float32, 16, False)[0] == 2
    assert mcs.tile_options(1280, np.float32, 20, True)[0] == 3
    assert mcs.tile_options(384, np.float64, 12, True)[0] == 5
    assert mcs.tile_options(1280, np.float64, 20, False)[0] == 3


def _planned(mcs, ext, prec, plan):
    p = mcs.PfcParams()
    p.ext = (ctypes.c_int * 3)(*ext)
    p.h = (ctypes.c_double * 3)(0.785, 0.785, 0.785)
    p.eps, p.dt, p.temp, p.seed, p.split, p.prec = 0.3, 0.5, 0.0, 0, 0, prec
    st = ctypes.c_void_p()
    return mcs.lib().mcs_stepper_create_planned(ctypes.byref(st), ctypes.byref(p), (ctypes.c_int * 6)(*plan))


@pytest.mark.parametrize("plan,why", [
    ((8, 8, 16, 8, 32, 32), "z tile of 32 lines x 16 threads = 512 threads > 256"),
    ((8, 8, 16, 8, 32, 0), "no lines"),
    ((8, 64, 16, 8, 32, 16), "x tile of 64 pairs x 32 threads"),
    ((8, 8, 12, 8, 32, 16), "12 points per thread on a power-of-two axis"),
    ((4, 8, 16, 8, 32, 16), "x tile of 8 pairs x 64 threads"),
])
def test_explicit_plan_outside_the_bounds_is_rejected_before_launch(mcs, plan, why):
    assert _planned(mcs, (512, 512, 512), 32, plan) == mcs.ERR_UNSUPPORTED, why
    err = mcs.lib().mcs_last_error()
    assert b"tile shape" in err or b"points per thread" in err
