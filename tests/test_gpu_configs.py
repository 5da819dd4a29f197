"""Configuration-level checks of the fused stepper.

- 3D 256^3 with conserved k-space noise (eps=0.3, dt=0.25, psi0=-0.2+U[-0.1,0.1] seed 7, T=1e-3, 2000 steps):
  the noise spectrum against <|eta_k|^2> = 2 T dt k^2 / dV, exact mass conservation, FP32 against the all-FP64
  form.
- The 512^3 precision study (eps=0.3, psi0=-0.25+U[-0.05,0.05] seed 11, dt=0.5, 500 steps, then peaks and
  12-NN q4/q6): at 512^3 against the FP64 oracle's field of the same run (tests/golden/c3_oracle_512.npz,
  written by tests/golden/gen_c3_oracle.py: every 8th cell per axis, a full-resolution 32^3 block and
  whole-field sums), and at 64^3 against a live oracle run; the full-field comparison with a live
  512^3 oracle run is opt-in (MCS_FULL_ORACLE=1, about 15 minutes of CPU time on 16 cores).
"""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = np.pi / 4
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def k2_grid(shp, h):
    """|k|^2 on the half spectrum of a real grid with numpy shape (.., nx)"""
    ks = [2 * np.pi * np.fft.fftfreq(n, d=h) for n in shp[:-1]] + [2 * np.pi * np.fft.rfftfreq(shp[-1], d=h)]
    grids = np.meshgrid(*ks, indexing="ij")
    return sum(g ** 2 for g in grids)


def noise_one_step(mcs, shp, eps, dt, temp, seed, step0, dtype):
    """one step from psi = 0: N = psi^3 = 0, so the new spectrum is eta_k / D_k"""
    st = mcs.Stepper(shp, [H] * 3, eps, dt, temp=temp, seed=seed, dtype=dtype)
    mcs.check(mcs.lib().mcs_stepper_set_step(st.ptr, step0))
    st.set_psi(np.zeros(shp))
    st.step(1)
    return np.fft.rfftn(st.get_psi().astype(np.float64))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noise_spectrum_matches_theory_256cubed(mcs, dtype):
    shp, eps, dt, temp = (256, 256, 256), 0.3, 0.25, 1e-3
    N = np.prod(shp)
    dV = H ** 3
    k2 = k2_grid(shp, H)
    D = 1 + dt * k2 * (-eps + (1 - k2) ** 2)
    spec = noise_one_step(mcs, shp, eps, dt, temp, 7, 0, dtype)
    # k = 0 carries no noise (mass is conserved); what remains is the rounding of the exported field
    assert abs(spec[0, 0, 0]) < 1e-5 * np.sqrt(np.mean(np.abs(spec) ** 2))
    nz = k2 > 0
    # |eta_k|^2 / <|eta_k|^2> per mode: E = 1 and Var = 1 (|xi|^2 is exponential; real modes: chi^2_1, Var 2)
    ratio = (np.abs(spec) ** 2 * D ** 2)[nz] / (2 * temp * dt / dV * N * k2[nz])
    assert abs(ratio.mean() - 1) < 5 / np.sqrt(ratio.size)
    # per |k| shell: the spectrum is proportional to k^2 at every scale
    kk = np.sqrt(k2[nz])
    edges = np.linspace(0, kk.max() + 1e-9, 17)
    which = np.digitize(kk, edges) - 1
    for b in range(16):
        r = ratio[which == b]
        if r.size >= 1000:
            assert abs(r.mean() - 1) < 6 / np.sqrt(r.size), (b, r.size, r.mean())
    # real and imaginary parts carry half the power each, and draws of different steps are independent
    mid = nz.copy()
    mid[:, :, 0] = False
    mid[:, :, -1] = False
    re = (spec.real * D)[mid] / np.sqrt(2 * temp * dt / dV * N * k2[mid])
    im = (spec.imag * D)[mid] / np.sqrt(2 * temp * dt / dV * N * k2[mid])
    assert abs(np.mean(re ** 2) - 0.5) < 0.01 and abs(np.mean(im ** 2) - 0.5) < 0.01
    assert abs(np.mean(re * im)) < 0.005
    other = noise_one_step(mcs, shp, eps, dt, temp, 7, 5, dtype)
    corr = np.vdot(spec[mid], other[mid]).real / np.sqrt(np.vdot(spec[mid], spec[mid]).real *
                                                         np.vdot(other[mid], other[mid]).real)
    assert abs(corr) < 0.01


def test_config_256cubed_noise_2000_steps(mcs, orc):
    shp, eps, dt, temp, seed = (256, 256, 256), 0.3, 0.25, 1e-3, 7
    psi = orc.init_field(shp, -0.2, 0.1, 7)
    sts = {}
    for dtype in (np.float32, np.float64):
        st = mcs.Stepper(shp, [H] * 3, eps, dt, temp=temp, seed=seed, dtype=dtype)
        st.set_psi(psi)
        assert st.mass() == pytest.approx(psi.astype(dtype).astype(np.float64).mean(), abs=1e-15)
        sts[dtype] = st
    # pointwise while the trajectories stay close: both forms draw the same Philox numbers
    for st in sts.values():
        st.step(500)
    a, b = sts[np.float32].get_psi().astype(np.float64), sts[np.float64].get_psi()
    r500 = rel_l2(a, b)
    assert r500 < 1e-4
    for st in sts.values():
        st.step(1500)
    out = {}
    for dtype, st in sts.items():
        m0 = psi.astype(dtype).astype(np.float64).mean()
        assert st.steps_taken() == 2000
        # the k = 0 mode is carried in FP64 and never touched by the update or the noise
        assert abs(st.mass() - m0) < 1e-15
        got = st.get_psi().astype(np.float64)
        assert np.isfinite(got).all()
        assert abs(got.mean() - m0) < (2e-8 if dtype == np.float32 else 1e-14)
        out[dtype] = (got, mcs.field_stats(got, [H] * 3, eps)[1])
    # by 2000 steps coarsening events amplify round-off (rel-L2 grows from 1e-4 at 1000 steps
    # to about 6e-3); the statistics of the two forms still agree
    (a, fa), (b, fb) = out[np.float32], out[np.float64]
    print(f"256^3 T=1e-3: rel-L2 FP32 vs FP64 {r500:.3e} at 500 steps, {rel_l2(a, b):.3e} at 2000; "
          f"variance {abs(a.var() - b.var()) / b.var():.2e}, free energy {abs(fa - fb) / abs(fb):.2e} apart")
    assert abs(a.var() - b.var()) / b.var() < 1e-5
    assert abs(fa - fb) / abs(fb) < 2e-5


def _c3(mcs, orc, n, steps, chunk=None):
    shp = (n, n, n)
    psi = orc.init_field(shp, -0.25, 0.05, 11)
    st = mcs.Stepper(shp, [H] * 3, 0.3, 0.5, dtype=np.float32)
    st.set_psi(psi)
    st.step(steps)
    got = st.get_psi()
    del st
    chunk = chunk or steps
    want = psi
    for s0 in range(0, steps, chunk):  # noise off: chunks compose exactly
        want = orc.pfc_run(want, [H] * 3, 0.3, 0.5, 0.0, 0, s0, min(chunk, steps - s0), 0)
        if chunk < steps:
            print(f"oracle {n}^3: {s0 + min(chunk, steps - s0)}/{steps} steps", file=sys.__stderr__, flush=True)
    r, m = rel_l2(got, want), float(np.abs(got - want).max())
    print(f"{n}^3 FP32 {steps} steps vs FP64 oracle: rel-L2 {r:.3e}, max-abs {m:.3e}", file=sys.__stderr__,
          flush=True)
    return got, want, r, m


def _peaks_and_knn(mcs, orc, field, n):
    thr = float(field.mean())
    pg = mcs.findpeaks(field, thr)
    assert np.array_equal(pg, orc.findpeaks(field, thr))
    ijk = np.column_stack([pg % n, (pg // n) % n, pg // (n * n)]).astype(np.int32)
    return ijk


def test_precision_study_64cubed_500_steps(mcs, orc):
    n = 64
    got, want, r, m = _c3(mcs, orc, n, 500)
    assert r < 1e-4 and m < 1e-3
    ijk = _peaks_and_knn(mcs, orc, got, n)
    assert ijk.shape[0] > 12
    nbr, q, h4, h6 = mcs.bond_order_knn(ijk, n, H, 12)
    onbr, oq = orc.bond_order_knn(ijk, n, H, 12)
    assert np.array_equal(nbr, onbr)
    assert np.abs(q - oq).max() < 1e-12
    assert h4.sum() == ijk.shape[0] and h6.sum() == ijk.shape[0]


@pytest.mark.parametrize("dtype,form", [(np.float32, 0), (np.float64, 0), (np.float32, 1)])
def test_precision_study_512cubed_500_steps_against_oracle_fixture(mcs, dtype, form):
    """the headline precision gate: FP32 (FFT and pointwise in FP32, twiddles from FP64, k = 0 in
    FP64) within rel-L2 1e-4 and max-abs 5e-3 of the FP64 oracle after 500 steps at 512^3; the
    all-FP64 GPU form within 1e-10 (cuda/kernels/eom.hpp:60-72 is the update both run).  form 0 is
    the stepper's default (spectral) state; the mixed form (form 1: the state passes through an x
    transform and back every step) is held to its own stated bound, rel-L2 2e-4"""
    g = np.load(os.path.join(GOLD, "c3_oracle_512.npz"))
    n, eps, dt, psi0, amp, seed, steps, _ = g["params"]
    n, seed, steps = int(n), int(seed), int(steps)
    stride, b0 = int(g["stride"]), int(g["block_origin"])
    bw = g["block"].shape[0]
    psi = mcs.init_field((n, n, n), psi0, amp, seed, dtype=np.float64)
    st = mcs.Stepper((n, n, n), [H] * 3, eps, dt, dtype=dtype, form=form)
    st.set_psi(psi)
    del psi
    st.step(steps)
    assert st.steps_taken() == steps
    got = st.get_psi().astype(np.float64)
    del st
    samp = got[::stride, ::stride, ::stride]
    blk = got[b0:b0 + bw, b0:b0 + bw, b0:b0 + bw]
    r_s, m_s = rel_l2(samp, g["sample"]), float(np.abs(samp - g["sample"]).max())
    r_b, m_b = rel_l2(blk, g["block"]), float(np.abs(blk - g["block"]).max())
    sumsq = float(np.sum(got * got))
    r_q = abs(sumsq - float(g["sumsq"])) / float(g["sumsq"])
    mean = float(got.mean())
    print(f"512^3 {np.dtype(dtype).name} form {form} {steps} steps vs FP64 oracle: sampled rel-L2 {r_s:.3e} max-abs {m_s:.3e}, "
          f"block rel-L2 {r_b:.3e} max-abs {m_b:.3e}, sum psi^2 {r_q:.2e} apart, max {got.max():.6f} "
          f"({float(g['vmax']):.6f}), min {got.min():.6f} ({float(g['vmin']):.6f})", file=sys.__stderr__, flush=True)
    if dtype == np.float32:
        # the gate is on the whole field (estimated from the 64^3 sample); the block is a local check
        # of one 32^3 region, where an interface moving through it makes the local error larger
        gate = 1e-4 if form == 0 else 2e-4
        assert r_s <= gate and m_s <= 5e-3 and r_b <= 5e-4 and m_b <= 5e-3
        assert r_q < 1e-5 and abs(got.max() - float(g["vmax"])) < 5e-3 and abs(got.min() - float(g["vmin"])) < 5e-3
        # the mean is carried exactly; what remains is the rounding of the exported FP32 field
        assert abs(mean - float(g["total"]) / n ** 3) < 2e-8
    else:
        assert r_s <= 1e-10 and r_b <= 1e-10 and m_s < 1e-9 and r_q < 1e-12
        assert abs(mean - float(g["total"]) / n ** 3) < 1e-14
    # the oracle's own mean stays at the initial one through every step (k = 0 untouched)
    assert np.abs(g["mean_hist"] - g["mean_hist"][0]).max() < 1e-15


@pytest.mark.skipif(os.environ.get("MCS_FULL_ORACLE") != "1",
                    reason="opt-in: the FP64 oracle needs about 15 minutes for 500 steps at 512^3")
def test_precision_study_512cubed_500_steps_full_oracle(mcs, orc):
    n = 512
    got, want, r, m = _c3(mcs, orc, n, 500, chunk=10)
    assert r < 1e-4
    del want
    pg = mcs.findpeaks(got, float(got.mean()))
    assert np.array_equal(pg, orc.findpeaks(got, float(got.mean())))


def test_precision_study_512cubed_peaks_and_bond_order(mcs, orc):
    # the full-size flow: 500 FP32 steps at 512^3, peaks of the result, 12-NN q4/q6 of every peak on
    # the device; peak list against the oracle's on the same field, neighbour lists and q values of a
    # sample of peaks against the oracle's all-pairs search over every peak
    n = 512
    psi = orc.init_field((n, n, n), -0.25, 0.05, 11)
    st = mcs.Stepper((n, n, n), [H] * 3, 0.3, 0.5, dtype=np.float32)
    st.set_psi(psi)
    del psi
    st.step(500)
    got = st.get_psi()
    del st
    ijk = _peaks_and_knn(mcs, orc, got, n)
    P = ijk.shape[0]
    print(f"512^3 after 500 steps: {P} peaks", file=sys.__stderr__, flush=True)
    assert P > 10000
    t0 = time.perf_counter()
    nbr, q, h4, h6 = mcs.bond_order_knn(ijk, n, H, 12)
    print(f"12-NN bond order of {P} peaks: {1e3 * (time.perf_counter() - t0):.1f} ms including transfers",
          file=sys.__stderr__, flush=True)
    assert (nbr >= 0).all() and np.isfinite(q).all()
    sample = np.random.default_rng(0).choice(P, 1500, replace=False)
    onbr, oq = orc.bond_order_knn_points(ijk, n, H, sample, 12)
    assert np.array_equal(nbr[sample], onbr)
    assert np.abs(q[sample] - oq).max() < 1e-12
    assert np.array_equal(h4, np.bincount(np.minimum((q[:, 0] * 64).astype(int), 63), minlength=64))
    assert np.array_equal(h6, np.bincount(np.minimum((q[:, 1] * 64).astype(int), 63), minlength=64))
