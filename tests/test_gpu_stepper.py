"""GPU parity of the fused PFC stepper against the FP64 oracle on seeded inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = np.pi / 4


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


FORMS = [0, 1]  # FORM_SPECTRAL, FORM_MIXED


def run_pair(mcs, orc, shp, eps, dt, psi0, amp, seed, steps, split=0, temp=0.0, noise_seed=0, dtype=np.float64,
             form=None):
    psi = orc.init_field(shp, psi0, amp, seed)
    h3 = [H, H, H]
    st = mcs.Stepper(shp, h3, eps, dt, temp=temp, seed=noise_seed, split=split, dtype=dtype, form=form)
    st.set_psi(psi)
    st.step(steps)
    got = st.get_psi()
    want = orc.pfc_run(psi, h3, eps, dt, temp, noise_seed, 0, steps, split)
    return got, want, psi, st


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("shp", [(32, 32, 32), (24, 20, 48), (64, 64), (48, 80)])
def test_fp64_matches_oracle(mcs, orc, shp, split, form):
    got, want, psi, st = run_pair(mcs, orc, shp, 0.3, 0.5, -0.25, 0.05, 11, 20, split=split, form=form)
    assert st.form() == form
    assert rel_l2(got, want) < 1e-12
    assert abs(got.mean() - psi.mean()) < 1e-14


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("shp", [(32, 32, 32), (64, 96)])
def test_fp64_with_noise_matches_oracle(mcs, orc, shp, form):
    got, want, psi, st = run_pair(mcs, orc, shp, 0.3, 0.25, -0.2, 0.1, 7, 15, temp=1e-3, noise_seed=7, form=form)
    assert rel_l2(got, want) < 1e-11
    assert abs(got.mean() - psi.mean()) < 1e-14


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("shp", [(32, 32, 32), (128, 128)])
def test_fp32_matches_oracle(mcs, orc, shp, form):
    got, want, psi, st = run_pair(mcs, orc, shp, 0.3, 0.5, -0.25, 0.05, 11, 50, dtype=np.float32, form=form)
    assert rel_l2(got, want) < 1e-5
    assert np.abs(got - want).max() < 1e-5
    assert abs(st.mass() - psi.astype(np.float32).astype(np.float64).mean()) < 1e-15
    assert abs(got.astype(np.float64).mean() - st.mass()) < 2e-8


@pytest.mark.parametrize("form", FORMS)
def test_fp32_with_noise_matches_oracle(mcs, orc, form):
    got, want, psi, st = run_pair(mcs, orc, (32, 32, 32), 0.3, 0.25, -0.2, 0.1, 7, 20, temp=1e-3, noise_seed=7,
                                  dtype=np.float32, form=form)
    assert rel_l2(got, want) < 5e-5


def test_default_form_is_spectral(mcs):
    assert mcs.get_step_form() == mcs.FORM_SPECTRAL
    st = mcs.Stepper((32, 32, 32), [H] * 3, 0.3, 0.5)
    assert st.form() == mcs.FORM_SPECTRAL
    mcs.set_step_form(mcs.FORM_MIXED)
    try:
        assert mcs.Stepper((32, 32, 32), [H] * 3, 0.3, 0.5).form() == mcs.FORM_MIXED
        with pytest.raises(mcs.McsError):
            mcs.set_step_form(2)
    finally:
        mcs.set_step_form(mcs.FORM_SPECTRAL)
    with pytest.raises(mcs.McsError):
        st.set_form(2)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
@pytest.mark.parametrize("shp", [(32, 32, 32), (64, 96), (24, 40, 60)])
def test_forms_agree_convert_and_restore(mcs, orc, shp, dtype, tol):
    # the two forms run the same update on the same modes: they differ by rounding only; a
    # stepper converted mid-run (both ways) continues the same run; the mixed state is the
    # inverse x transform of the spectral state over nx/2 (numpy's normalised ifft), and a
    # restored mixed state continues bit for bit
    psi = orc.init_field(shp, -0.2, 0.1, 7)
    kw = dict(temp=1e-3, seed=7, dtype=dtype)
    a = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, form=mcs.FORM_SPECTRAL, **kw)
    b = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, form=mcs.FORM_MIXED, **kw)
    c = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, form=mcs.FORM_SPECTRAL, **kw)
    for st in (a, b, c):
        st.set_psi(psi)
    a.step(12)
    b.step(12)
    c.step(5)
    c.set_form(mcs.FORM_MIXED)
    assert c.form() == mcs.FORM_MIXED
    c.step(4)
    c.set_form(mcs.FORM_SPECTRAL)
    c.step(3)
    pa = a.get_psi()
    assert rel_l2(b.get_psi(), pa) < tol
    assert rel_l2(c.get_psi(), pa) < tol
    sa, ma, ka = a.get_state()
    sb, mb, kb = b.get_state()
    assert ma == mb and ka == kb == 12
    want = np.fft.ifft(sa.astype(np.complex128), axis=-1)
    assert np.linalg.norm(sb - want) / np.linalg.norm(want) < tol * 10
    d = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, form=mcs.FORM_MIXED, **kw)
    d.set_state(sb, mb, kb)
    d.step(3)
    b.step(3)
    assert np.array_equal(d.get_state()[0].view(np.uint8), b.get_state()[0].view(np.uint8))


def test_config1_2d_256_fp64_1000_steps(mcs, orc):
    """2D 256^2 FP64, eps=0.25, dt=0.5, psi0=-0.25 +- U[-0.05,0.05] seed 1, noise off, 1000 steps."""
    shp = (256, 256)
    got, want, psi, st = run_pair(mcs, orc, shp, 0.25, 0.5, -0.25, 0.05, 1, 1000)
    assert rel_l2(got, want) < 1e-10
    assert abs(got.mean() - want.mean()) < 1e-14
    pg = mcs.findpeaks(got, -1e9)
    po = orc.findpeaks(want, -1e9)
    assert np.array_equal(pg, po)
    assert np.array_equal(pg, orc.findpeaks(got, -1e9))


def test_resume_matches_uninterrupted_run(mcs, orc):
    shp = (32, 32, 32)
    psi = orc.init_field(shp, -0.2, 0.1, 3)
    kw = dict(temp=1e-3, seed=5, dtype=np.float64)
    a = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, **kw)
    a.set_psi(psi)
    a.step(10)
    b = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, **kw)
    b.set_psi(psi)
    b.step(4)
    mid = b.get_psi()
    c = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, **kw)
    mcs.check(mcs.lib().mcs_stepper_set_step(c.ptr, 4))
    c.set_psi(mid)
    c.step(6)
    assert rel_l2(c.get_psi(), a.get_psi()) < 1e-12
    assert a.steps_taken() == 10 and c.steps_taken() == 10


def test_single_mode_linear_decay(mcs):
    # small-amplitude single mode, linear regime: psi_k(n) = psi_k(0) / (1 + dt k^2 L(k))^n
    shp = (32, 64)
    y, x = np.meshgrid(np.arange(32), np.arange(64), indexing="ij")
    kx = 2 * np.pi * 3 / (64 * H)
    psi = 1e-6 * np.cos(kx * x * H)
    st = mcs.Stepper(shp, [H, H, H], 0.3, 0.5, dtype=np.float64)
    st.set_psi(psi)
    st.step(10)
    k2 = kx ** 2
    g = 1.0 / (1 + 0.5 * k2 * (-0.3 + (1 - k2) ** 2))
    assert np.abs(st.get_psi() - psi * g ** 10).max() < 1e-17 + 1e-9 * np.abs(psi).max()


def test_unsupported_sizes(mcs):
    for shp in [(7, 9), (30, 30, 30), (32, 14)]:
        with pytest.raises(mcs.McsError) as e:
            mcs.Stepper(shp, [H] * 3, 0.3, 0.5)
        assert e.value.code == mcs.ERR_UNSUPPORTED


# the sizes cuda/math.hpp:139-149's roundup235 yields beyond 2^a, 3*2^a and 5*2^a, among them three of the
# reference's FFT test sizes (cufft_unittest/my_unit_test.py:161-163): nx/2, ny, nz of the forms 9*2^a
# (1152 -> 576 and 1152: 12-point plans), 25*2^a (1600 -> 800 and 1600: 20 points) and 3*5^2*2^a or
# 3^2*5*2^a (1200 -> 600 and 1200, 360 -> 180 and 360: 30-point plans)
@pytest.mark.parametrize("shp,steps", [((1152, 1152), 6), ((1200, 1200), 6), ((1600, 1600), 4), ((360, 360, 360), 3),
                                       ((90, 150, 72), 8)])
def test_mixed_radix_sizes_match_oracle(mcs, orc, shp, steps):
    for L in set(shp) | {shp[-1] // 2}:
        P = mcs.plan_options(L, np.float32)[0]
        assert L % P == 0 and P in (6, 10, 12, 20, 30), (L, P)
    got, want, psi, st = run_pair(mcs, orc, shp, 0.3, 0.5, -0.25, 0.05, 11, steps)
    assert rel_l2(got, want) < 1e-12
    g32, w32, _, st32 = run_pair(mcs, orc, shp, 0.3, 0.5, -0.25, 0.05, 11, steps, dtype=np.float32)
    assert rel_l2(g32, w32) < 5e-6
    assert abs(st32.mass() - psi.astype(np.float32).astype(np.float64).mean()) < 1e-15


def test_mixed_radix_with_noise_fp32(mcs, orc):
    got, want, psi, st = run_pair(mcs, orc, (60, 72, 120), 0.3, 0.25, -0.2, 0.1, 7, 10, temp=1e-3, noise_seed=3,
                                  dtype=np.float32)
    assert rel_l2(got, want) < 5e-5
