"""GPU parity of the alternative integrators (exponential time differencing, hpfc noise,
polynomial operator) with the FP64 oracle's restatements."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(shp, dt, seed=0):
    rng = np.random.default_rng(seed)
    ct = np.complex128 if dt == np.float64 else np.complex64
    kp = np.fft.rfftn(rng.standard_normal(shp)).astype(ct)
    kn = np.fft.rfftn(0.2 * rng.standard_normal(shp)).astype(ct)
    lens = [n * 0.7 for n in shp][::-1] + [1.0] * (3 - len(shp))
    return kp, kn, lens


def close(a, b, tol):
    return np.abs(a - b).max() <= tol * np.abs(b).max()


TOL = [(np.float64, 1e-13), (np.float32, 3e-6)]


@pytest.mark.parametrize("dt,tol", TOL)
@pytest.mark.parametrize("shp", [(8, 12, 16), (24, 20)])
def test_etd_conserved(mcs, orc, shp, dt, tol):
    kp, kn, lens = case(shp, dt)
    kop = mcs.etd_op(shp, lens, 0.6, 0.25, False, dt)
    want_op = orc.etd_op(shp, lens, 0.6, 0.25, 0)
    # both sides take expf (single precision, as the reference does in both builds); device and host expf
    # may round differently by an ulp
    ttab = 2.4e-7 if dt == np.float64 else 5e-6  # FP32 also rounds (E - 1) / linop in single precision
    assert close(kop.real, want_op.real, ttab) and close(kop.imag, want_op.imag, ttab)
    xi = orc.kspace_noise(shp, 4, 0).astype(kp.dtype)
    n = int(np.prod(shp))
    namp_eff = 0.02 * float(np.sqrt(np.float32(n))) / np.sqrt(n)  # sqrtf(N) in both builds
    got = mcs.update_eom_etd(kop, kp, kn, xi, shp, lens, 0.02)
    want = orc.update_eom_etd(kop.astype(np.complex128), kp, kn, xi, shp, lens, namp_eff)
    assert close(got, want, tol * 4)
    got2, lpsi = mcs.update_eom_etd(kop, kp, kn, xi, shp, lens, 0.02, eps=0.25, with_fen=True)
    assert np.array_equal(got2, got)
    _, want_l = orc.update_eom_fen(kp, kn, np.zeros_like(kp), shp, lens, 0.0, 0.0, 0.25)
    assert close(lpsi, want_l, tol * 4)


@pytest.mark.parametrize("dt,tol", TOL)
def test_etd_nonconserved(mcs, orc, dt, tol):
    shp = (16, 24, 32)
    kp, kn, lens = case(shp, dt, 1)
    kop = mcs.etd_op(shp, lens, 0.3, 0.2, True, dt)
    assert close(kop, orc.etd_op(shp, lens, 0.3, 0.2, 1), 2.4e-7 if dt == np.float64 else 5e-6)
    got, m = mcs.update_el_etd(kop, kp, kn, shp)
    want, wm = orc.update_el_etd(kop.astype(np.complex128), kp, kn, shp)
    assert close(got, want, tol * 4) and abs(m - wm) <= tol * 4 * wm


@pytest.mark.parametrize("dt,tol", TOL)
def test_hpfc_noise_update(mcs, orc, dt, tol):
    shp = (48, 64)
    kp, kn, lens = case(shp, dt, 2)
    xs = [orc.kspace_noise(shp, 7 + j, 0).astype(kp.dtype) for j in range(3)]
    n = int(np.prod(shp))
    namp_eff = 0.05 * float(np.sqrt(np.float32(n))) / np.sqrt(n)
    got = mcs.update_eom_hpfc(kp, kn, *xs, shp, lens, 0.05, 0.3, 0.2)
    want = orc.update_eom_hpfc(kp, kn, *xs, shp, lens, namp_eff, 0.3, 0.2)
    assert close(got, want, tol * 4)
    with pytest.raises(mcs.McsError):
        mcs.update_eom_hpfc(*(case((4, 6, 8), dt)[:2]), *[case((4, 6, 8), dt)[1]] * 3, (4, 6, 8), [1, 1, 1], 0.1,
                            0.1, 0.1)


@pytest.mark.parametrize("dt,tol", TOL)
@pytest.mark.parametrize("shp", [(8, 12, 16), (24, 20)])
def test_polynomial_operator(mcs, orc, shp, dt, tol):
    kp, kn, lens = case(shp, dt, 3)
    xi = orc.kspace_noise(shp, 5, 0).astype(kp.dtype)
    n = int(np.prod(shp))
    namp_eff = 0.01 * float(np.sqrt(np.float32(n))) / np.sqrt(n)
    got = mcs.update_eom_poly(kp, kn, xi, shp, lens, 0.01, 0.2, 0.3)
    want = orc.update_eom_poly(kp, kn, xi, shp, lens, namp_eff, 0.2, 0.3)
    # the fitted polynomial cancels heavily (terms of several thousand for a value of 0.3 at k^2 = 3.8),
    # so an FP32 evaluation, the reference's included, keeps about 4 digits of the update (FP64 about 12)
    assert close(got, want, 1e-10 if dt == np.float64 else 5e-4)
