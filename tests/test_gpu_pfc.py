"""GPU parity of the pointwise PFC kernels, Philox streams and k-space noise against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_philox_streams_bit_exact(mcs, orc):
    for seed, c0, c2, c3 in [(0, 0, 0, 0), (0xDEADBEEF12345678, 2 ** 32 - 5, 17, 3)]:
        got = mcs.philox_blocks(seed, c0, c2, c3, 4096)
        want = orc.philox_stream(seed, c0 & 0xFFFFFFFF, c2, c3, 4096) if c0 < 2 ** 32 - 5 else None
        if want is None:  # counter crosses 2^32: compare block by block
            for i in (0, 3, 4, 5, 4095):
                c = c0 + i
                w = orc.philox([c & 0xFFFFFFFF, c >> 32, c2, c3], [seed & 0xFFFFFFFF, seed >> 32])
                assert list(got[4 * i:4 * i + 4]) == list(w)
        else:
            assert np.array_equal(got, want)


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
@pytest.mark.parametrize("shp", [(6, 8, 10), (5, 7, 9), (16, 12), (9,)])
def test_kspace_noise_matches_oracle(mcs, orc, shp, dt, tol):
    got = mcs.kspace_noise(shp, seed=123, draw=4, dtype=dt)
    want = orc.kspace_noise(shp, 123, 4)
    assert np.abs(got - want).max() < tol * 8
    # Hermitian pairs on the x = 0 (and nx/2) planes are exact conjugates on the GPU too
    r = mcs.irfftn(got.astype(np.complex128), shp) / np.prod(shp)
    assert np.abs(np.fft.rfftn(r) - got).max() < 1e-5


def test_real_noise_statistics(mcs):
    z = mcs.fill_normal(1 << 20, 0.5, seed=9, draw=0)
    assert abs(z.mean()) < 2e-3 and abs(z.std() - 0.5) < 2e-3
    z2 = mcs.fill_normal(1 << 20, 0.5, seed=9, draw=1)
    assert abs(np.mean(z * z2)) < 2e-3


@pytest.mark.parametrize("shp", [(4, 5, 6), (64, 64, 64), (256, 256)])
def test_init_field_bit_exact(mcs, orc, shp):
    want = orc.init_field(shp, -0.25, 0.05, 11)
    assert np.array_equal(mcs.init_field(shp, -0.25, 0.05, 11, np.float64), want)
    assert np.array_equal(mcs.init_field(shp, -0.25, 0.05, 11, np.float32), want.astype(np.float32))


def _spectra(shp, dt, seed=0):
    rng = np.random.default_rng(seed)
    hs = shp[:-1] + (shp[-1] // 2 + 1,)
    ct = np.complex128 if dt == np.float64 else np.complex64
    kp = np.fft.rfftn(rng.standard_normal(shp)).astype(ct)
    kn = np.fft.rfftn(0.1 * rng.standard_normal(shp)).astype(ct)
    assert kp.shape == hs
    return kp, kn


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-13), (np.float32, 3e-6)])
@pytest.mark.parametrize("shp", [(8, 12, 16), (24, 20)])
def test_update_eom_reference_form(mcs, orc, shp, dt, tol):
    kp, kn = _spectra(shp, dt)
    xi = orc.kspace_noise(shp, 3, 0)
    h = [np.pi / 4] * 3
    lens = [n * hh for n, hh in zip(shp[::-1], h)] + [1.0] * (3 - len(shp))
    N = np.prod(shp)
    namp, dtt, eps = 0.02, 0.5, 0.3
    got = mcs.update_eom(kp, kn, xi.astype(kp.dtype), shp, namp, dtt, eps, lens)
    # eom.hpp:60-72: knl enters with dt*dop.x = -dt k^2, noise scaled by dt*namp*sqrtf(N)*|k|
    # (eom.hpp:62 takes the square root in single precision in both builds)
    sqrtf_n = float(np.sqrt(np.float32(N)))
    want = orc.update(kp, kn, xi, shp, lens, dtt, eps, dtt * namp * sqrtf_n, orc.SPLIT_REFERENCE)
    assert np.abs(got - want).max() / np.abs(want).max() < tol
    assert got.reshape(-1)[0] == kp.reshape(-1)[0]  # k = 0 untouched bit for bit
    got0 = mcs.update_eom(kp, kn, None, shp, namp, dtt, eps, lens)
    want0 = orc.update(kp, kn, None, shp, lens, dtt, eps, 0.0, orc.SPLIT_REFERENCE)
    assert np.abs(got0 - want0).max() / np.abs(want0).max() < tol


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_update_pfc_with_philox_noise(mcs, orc, split, dt, tol):
    shp = (10, 12, 16)
    kp, kn = _spectra(shp, dt, 1)
    h = np.pi / 4
    lens = [16 * h, 12 * h, 10 * h]
    temp, dtt, eps, seed, draw = 1e-3, 0.25, 0.3, 77, 5
    got = mcs.update_pfc(kp, kn, shp, lens, dtt, eps, temp, seed, draw, split)
    amp = np.sqrt(2 * temp * dtt / h ** 3) * np.sqrt(np.prod(shp))
    want = orc.update(kp, kn, orc.kspace_noise(shp, seed, draw), shp, lens, dtt, eps, amp, split)
    assert np.abs(got - want).max() / np.abs(want).max() < tol


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_psi3_split_and_norm(mcs, dt):
    a = np.random.default_rng(2).standard_normal((6, 7, 8)).astype(dt)
    mu = 0.125
    got = mcs.psi3_split(a, mu=mu, norm=True)
    t = a.astype(np.float64) / a.size
    assert np.allclose(got, t ** 3 - t - mu, rtol=1e-6 if dt == np.float32 else 1e-14, atol=1e-7)
    pot = np.full_like(a, 0.5)
    got2 = mcs.psi3_split(a, mu=0.0, norm=False, pot=pot)
    t = a.astype(np.float64)
    assert np.allclose(got2, t ** 3 - t + 0.5, rtol=1e-6 if dt == np.float32 else 1e-14, atol=1e-6)
    assert np.allclose(mcs.norm(a), a.astype(np.float64) / a.size, rtol=1e-6)


@pytest.mark.parametrize("shp", [(16, 16, 16), (32, 24)])
def test_field_stats_fp64_reductions(mcs, orc, shp):
    psi = orc.init_field(shp, -0.2, 0.3, 5)
    h3 = [np.pi / 4] * 3
    mean, fen = mcs.field_stats(psi, h3, 0.25)
    assert abs(mean - psi.mean()) < 1e-15
    want = orc.free_energy(psi, h3, 0.25)
    assert abs(fen - want) < 1e-12 * abs(want)
    mean32, fen32 = mcs.field_stats(psi.astype(np.float32), h3, 0.25)
    assert abs(mean32 - psi.astype(np.float32).astype(np.float64).mean()) < 1e-15
    assert abs(fen32 - want) < 1e-5 * abs(want)


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-13), (np.float32, 3e-6)])
@pytest.mark.parametrize("shp", [(8, 12, 16), (24, 20)])
def test_update_eom_with_fen_and_calc_fen(mcs, orc, shp, dt, tol):
    kp, kn = _spectra(shp, dt)
    xi = orc.kspace_noise(shp, 3, 0).astype(kp.dtype)
    lens = [n * 0.7 for n in shp][::-1] + [1.0] * (3 - len(shp))
    got_k, got_l = mcs.update_eom_with_fen(kp, kn, xi, shp, 0.01, 0.4, 0.25, lens)
    # eom.hpp:90 scales namp by sqrtf(N) (single precision) in both builds
    n = int(np.prod(shp))
    namp_eff = 0.01 * float(np.sqrt(np.float32(n))) / np.sqrt(n)
    want_k, want_l = orc.update_eom_fen(kp, kn, xi, shp, lens, namp_eff, 0.4, 0.25)
    assert np.abs(got_k - want_k).max() < tol * np.abs(want_k).max()
    assert np.abs(got_l - want_l).max() < tol * np.abs(want_l).max()
    # the free-energy density of a field and its L psi, summed in FP64
    h = [0.7, 0.7, 0.7]
    psi = orc.init_field(shp, -0.2, 0.2, 5)
    lpsi = orc.c2r(want_l, shp) / psi.size
    f, dens = mcs.calc_fen(psi.astype(dt), lpsi.astype(dt), h)
    wf, wd = orc.calc_fen(psi.astype(dt).astype(np.float64), lpsi.astype(dt).astype(np.float64), h)
    assert abs(f - wf) < tol * 10 * np.abs(wd).sum() * np.prod(h)
    assert np.abs(dens - wd).max() < tol * np.abs(wd).max() + 1e-30


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-13), (np.float32, 3e-6)])
def test_update_el_matches_oracle(mcs, orc, dt, tol):
    shp = (16, 24, 32)
    kp, kn = _spectra(shp, dt)
    lens = [n * 0.7 for n in shp][::-1]
    got, m = mcs.update_el(kp, kn, shp, lens, 0.2, 0.3)
    want, wm = orc.update_el(kp, kn, shp, lens, 0.2, 0.3)
    assert np.abs(got - want).max() < tol * np.abs(want).max()
    assert abs(m - wm) < tol * wm
