"""Pins the FP64 CPU oracle (the checker) against golden vectors, published
known-answer tests and closed forms.  CPU only."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- Philox4x32-10
# Random123 known-answer vectors (kat_vectors, philox4x32_10)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(orc, ctr, key, want):
    assert [int(v) for v in orc.philox(ctr, key)] == want


def test_philox_stream_matches_single_calls(orc):
    seed = 0x0123456789ABCDEF
    s = orc.philox_stream(seed, 5, 7, 3, 10)
    for i in range(10):
        one = orc.philox([5 + i, 0, 7, 3], [seed & 0xFFFFFFFF, seed >> 32])
        assert list(s[4 * i:4 * i + 4]) == list(one)


# ---------------------------------------------------------------- FFT
def test_fft_small_golden(orc):
    g = np.load(os.path.join(GOLD, "fft_small.npz"))
    i = 0
    while f"x{i}" in g:
        x, k = g[f"x{i}"], g[f"k{i}"]
        got = orc.r2c(x)
        # cufft_unittest/my_unit_test.py:138-141 metric and FP64 tolerance
        assert np.abs(got - k).max() / np.sqrt(x.size) < 1e-14
        back = orc.c2r(k, x.shape)
        assert np.abs(back / x.size - x).max() < 1e-14
        i += 1
    assert i >= 10


def test_fft_reference_shapes_golden(orc):
    g = np.load(os.path.join(GOLD, "fft_refshapes.npz"))
    for i in range(int(g["count"])):
        shp = tuple(int(v) for v in g[f"shape{i}"])
        x = np.random.default_rng(int(g[f"seed{i}"])).standard_normal(shp)
        k = orc.r2c(x)
        tol = 1e-14 if len(shp) == 3 else 1e-13  # my_unit_test.py:97,141,185
        assert np.abs(k.reshape(-1)[g[f"sel{i}"]] - g[f"val{i}"]).max() / np.sqrt(x.size) < tol
        assert abs(np.linalg.norm(k) - float(g[f"norm{i}"])) < 1e-10 * float(g[f"norm{i}"])
        if len(shp) == 3:
            y = orc.c2r(k, shp)
            assert np.abs(x - y / x.size).max() < 1e-14  # round trip, my_unit_test.py:95-97


def test_fft_closed_forms(orc):
    n = (6, 10, 12)
    d = np.zeros(n)
    d[0, 0, 0] = 1.0
    assert np.allclose(orc.r2c(d), 1.0, atol=1e-15)  # delta -> constant
    i0 = np.arange(12)
    c = np.broadcast_to(np.cos(2 * np.pi * 3 * i0 / 12), n).copy()  # cos along x -> spike at kx=3
    k = orc.r2c(c)
    want = np.zeros_like(k)
    want[0, 0, 3] = c.size / 2
    assert np.abs(k - want).max() < 1e-12
    x = np.random.default_rng(1).standard_normal(n)
    kx = orc.r2c(x)
    full = np.sum(np.abs(kx[..., 0]) ** 2) + np.sum(np.abs(kx[..., 6]) ** 2) + 2 * np.sum(np.abs(kx[..., 1:6]) ** 2)
    assert abs(full / x.size - np.sum(x ** 2)) < 1e-10  # Parseval
    y = np.random.default_rng(2).standard_normal(n)
    assert np.abs(orc.r2c(2 * x - 3 * y) - (2 * kx - 3 * orc.r2c(y))).max() < 1e-12  # linearity


def test_fft_c2c_prime_lengths(orc):
    for nn in (7, 11, 13, 49, 97):
        x = np.random.default_rng(nn).standard_normal(nn) + 1j * np.random.default_rng(nn + 1).standard_normal(nn)
        assert np.abs(orc.fft_c2c(x, -1) - np.fft.fft(x)).max() < 1e-12
        assert np.abs(orc.fft_c2c(x, +1) - np.fft.ifft(x) * nn).max() < 1e-12


def test_c2r_ignores_imag_of_self_conjugate_x_modes(orc):
    k = np.random.default_rng(4).standard_normal((4, 6, 5)) + 1j * np.random.default_rng(5).standard_normal((4, 6, 5))
    got = orc.c2r(k, (4, 6, 8))
    assert np.abs(got / 192 - np.fft.irfftn(k, s=(4, 6, 8), axes=(0, 1, 2))).max() < 1e-14


# ---------------------------------------------------------------- noise
def _mirror(i, n):
    return (n - i) % n


def test_kspace_noise_hermitian_exact(orc):
    for shp in [(6, 8, 10), (5, 7, 9), (8, 12)]:
        xi = orc.kspace_noise(shp, seed=77, step=3)
        nx = shp[-1]
        planes = [0] + ([nx // 2] if nx % 2 == 0 else [])
        full = shp if len(shp) == 3 else (1,) + shp
        x3 = xi.reshape(full[0], full[1], -1)
        for i0 in planes:
            for i2 in range(full[0]):
                for i1 in range(full[1]):
                    a = x3[i2, i1, i0]
                    b = x3[_mirror(i2, full[0]), _mirror(i1, full[1]), i0]
                    assert a == np.conj(b)
        # a Hermitian half spectrum survives c2r -> r2c exactly up to rounding
        r = orc.c2r(xi, shp)
        assert np.abs(orc.r2c(r) / r.size - xi).max() < 1e-12


def test_kspace_noise_statistics(orc):
    xi = orc.kspace_noise((32, 32, 32), seed=5, step=0)
    inner = xi[:, :, 1:16].reshape(-1)
    assert abs(np.mean(np.abs(inner) ** 2) - 1.0) < 0.01
    assert abs(np.mean(inner.real)) < 0.01 and abs(np.mean(inner.imag)) < 0.01
    assert abs(np.mean(inner.real ** 2) - 0.5) < 0.01
    assert abs(np.mean(inner.real * inner.imag)) < 0.01
    other = orc.kspace_noise((32, 32, 32), seed=5, step=1)
    assert abs(np.mean(inner * np.conj(other[:, :, 1:16].reshape(-1)))) < 0.02  # steps independent


@pytest.mark.parametrize("shp", [(6, 8, 10), (4, 6, 12), (8, 12), (10, 16)])
def test_stepper_noise_hermitian_and_full_rank(orc, shp):
    xi = orc.stepper_noise(shp, seed=77, step=3)
    nx = shp[-1]
    full = shp if len(shp) == 3 else (1,) + shp
    x3 = xi.reshape(full[0], full[1], -1)
    nreal = 0
    for i0 in (0, nx // 2):
        for i2 in range(full[0]):
            for i1 in range(full[1]):
                a = x3[i2, i1, i0]
                m2, m1 = _mirror(i2, full[0]), _mirror(i1, full[1])
                assert a == np.conj(x3[m2, m1, i0])
                if (m2, m1) == (i2, i1):
                    assert a.imag == 0
                    nreal += 1
    r = orc.c2r(xi, shp)
    assert np.abs(orc.r2c(r) / r.size - xi).max() < 1e-12
    # full rank: every real degree of freedom of the grid gets its own normal, so the
    # magnitudes are distinct except for Hermitian partners (a reused Philox output would repeat one)
    n = int(np.prod(shp))
    assert np.unique(np.round(np.abs(xi), 12)).size == (n - nreal) // 2 + nreal


def test_stepper_noise_statistics(orc):
    xi = orc.stepper_noise((32, 32, 32), seed=5, step=0)
    inner = xi[:, :, 1:16].reshape(-1)
    assert abs(np.mean(np.abs(xi) ** 2) - 1.0) < 0.01
    assert abs(np.mean(inner.real)) < 0.01 and abs(np.mean(inner.imag)) < 0.01
    assert abs(np.mean(inner.real ** 2) - 0.5) < 0.01
    assert abs(np.mean(inner.real * inner.imag)) < 0.01
    other = orc.stepper_noise((32, 32, 32), seed=5, step=1)
    assert abs(np.mean(inner * np.conj(other[:, :, 1:16].reshape(-1)))) < 0.02
    # independent of the boundary layout's stream (a different Philox tag)
    assert abs(np.mean(inner * np.conj(orc.kspace_noise((32, 32, 32), 5, 0)[:, :, 1:16].reshape(-1)))) < 0.02


# ---------------------------------------------------------------- initial field
def test_init_field_from_philox(orc):
    shp = (4, 5, 6)
    psi = orc.init_field(shp, -0.25, 0.05, 11)
    s = orc.philox_stream(11, 0, 0, 1, psi.size).reshape(-1, 4).astype(np.float64)
    u = (np.floor(s[:, 0] / 32) * 67108864.0 + np.floor(s[:, 1] / 64)) / 9007199254740992.0
    assert np.array_equal(psi.reshape(-1), -0.25 + 0.05 * (2 * u - 1))
    big = orc.init_field((64, 64, 64), -0.25, 0.05, 11)
    assert big.min() >= -0.3 and big.max() <= -0.2 and abs(big.mean() + 0.25) < 1e-3


# ---------------------------------------------------------------- update and step
def _k2(shp, h):
    ks = [2 * np.pi * np.fft.fftfreq(n, d=hh) for n, hh in zip(shp, h)]
    ks[-1] = 2 * np.pi * np.fft.rfftfreq(shp[-1], d=h[-1])
    grids = np.meshgrid(*ks, indexing="ij")
    return sum(g ** 2 for g in grids)


def _numpy_step(psi, h, eps, dt, split):
    """The semi-implicit scheme written with numpy.fft from the PDE
    d psi/dt = lap[(-eps + (1+lap)^2) psi + psi^3], splitting L as the mode says."""
    k2 = _k2(psi.shape, h)
    kp = np.fft.rfftn(psi)
    if split == 0:
        nl, lin = psi ** 3, -eps + (1 - k2) ** 2
    else:
        nl, lin = psi ** 3 - psi, 1 - eps + (1 - k2) ** 2
    kp = (kp - dt * k2 * np.fft.rfftn(nl)) / (1 + dt * k2 * lin)
    return np.fft.irfftn(kp, s=psi.shape, axes=tuple(range(psi.ndim)))


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("shp", [(16, 12), (8, 10, 12)])
def test_step_matches_numpy_scheme(orc, split, shp):
    h = [np.pi / 4] * len(shp)
    psi0 = orc.init_field(shp, -0.25, 0.1, 9)
    want = psi0.copy()
    for _ in range(5):
        want = _numpy_step(want, h, 0.3, 0.5, split)
    h3 = list(h)[::-1] + [1.0] * (3 - len(h))
    got = orc.pfc_run(psi0, h3, 0.3, 0.5, 0.0, 0, 0, 5, split)
    assert np.abs(got - want).max() < 1e-13


def test_split_forms_agree_to_first_order(orc):
    # the two splits discretise the same PDE: their difference shrinks linearly with dt
    shp, h = (8, 8, 8), [np.pi / 4] * 3
    psi0 = orc.init_field(shp, -0.25, 0.1, 3)
    diffs = []
    for dt, n in [(0.02, 10), (0.01, 20)]:
        a = orc.pfc_run(psi0, h, 0.3, dt, 0.0, 0, 0, n, 0)
        b = orc.pfc_run(psi0, h, 0.3, dt, 0.0, 0, 0, n, 1)
        diffs.append(np.abs(a - b).max())
    assert diffs[1] < 0.6 * diffs[0]


def test_single_mode_linear_decay_converges_to_pde(orc):
    # linear PDE: psi_k(t) = psi_k(0) exp(-k^2 L(k) t); implicit Euler converges at O(dt)
    shp = (16, 16)
    lens = [16 * np.pi / 4] * 2 + [1.0]
    kpsi = np.zeros((16, 9), complex)
    kpsi[0, 2] = 1.0
    kx = 2 * np.pi * 2 / lens[0]
    rate = kx ** 2 * (-0.3 + (1 - kx ** 2) ** 2)
    errs = []
    for dt in (0.1, 0.05):
        k = kpsi.copy()
        for _ in range(int(round(1.0 / dt))):
            k = orc.update(k, np.zeros_like(k), None, shp, lens, dt, 0.3, 0.0, 0)
        errs.append(abs(k[0, 2] - np.exp(-rate)))
    assert errs[1] < 0.55 * errs[0] and errs[1] < 0.05


def test_update_k0_bit_exact_and_noise_scaling(orc):
    shp, lens = (6, 8, 10), [6.0, 7.0, 8.0]
    rng = np.random.default_rng(1)
    kp = rng.standard_normal((6, 8, 6)) + 1j * rng.standard_normal((6, 8, 6))
    kn = rng.standard_normal((6, 8, 6)) + 1j * rng.standard_normal((6, 8, 6))
    xi = orc.kspace_noise(shp, 1, 0)
    for split in (0, 1):
        out = orc.update(kp, kn, xi, shp, lens, 0.3, 0.2, 2.0, split)
        assert out[0, 0, 0] == kp[0, 0, 0]
    # noise enters linearly with |k|
    a = orc.update(kp, kn, xi, shp, lens, 0.3, 0.2, 2.0, 0)
    b = orc.update(kp, kn, None, shp, lens, 0.3, 0.2, 0.0, 0)
    k2 = _k2((6, 8, 10), [l / n for l, n in zip(lens[::-1], (6, 8, 10))])
    D = 1 + 0.3 * k2 * (-0.2 + (1 - k2) ** 2)
    assert np.abs((a - b) * D - 2.0 * np.sqrt(k2) * xi).max() < 1e-12


def test_mass_conserved_with_noise(orc):
    psi0 = orc.init_field((8, 8, 8), -0.2, 0.1, 7)
    psi, mass = orc.pfc_run(psi0, [np.pi / 4] * 3, 0.3, 0.25, 1e-3, 7, 0, 20, 0, want_mass=True)
    assert np.abs(mass - mass[0]).max() < 1e-15
    assert abs(psi.mean() - mass[0]) < 1e-15


def test_free_energy_uniform_closed_form(orc):
    psi = np.full((8, 8, 8), -0.3)
    h = [0.5, 0.5, 0.5]
    eps = 0.25
    f = orc.free_energy(psi, h, eps)
    V = 8 ** 3 * 0.125
    assert abs(f / V - (0.5 * (1 - eps) * 0.09 + 0.25 * 0.3 ** 4)) < 1e-14


def test_free_energy_decreases_noise_free(orc):
    h = [np.pi / 4] * 2 + [1.0]
    psi = orc.init_field((32, 32), -0.25, 0.1, 1)
    f = [orc.free_energy(psi, h, 0.25)]
    for _ in range(5):
        psi = orc.pfc_run(psi, h, 0.25, 0.5, 0.0, 0, 0, 20, 0)
        f.append(orc.free_energy(psi, h, 0.25))
    assert all(b < a for a, b in zip(f, f[1:]))


def test_update_eom_fen_is_the_reference_split_update(orc):
    shp, lens = (6, 8, 10), [6 * 0.7, 8 * 0.7, 10 * 0.6]
    rng = np.random.default_rng(3)
    kp = np.fft.rfftn(rng.standard_normal(shp))
    kn = np.fft.rfftn(0.1 * rng.standard_normal(shp))
    xi = orc.kspace_noise(shp, 5, 0)
    namp, dt, eps = 0.02, 0.3, 0.2
    got, lpsi = orc.update_eom_fen(kp, kn, xi, shp, lens[::-1], namp, dt, eps)
    # the same step as the reference-split update with noise amplitude dt*namp*sqrt(N)
    want = orc.update(kp, kn, xi, shp, lens[::-1], dt, eps, dt * namp * np.sqrt(np.prod(shp)), orc.SPLIT_REFERENCE)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    k2 = _k2(shp, [l / n for l, n in zip(lens, shp)])
    want_l = ((1 - k2) ** 2 + 1 - eps) * kp
    assert np.abs(lpsi - want_l).max() < 1e-13 * np.abs(want_l).max()


def test_calc_fen_is_the_free_energy(orc):
    # with lpsi = L psi (the real field of lk * psi_k), calc_fen is the free energy functional
    h = [0.6, 0.7, 0.5]
    shp = (8, 10, 12)
    psi = orc.init_field(shp, -0.2, 0.2, 9)
    eps = 0.25
    k2 = _k2(shp, h[::-1])
    lpsi = np.fft.irfftn(((1 - k2) ** 2 + 1 - eps) * np.fft.rfftn(psi), s=shp, axes=(0, 1, 2))
    f, dens = orc.calc_fen(psi, lpsi, h)
    assert abs(f - orc.free_energy(psi, h, eps)) < 1e-12 * abs(f)
    assert np.allclose(dens, 0.5 * psi * (lpsi - psi + 0.5 * psi ** 3), rtol=1e-14, atol=1e-16)
    # uniform field closed form (common/pfc.hpp:4-5 without mu)
    u = np.full(shp, -0.3)
    fu, _ = orc.calc_fen(u, (2 - eps) * u, h)
    V = np.prod(shp) * np.prod(h)
    assert abs(fu / V - (0.5 * (1 - eps) * 0.09 + 0.25 * 0.3 ** 4)) < 1e-14


def test_update_el_relaxes(orc):
    shp, h = (32, 32), [np.pi / 4, np.pi / 4, 1.0]
    lens = [32 * h[0], 32 * h[1], 1.0]
    # single mode, no nonlinearity: psi_k / denom and the change closed form
    kp = np.zeros((32, 17), complex)
    kp[3, 5] = 2.0 + 1.0j
    k2 = (2 * np.pi * 3 / lens[1]) ** 2 + (2 * np.pi * 5 / lens[0]) ** 2
    den = 1 + 0.1 * ((1 - k2) ** 2 + 1 - 0.3)
    got, m = orc.update_el(kp, np.zeros_like(kp), shp, lens, 0.1, 0.3)
    assert abs(got[3, 5] - kp[3, 5] / den) < 1e-15 and abs(m - abs(kp[3, 5]) * (1 - 1 / den)) < 1e-15
    # with the nonlinearity psi^3 - psi the relaxation lowers F and the change shrinks
    psi = orc.init_field(shp, -0.3, 0.2, 4)
    f0, dm = orc.free_energy(psi, h, 0.3), []
    fs = [f0]
    for _ in range(40):
        kp = np.fft.rfftn(psi)
        got, m = orc.update_el(kp, np.fft.rfftn(psi ** 3 - psi), shp, lens, 0.1, 0.3)
        psi = np.fft.irfftn(got, s=shp, axes=(0, 1))
        fs.append(orc.free_energy(psi, h, 0.3))
        dm.append(m)
    assert all(b < a for a, b in zip(fs, fs[1:]))
    assert dm[-1] < dm[0]


def _etd_case():
    shp = (6, 8, 10)
    h = [0.7, 0.6, 0.5]
    lens = [n * hh for n, hh in zip(shp[::-1], h)]
    rng = np.random.default_rng(11)
    kp = np.fft.rfftn(rng.standard_normal(shp))
    nl = np.fft.rfftn(0.3 * rng.standard_normal(shp))
    k2 = _k2(shp, h[::-1])
    return shp, lens, kp, nl, k2


def test_etd_conserved_is_exact_for_constant_forcing(orc):
    # d psi_k/dt = -k^2 L' psi_k - k^2 N_k with N fixed: one ETD1 step of any length is the exact solution
    shp, lens, kp, nl, k2 = _etd_case()
    eps, dt = 0.25, 0.7
    lp = (1 - k2) ** 2 + 1 - eps
    kop = orc.etd_op(shp, lens, dt, eps, 0)
    got = orc.update_eom_etd(kop, kp, nl, None, shp, lens, 0.0)
    lam = k2 * lp
    e = np.exp(-lam * dt)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(lam == 0, -dt, (e - 1) / lam)
    want = e * kp + s * k2 * nl
    assert np.abs(got - want).max() < 2e-7 * np.abs(want).max()  # the table holds single-precision exponentials
    assert kop[0, 0, 0] == 1 - 1j * dt and got[0, 0, 0] == kp[0, 0, 0]  # k = 0: mass unchanged, S = -dt


def test_etd_nonconserved_is_exact_for_constant_forcing(orc):
    shp, lens, kp, nl, k2 = _etd_case()
    eps, dt = 0.25, 0.4
    lp = (1 - k2) ** 2 + 1 - eps
    kop = orc.etd_op(shp, lens, dt, eps, 1)
    got, m = orc.update_el_etd(kop, kp, nl, shp)
    e = np.exp(-lp * dt)
    want = e * kp + (e - 1) / lp * nl
    assert np.abs(got - want).max() < 2e-7 * np.abs(want).max()
    assert abs(m - np.abs(got - kp).max()) < 1e-12 * m


def test_hpfc_update_reduces_to_reference_split_and_filters(orc):
    shp = (48, 64)
    lens = [64 * 0.7, 48 * 0.7, 1.0]
    rng = np.random.default_rng(2)
    kp = np.fft.rfftn(rng.standard_normal(shp))
    nl = np.fft.rfftn(0.2 * rng.standard_normal(shp))
    z = np.zeros_like(kp)
    got = orc.update_eom_hpfc(kp, nl, z, z, z, shp, lens, 0.05, 0.3, 0.2)
    want = orc.update(kp, nl, None, shp, lens, 0.3, 0.2, 0.0, orc.SPLIT_REFERENCE)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_hpfc_noise_is_isotropic_with_the_fixed_total_weight(orc):
    """eom.hpp:16-46: the three noise arrays enter with weights w1, w2, w3 per mode; for
    independent unit xi the per-mode variance w1^2 + w2^2 + w3^2 must not depend on the direction
    of k.  For weights (c1 kx^2 + c2 ky^2, c2 kx^2 + c1 ky^2, 2 kx ky) with c1 c2 = kappa/2 that
    holds exactly when (c2 - c1)^2 = 2, and then the variance is (kappa + 2)|k|^4, kappa = 4/3.
    Each weight is read off by feeding a unit field into one array at a time; the filter keeps
    |k| <= 1/4 only; the noise is linear in the three arrays."""
    shp = (64, 64)
    hh = 0.7
    lens = [64 * hh, 64 * hh, 1.0]
    dt, namp, eps = 0.3, 0.05, 0.2
    rng = np.random.default_rng(2)
    kp = np.fft.rfftn(rng.standard_normal(shp))
    nl = np.fft.rfftn(0.2 * rng.standard_normal(shp))
    z = np.zeros_like(kp)
    one = np.ones_like(kp)
    base = orc.update_eom_hpfc(kp, nl, z, z, z, shp, lens, namp, dt, eps)
    ky = 2 * np.pi * np.fft.fftfreq(64, d=hh)[:, None]
    kx = 2 * np.pi * np.fft.rfftfreq(64, d=hh)[None, :]
    k2 = kx ** 2 + ky ** 2
    den = 1 + dt * k2 * ((1 - k2) ** 2 + 1 - eps)
    scale = dt * namp * np.sqrt(64 * 64)
    w = []
    for j in range(3):
        xs = [z, z, z]
        xs[j] = one
        w.append(((orc.update_eom_hpfc(kp, nl, *xs, shp, lens, namp, dt, eps) - base) * den / scale).real)
    inside = np.sqrt(k2) <= 0.25
    assert np.any(inside & (k2 > 0)) and np.any(~inside)
    var = w[0] ** 2 + w[1] ** 2 + w[2] ** 2
    kappa = 4.0 / 3.0
    assert np.abs(var[inside] - (kappa + 2) * k2[inside] ** 2).max() < 1e-12 * ((kappa + 2) * 0.25 ** 4)
    assert np.all(var[~inside] == 0)
    # the cross array carries 2 kx ky; the other two swap roles under kx <-> ky (a square box)
    assert np.abs(w[2][inside] - (2 * kx * ky * np.ones_like(k2))[inside]).max() < 1e-14
    for iy in range(1, 5):
        for ix in range(1, 5):
            if np.sqrt(kx[0, ix] ** 2 + ky[iy, 0] ** 2) <= 0.25:
                assert abs(w[0][iy, ix] - w[1][ix, iy]) < 1e-14
    # linear in the three arrays
    xs = [orc.kspace_noise(shp, 7 + j, 0) for j in range(3)]
    got = orc.update_eom_hpfc(kp, nl, *xs, shp, lens, namp, dt, eps)
    assert np.abs((got - base) * den / scale - (w[0] * xs[0] + w[1] * xs[1] + w[2] * xs[2])).max() < 1e-12


def test_polynomial_operator_update(orc):
    """eom.hpp:151-171 evaluates 1 - eps - P(k^2) with a fitted degree-6 P.  With P(q) = -(1 - q)^2 =
    -1 + 2q - q^2 the operator is the PFC one, 1 - eps + (1 - k^2)^2, so the update must equal the
    reference-split semi-implicit update (eom.hpp:60-72) term for term, noise included; with the
    fitted coefficients a single mode decays by 1/denom and k = 0 is unchanged."""
    shp, lens, kp, nl, k2 = _etd_case()
    xi = orc.kspace_noise(shp, 3, 0)
    n = int(np.prod(shp))
    dt, eps, namp = 0.2, 0.3, 0.07
    got = orc.update_eom_poly(kp, nl, xi, shp, lens, namp, dt, eps, coef=[-1.0, 2.0, -1.0])
    want = orc.update(kp, nl, xi, shp, lens, dt, eps, dt * namp * np.sqrt(n), orc.SPLIT_REFERENCE)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # the same with the trailing zero coefficients of a degree-6 polynomial
    got6 = orc.update_eom_poly(kp, nl, xi, shp, lens, namp, dt, eps, coef=[-1.0, 2.0, -1.0, 0, 0, 0, 0])
    assert np.abs(got6 - want).max() < 1e-13 * np.abs(want).max()
    c = orc.poly_coefficients()
    assert c.size == 7
    got = orc.update_eom_poly(kp, np.zeros_like(nl), None, shp, lens, 0.0, dt, eps)
    den = 1 + dt * k2 * (1 - eps - np.polynomial.polynomial.polyval(k2, c))
    assert np.abs(got - kp / den).max() < 1e-12 * np.abs(kp).max()
    assert got[0, 0, 0] == kp[0, 0, 0]


# ---------------------------------------------------------------- peaks
def _brute_peaks(v, ths):
    nd = v.ndim
    peak = v >= ths
    import itertools
    for d in itertools.product((-1, 0, 1), repeat=nd):
        if all(x == 0 for x in d):
            continue
        peak &= v > np.roll(v, shift=[-x for x in d], axis=tuple(range(nd)))
    return np.flatnonzero(peak.reshape(-1))


@pytest.mark.parametrize("shp", [(7, 9, 11), (12, 10), (3, 4, 5), (2, 2, 2)])
def test_peaks_brute_force(orc, shp):
    v = np.random.default_rng(sum(shp)).standard_normal(shp)
    for ths in (-10.0, 0.5):
        assert np.array_equal(orc.findpeaks(v, ths), _brute_peaks(v, ths))


def test_peaks_special_cases(orc):
    v = np.zeros((6, 6, 6))
    assert orc.findpeaks(v, -1).size == 0  # plateau: strict comparison
    v[5, 5, 5] = 1.0  # corner, periodic wrap
    assert list(orc.findpeaks(v, 1.0)) == [215]  # threshold equality counts
    assert orc.findpeaks(v, 1.0000001).size == 0
    v[0, 0, 0] = 1.0  # periodic neighbour of the corner: equal values, neither is strict
    assert orc.findpeaks(v, 0.5).size == 0
    c = np.zeros((8, 8))
    c[::2, ::2] = 1.0  # one peak per 2x2 block: the N/4 bound
    assert orc.findpeaks(c, 0.5).size == 16


# ---------------------------------------------------------------- bond order
def _lattice(kind, m=4):
    b = {"sc": [(0, 0, 0)], "bcc": [(0, 0, 0), (1, 1, 1)],
         "fcc": [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]}[kind]
    a = 2
    pts = [((i * a + x) % (m * a), (j * a + y) % (m * a), (k * a + z) % (m * a))
           for i in range(m) for j in range(m) for k in range(m) for (x, y, z) in b]
    return np.array(pts), m * a


# Steinhardt, Nelson & Ronchetti (1983), Table I
LATTICE = {"fcc": (12, 0.19094, 0.57452), "bcc": (8, 0.50918, 0.62854), "sc": (6, 0.76376, 0.35355)}


@pytest.mark.parametrize("kind", ["fcc", "bcc", "sc"])
def test_bond_order_lattices(orc, kind):
    pts, n = _lattice(kind)
    k, q4, q6 = LATTICE[kind]
    nbr, out = orc.bond_order_knn(pts, n, 1.0, k)
    assert np.allclose(out[:, 0], q4, atol=5e-5) and np.allclose(out[:, 1], q6, atol=5e-5)
    rlim = {"fcc": 1.5, "bcc": 1.8, "sc": 2.1}[kind]
    nb, o9 = orc.bond_order_cutoff(pts.astype(float), float(n), rlim)
    assert np.all(nb == k)
    assert np.allclose(o9[:, 0], q4, atol=5e-5) and np.allclose(o9[:, 1], q6, atol=5e-5)
    assert np.allclose(o9[:, 4], q4, atol=5e-5)  # neighbour average of identical environments


def test_bond_order_hcp_and_icosahedron(orc):
    # hcp: the 12 neighbours of one atom in an ideal hcp crystal
    a = 1.0
    nb = [(a * np.cos(t), a * np.sin(t), 0.0) for t in np.arange(6) * np.pi / 3]
    c = np.sqrt(2.0 / 3.0) * a
    for s in (1, -1):
        for t in np.arange(3) * 2 * np.pi / 3 + np.pi / 6:
            nb.append((a / np.sqrt(3) * np.cos(t), a / np.sqrt(3) * np.sin(t), s * c))
    # icosahedron vertices
    g = (1 + 5 ** 0.5) / 2
    ico = [(0, s1, s2 * g) for s1 in (1, -1) for s2 in (1, -1)]
    ico = ico + [(y, z, x) for x, y, z in ico] + [(z, x, y) for x, y, z in ico]
    for pts, q4w, q6w in [(nb, 0.09722, 0.48476), (ico, 0.0, 0.66332)]:
        v = np.array(pts, float)
        r = np.linalg.norm(v, axis=1)
        ct, ph = v[:, 2] / r, np.arctan2(v[:, 1], v[:, 0])
        ql = []
        for l in (4, 6):
            qm = [np.mean([orc.ylm(l, m, a_, b_) for a_, b_ in zip(ct, ph)]) for m in range(-l, l + 1)]
            ql.append(np.sqrt(4 * np.pi / (2 * l + 1) * np.sum(np.abs(qm) ** 2)))
        assert abs(ql[0] - q4w) < 5e-5 and abs(ql[1] - q6w) < 5e-5


def test_ylm_against_scipy(orc):
    sp = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(0)
    for _ in range(50):
        l = int(rng.choice([4, 6, 8]))
        m = int(rng.integers(-l, l + 1))
        th, ph = rng.uniform(0, np.pi), rng.uniform(-np.pi, np.pi)
        assert abs(orc.ylm(l, m, np.cos(th), ph) - sp.sph_harm_y(l, m, th, ph)) < 1e-12


def test_wigner3j_orthogonality(orc):
    assert abs(orc.wigner3j(4, 4, 4, 0, 0, 0) - np.sqrt(18 / 1001)) < 1e-15
    s = sum(orc.wigner3j(4, 4, 4, m1, m2, -m1 - m2) ** 2 for m1 in range(-4, 5) for m2 in range(-4, 5))
    assert abs(s - 1.0) < 1e-14  # sum over m1, m2 of (3j)^2 = 1 (the 2*4+1 values of m3 each give 1/9)


def test_bond_order_cutoff_matches_reference_build(orc):
    g = np.load(os.path.join(GOLD, "bond_order_ref.npz"))
    for c in range(int(g["count"])):
        pts, box, rlim, table = g[f"pts{c}"], float(g[f"box{c}"]), float(g[f"rlim{c}"]), g[f"table{c}"]
        nb, out = orc.bond_order_cutoff(pts, box, rlim)
        m = nb > 0
        mine = np.column_stack([pts[m], nb[m], out[m][:, 8], out[m][:, :8]])
        assert mine.shape == table.shape
        assert np.array_equal(mine[:, :5], table[:, :5])  # positions, neighbour counts, xi exactly
        # the reference prints %.6f; its closed-form Y_lm (bond-order.cpp:135-175) differ from the
        # recurrence used here by ~1e-8, which can move the last printed digit
        # coincident random points give 0/0 bond angles in both codes: NaN where the reference has NaN
        a, b = mine[:, 5:], table[:, 5:]
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.nanmax(np.abs(a - b)) <= 6e-7


def test_knn_neighbours_brute_force(orc):
    rng = np.random.default_rng(8)
    n = 20
    pts = rng.choice(n ** 3, 150, replace=False)
    ijk = np.column_stack([pts % n, (pts // n) % n, pts // (n * n)]).astype(np.int32)
    nbr, _ = orc.bond_order_knn(ijk, n, 1.0, 12)
    d = ijk[None, :, :] - ijk[:, None, :]
    d = (d + n // 2) % n - n // 2
    d2 = (d ** 2).sum(-1)
    for i in range(len(ijk)):
        order = sorted((int(d2[i, j]), j) for j in range(len(ijk)) if j != i)[:12]
        assert list(nbr[i]) == [j for _, j in order]


def test_knn_points_equals_full_rows(orc):
    # the query-subset form (used to check full-size GPU runs on a sample) is the full form's rows
    rng = np.random.default_rng(5)
    n = 24
    pts = rng.choice(n ** 3, 300, replace=False)
    ijk = np.column_stack([pts % n, (pts // n) % n, pts // (n * n)]).astype(np.int32)
    nbr, out = orc.bond_order_knn(ijk, n, 0.5, 12)
    q = rng.choice(300, 40, replace=False)
    snbr, sout = orc.bond_order_knn_points(ijk, n, 0.5, q, 12)
    assert np.array_equal(snbr, nbr[q])
    assert np.array_equal(sout, out[q])
