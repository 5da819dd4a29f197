"""Pins of the oracle's post-processing restatements (CPU only):

* the Gaussian filter table against what its mathematics fixes: the blurred
  delta keeps unit sum and has per-axis variance (1.1 divs_per_sigma)^2 in grid
  units (anisotropic spacings catch a mixed-up axis), the high-pass part removes
  the mean, the table factorises per axis;
* the low-pass amplitude of a single cosine mode against its closed form, and
  the text dump's header and kept modes;
* pair distances against the reference's own particlepairs tool built from its
  source (oracle/_ref), against the shells of a simple cubic lattice, and the
  minimum-image tie at exactly half the box.
"""
import re
import subprocess

import numpy as np
import pytest


def grid_coord(n):
    """minimum-image grid coordinate of each index, in [-n/2, n/2)"""
    j = np.arange(n)
    return np.where(j < (n + 1) // 2, j, j - n).astype(np.float64)


@pytest.mark.parametrize("divs", [2.5, 3.0])
def test_gaussian_blur_of_delta_has_unit_sum_and_variance(orc, divs):
    # widths of 2.75 and 3.3 cells: the filter is below 1e-16 at the Nyquist modes and the
    # blurred delta below 1e-11 at half the box, so the discrete moments are the continuous ones
    shp = (48, 56, 64)  # (nz, ny, nx)
    h = [0.7, 0.9, 1.3]  # hx, hy, hz
    lens = [shp[2] * h[0], shp[1] * h[1], shp[0] * h[2]]
    kg = orc.gaussian_init(shp, lens, h, divs)
    assert kg[0, 0, 0] == 1.0 + 0.0j
    delta = np.zeros(shp)
    delta[0, 0, 0] = 1.0
    k = orc.r2c(delta)
    low = orc.c2r(k * kg.real, shp) / delta.size
    high = orc.c2r(k * kg.imag, shp) / delta.size
    assert abs(low.sum() - 1.0) < 1e-13
    assert abs(high.sum()) < 1e-13
    np.testing.assert_allclose(low + high, delta, atol=1e-15)
    s2 = (1.1 * divs) ** 2  # grid units: s_i k_i = 1.1 divs 2 pi j / n_i for lens_i = n_i h_i
    for axis in range(3):
        c = grid_coord(shp[axis])
        shape = [1, 1, 1]
        shape[axis] = -1
        var = (low * c.reshape(shape) ** 2).sum()
        assert abs(var - s2) < 1e-8 * s2, (axis, var, s2)


def test_gaussian_table_factorises_per_axis(orc):
    shp = (6, 10, 16)
    h = [0.5, 1.25, 2.0]
    lens = [16 * 0.5, 10 * 1.25, 6 * 2.0]
    kg = orc.gaussian_init(shp, lens, h, 2.0).real
    ex = kg[0, 0, :]
    ey = kg[0, :, 0]
    ez = kg[:, 0, 0]
    np.testing.assert_allclose(kg, ez[:, None, None] * ey[None, :, None] * ex[None, None, :], rtol=1e-13, atol=1e-300)
    # one mode by hand: k = 2 pi j / L, s = 1.1 divs h
    kx = 2 * np.pi * 3 / lens[0]
    assert abs(ex[3] - np.exp(-0.5 * (1.1 * 2.0 * h[0] * kx) ** 2)) < 1e-16
    ky = 2 * np.pi * (7 - 10) / lens[1]  # index 7 of 10 is the wavenumber -3
    assert abs(ey[7] - np.exp(-0.5 * (1.1 * 2.0 * h[1] * ky) ** 2)) < 1e-16


def cosine_case(m=3, a=0.4, divs=0.8):
    ny, nx, h = 32, 64, np.pi / 4
    lens = [nx * h, ny * h, 1.0 * h]
    x = np.arange(nx) * h
    q = 2 * np.pi * m / lens[0]
    psi = np.tile(a * np.cos(q * x), (ny, 1))
    s = 1.1 * divs * h
    eq, e2q = np.exp(-0.5 * (s * q) ** 2), np.exp(-0.5 * (s * 2 * q) ** 2)
    amp = a * a * (1 - eq) ** 2 / 2 * (1 + e2q * np.cos(2 * q * x))
    return psi, lens, [h, h, h], divs, np.tile(amp, (ny, 1)), (a * a * (1 - eq) ** 2 / 2, e2q, m)


def test_lowpass_amplitude_of_a_cosine_matches_closed_form(orc):
    psi, lens, h, divs, want, _ = cosine_case()
    spec, amp = orc.amplitude_lowpass(psi, lens, h, divs)
    np.testing.assert_allclose(amp, want, atol=1e-14)


def test_lowpass_text_header_and_modes(orc):
    psi, lens, h, divs, _, (mean, e2q, m) = cosine_case()
    spec, _ = orc.amplitude_lowpass(psi, lens, h, divs)
    txt = orc.lowpass_text(spec, psi.shape, lens[:2])
    lines = txt.split("\n")
    assert lines[0] == "LOWPASS"
    head = "\n".join(lines[2:11]) + "\n"
    assert int(lines[1]) == len(head)
    got = re.fullmatch(r"\(S'n0'\nI(\d+)\nS'n1'\nI(\d+)\nS'l0'\nF([0-9.]+)\nS'l1'\nF([0-9.]+)\nd\.\n", head)
    assert got and (int(got[1]), int(got[2])) == (64, 32)
    assert abs(float(got[3]) - lens[0]) < 1e-6 and abs(float(got[4]) - lens[1]) < 1e-6
    rows = [ln.split() for ln in lines[11:] if ln]
    N = psi.size
    # the mean (0, 0) and the 2q mode: (2m, 0) with N mean e2q / 2
    assert [(int(r[0]), int(r[1])) for r in rows] == [(0, 0), (2 * m, 0)]
    assert abs(float(rows[0][2]) - N * mean) < 1e-5 * N * mean
    assert abs(float(rows[1][2]) - N * mean * e2q / 2) < 1e-5 * N * mean * e2q / 2


def test_lowpass_text_threshold_uses_squared_modulus(orc):
    spec = np.zeros((4, 3), np.complex128)
    spec[1, 2] = 0.0011  # 1.21e-6: kept
    spec[2, 1] = 0.0009j  # 8.1e-7: dropped
    spec[3, 0] = 0.0008 + 0.0007j  # 1.13e-6: kept
    txt = orc.lowpass_text(spec, (4, 4), [1.0, 2.0])
    body = txt.split("d.\n", 1)[1]
    assert body == "0 3 0.0008 0.0007\n2 1 0.0011 0\n"


def write_xyz(path, xyz):
    with open(path, "w") as f:
        f.write("%d\nframe\n" % len(xyz))
        for p in xyz:
            f.write("X %.17g %.17g %.17g\n" % tuple(p))


def test_pair_distances_match_reference_tool(orc, tmp_path):
    exe = orc.ref_particlepairs_binary()
    if exe is None:
        pytest.skip("oracle/_ref/particlepairs not built (needs the reference sources)")
    rng = np.random.default_rng(11)
    box = 17.5
    xyz = rng.uniform(0, box, (60, 3))
    xyz[1] = xyz[0] + [box / 2, 0, 0]  # the tie at exactly half the box
    xyz[1] %= box
    f = tmp_path / "pts.xyz"
    write_xyz(f, xyz)
    with open(f) as fin:
        r = subprocess.run([exe, "%.17g" % box], stdin=fin, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    ref = np.array([float(v) for v in r.stdout.split()])
    got = orc.pair_distances(xyz, box)
    assert ref.shape == got.shape == (60 * 59,)
    np.testing.assert_allclose(got, ref, rtol=6e-6)  # the tool prints 6 significant digits


def test_pair_distances_simple_cubic_shells(orc):
    m, a = 5, 1.5
    g = np.arange(m) * a
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    d = orc.pair_distances(xyz, m * a)
    P = len(xyz)
    for r, per_point in [(1.0, 6), (np.sqrt(2), 12), (np.sqrt(3), 8), (2.0, 6)]:
        assert np.sum(np.abs(d - r * a) < 1e-12) == P * per_point, r
    assert d.max() <= np.sqrt(3) * (m // 2) * a + 1e-12


def test_pair_distances_half_box_tie_and_small_sets(orc):
    box = 10.0
    d = orc.pair_distances(np.array([[0.0, 0, 0], [5.0, 0, 0]]), box)
    assert list(d) == [5.0, 5.0]
    d = orc.pair_distances(np.array([[1.0, 0, 0], [9.0, 0, 0]]), box)
    np.testing.assert_allclose(d, [2.0, 2.0], rtol=0, atol=1e-15)
    assert orc.pair_distances(np.zeros((1, 3)), box).size == 0
    assert orc.pair_distances(np.zeros((0, 3)), box).size == 0


def test_pair_histogram_counts_unordered_pairs(orc):
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0, 8, (40, 3))
    h = orc.pair_histogram(xyz, 8.0, 4.0, 16)
    d = orc.pair_distances(xyz, 8.0)
    assert h.sum() * 2 == np.sum(d < 4.0)


def test_package_writer_matches_oracle_text(mcs, orc, tmp_path):
    """the package's gz writer (host code) produces the oracle's text for the same spectrum"""
    import gzip
    rng = np.random.default_rng(1)
    for ct, rt in [(np.complex128, np.float64), (np.complex64, np.float32)]:
        spec = (rng.standard_normal((12, 9)) * 2e-3 + 1j * rng.standard_normal((12, 9)) * 2e-3).astype(ct)
        fn = str(tmp_path / "lp.gz")
        mcs.write_lowpass_2d(spec, (12, 16), [3.5, 2.25], fn)
        with gzip.open(fn, "rt") as f:
            assert f.read() == orc.lowpass_text(spec, (12, 16), [3.5, 2.25], rt)
