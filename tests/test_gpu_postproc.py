"""Post-processing on the GPU against the oracle: the Gaussian filter table, the spectral
products, squaring, the 2D low-pass amplitude with its gz text dump (Python binding and the
kept C++ header), and minimum-image pair distances and their histogram."""
import gzip
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("shp", [(12, 20, 32), (30, 48), (5, 7, 9)])
def test_gaussian_table(mcs, orc, shp):
    nd = len(shp)
    h = [0.7, 0.9, 1.3][:nd] + [1.0] * (3 - nd)
    ext = list(shp[::-1]) + [1] * (3 - nd)
    lens = [e * hh for e, hh in zip(ext, h)]
    want = orc.gaussian_init(shp, lens, h, 1.7)
    got64 = mcs.gaussian_init(shp, lens, h, 1.7, np.float64)
    assert np.abs(got64 - want).max() < 4e-16
    got32 = mcs.gaussian_init(shp, lens, h, 1.7, np.float32)
    assert got32.dtype == np.complex64
    assert np.abs(got32 - want).max() < 1e-6
    assert got32[(0,) * nd] == 1.0 + 0.0j


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_kmult_and_square_are_exact(mcs, dt):
    shp = (6, 10, 14)
    ct = np.complex64 if dt == np.float32 else np.complex128
    rng = np.random.default_rng(4)
    hs = (6, 10, 8)
    k1 = (rng.standard_normal(hs) + 1j * rng.standard_normal(hs)).astype(ct)
    k2 = (rng.standard_normal(hs) + 1j * rng.standard_normal(hs)).astype(ct)
    for part, f in [(0, k2.real), (1, k2.imag)]:
        got = mcs.kmult(k1, k2, shp, part)
        assert np.array_equal(got.real, k1.real * f) and np.array_equal(got.imag, k1.imag * f)
    a = rng.standard_normal(shp).astype(dt)
    assert np.array_equal(mcs.square(a), a * a)


def cosine_field(nx=64, ny=32, m=3, a=0.4):
    h = np.pi / 4
    x = np.arange(nx) * h
    psi = np.tile(a * np.cos(2 * np.pi * m / (nx * h) * x), (ny, 1))
    rng = np.random.default_rng(8)
    return psi + 0.05 * rng.standard_normal(psi.shape), [nx * h, ny * h, h], [h, h, h]


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_amplitude_lowpass(mcs, orc, dt, tol):
    psi, lens, h = cosine_field()
    psi = psi.astype(dt)
    want_k, want = orc.amplitude_lowpass(psi.astype(np.float64), lens, h, 0.8)
    got_k, got = mcs.amplitude_lowpass(psi, lens, h, 0.8)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
    assert np.abs(got_k - want_k).max() < tol * np.abs(want_k).max()


def parse_lowpass(txt):
    lines = txt.split("\n")
    assert lines[0] == "LOWPASS"
    nh = int(lines[1])
    rest = "\n".join(lines[2:])
    head, body = rest[:nh], rest[nh:]
    rows = [ln.split() for ln in body.split("\n") if ln]
    return head, {(int(r[0]), int(r[1])): complex(float(r[2]), float(r[3])) for r in rows}


def compare_dumps(got_txt, want_txt, want_k, arel):
    """same header; same kept modes except those within the value error of the threshold;
    values within that error (arel * the largest mode) plus the rounding of 6 printed digits"""
    gh, gm = parse_lowpass(got_txt)
    wh, wm = parse_lowpass(want_txt)
    assert gh == wh
    atol = arel * np.abs(want_k).max()
    near = {(i0, i1) for (i1, i0), v in np.ndenumerate(np.abs(want_k)) if abs(v - 1e-3) < 2 * atol + 1e-8}
    assert set(gm) - near == set(wm) - near
    assert len(set(gm) & set(wm)) > 10
    for key in set(gm) & set(wm):
        for g, w in [(gm[key].real, wm[key].real), (gm[key].imag, wm[key].imag)]:
            assert abs(g - w) <= atol + 1.1e-5 * abs(w), (key, g, w)


@pytest.mark.parametrize("dt,arel", [(np.float64, 1e-11), (np.float32, 5e-6)])
def test_lowpass_dump_from_python(mcs, orc, tmp_path, dt, arel):
    psi, lens, h = cosine_field()
    k, _ = mcs.amplitude_lowpass(psi.astype(dt), lens, h, 0.8)
    fn = str(tmp_path / "lp.gz")
    mcs.write_lowpass_2d(k, psi.shape, lens[:2], fn)
    with gzip.open(fn, "rt") as f:
        got = f.read()
    want_k, _ = orc.amplitude_lowpass(psi.astype(dt).astype(np.float64), lens, h, 0.8)
    compare_dumps(got, orc.lowpass_text(want_k, psi.shape, lens[:2], dt), want_k, arel)


@pytest.mark.parametrize("prec,dt,tol,arel", [("f64", np.float64, 1e-12, 1e-11), ("f32", np.float32, 2e-5, 5e-6)])
def test_lowpass_kept_header_flow(orc, tmp_path, prec, dt, tol, arel):
    """call_kernel_gaussian_init -> r2c -> calc_amplitude_lowpass through the kept C++ header"""
    exe = os.path.join(REPO, "tests", "cpp", "bin", f"kept_api_driver_{prec}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (make driver)")
    psi, lens, h = cosine_field(nx=48, ny=40, m=2)
    psi = psi.astype(dt)
    fin, fgz, famp = str(tmp_path / "in.npy"), str(tmp_path / "lp.gz"), str(tmp_path / "amp.npy")
    np.save(fin, psi)
    r = subprocess.run([exe, "lowpass", "48", "40", "1", fin, fgz, famp, repr(h[0]), "0.8"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # FP32 runs see the box lengths and spacing as floats
    hh = float(dt(h[0]))
    lens_f = [float(dt(48 * dt(hh))), float(dt(40 * dt(hh))), float(dt(1 * dt(hh)))]
    want_k, want = orc.amplitude_lowpass(psi.astype(np.float64), lens_f, [hh] * 3, float(dt(0.8)))
    got = np.load(famp).reshape(psi.shape)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
    with gzip.open(fgz, "rt") as f:
        txt = f.read()
    compare_dumps(txt, orc.lowpass_text(want_k, psi.shape, lens_f[:2], dt), want_k, arel)


@pytest.mark.parametrize("P,box", [(2, 10.0), (3, 4.0), (50, 17.5), (257, 9.0)])
def test_pair_distances_bit_exact(mcs, orc, P, box):
    rng = np.random.default_rng(P)
    xyz = rng.uniform(0, box, (P, 3))
    if P >= 2:
        xyz[1] = (xyz[0] + [box / 2, 0, 0]) % box  # minimum-image tie
    got = mcs.pair_distances(xyz, box)
    want = orc.pair_distances(xyz, box)
    assert got.shape == want.shape == (P * (P - 1),)
    assert np.array_equal(got, want)


def test_pair_distances_edges(mcs):
    assert mcs.pair_distances(np.zeros((0, 3)), 1.0).size == 0
    assert mcs.pair_distances(np.zeros((1, 3)), 1.0).size == 0
    with pytest.raises(mcs.McsError) as e:
        mcs.pair_distances(np.zeros((2, 3)), 0.0)
    assert e.value.code == mcs.ERR_INVALID
    out = np.zeros(1)
    cnt = mcs.ctypes.c_int64()
    xyz = np.zeros((3, 3))
    rc = mcs.lib().mcs_pair_distances(xyz.ctypes.data_as(mcs._dp), 3, 1.0, out.ctypes.data_as(mcs._dp), 5,
                                      mcs.ctypes.byref(cnt))
    assert rc == mcs.ERR_CAPACITY and cnt.value == 6


@pytest.mark.parametrize("P,nbins", [(2, 8), (300, 64), (1001, 100)])
def test_pair_histogram_matches_oracle(mcs, orc, P, nbins):
    rng = np.random.default_rng(3 + P)
    box = 12.0
    xyz = rng.uniform(0, box, (P, 3))
    got = mcs.pair_histogram(xyz, box, 5.0, nbins)
    assert np.array_equal(got, orc.pair_histogram(xyz, box, 5.0, nbins))


def test_pair_histogram_simple_cubic_shells_at_size(mcs):
    """64000 lattice points (2.0e9 pairs): the first four shells hold 3P, 6P, 4P and 3P pairs"""
    m = 40
    g = np.arange(m, dtype=np.float64)
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    P = len(xyz)
    # bins of 0.055, every shell inside a bin: 1 -> 18, 1.414 -> 25, 1.732 -> 31, 2 -> 36 (sqrt 5 > 2.2)
    hist = mcs.pair_histogram(xyz, float(m), 2.2, 40)
    want = np.zeros(40, np.int64)
    want[[18, 25, 31, 36]] = [3 * P, 6 * P, 4 * P, 3 * P]
    assert np.array_equal(hist, want)
