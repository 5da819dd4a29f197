"""Drop-in check: the C++ driver written against the kept host API (tests/cpp/kept_api_driver.cpp)
reproduces the reference drivers' flows on the GPU, checked against numpy and the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def drv(prec):
    p = os.path.join(REPO, "tests", "cpp", "bin", f"kept_api_driver_{prec}")
    if not os.path.exists(p):
        pytest.fail(f"{p} not built (make driver)")
    return p


def run(prec, *args):
    r = subprocess.run([drv(prec)] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("prec,dt,tol", [("f32", np.float32, 1e-5), ("f64", np.float64, 2e-14)])
@pytest.mark.parametrize("shp", [(36, 75, 60), (625, 10, 9), (27, 36, 27), (1, 64, 48)])
def test_fft_flow(tmp_path, prec, dt, tol, shp):
    x = np.random.default_rng(3).standard_normal(shp).astype(dt)
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, x)
    dims = [s for s in shp if True]
    n0, n1, n2 = shp[2], shp[1], shp[0]  # my_unit_test.py:49 passes the axes reversed
    if shp[0] == 1:
        np.save(fin, x[0])
    run(prec, "fft", n0, n1, n2, fin, fout, 0)
    y = np.load(fout)
    want = np.fft.rfftn(x.astype(np.float64)).reshape(y.shape)
    assert np.abs(want - y).max() / np.sqrt(x.size) < tol
    run(prec, "fft", n0, n1, n2, fin, fout, 1)
    z = np.load(fout).reshape(x.shape)
    assert np.abs(x - z / x.size).max() < tol
    del dims


def test_noise_flow(tmp_path, orc):
    out = str(tmp_path / "xi.npy")
    run("f64", "noise", 10, 8, 6, 99, out)
    assert np.abs(np.load(out) - orc.kspace_noise((6, 8, 10), 99, 0)).max() < 1e-13


@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 2e-5)])
@pytest.mark.parametrize("shp", [(16, 16, 16), (32, 48)])
def test_pfc_simulator_loop(tmp_path, orc, prec, tol, shp):
    """c2r -> psi3_split(norm) -> r2c -> fill_kspace -> update_eom with the reference split and noise."""
    psi = orc.init_field(shp, -0.2, 0.1, 7)
    dt = np.float64 if prec == "f64" else np.float32
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, psi.astype(dt))
    steps, eps, dtt, namp, seed, h = 10, 0.3, 0.25, 0.05, 5, np.pi / 4
    ext = list(shp[::-1]) + [1] * (3 - len(shp))
    run(prec, "pfc", *ext, fin, fout, steps, eps, dtt, namp, seed, h)
    got = np.load(fout)
    # the same loop on the oracle (eom.hpp:60-72 noise scale dt*namp*sqrtf(N)*|k|)
    N = psi.size
    lens = [e * h for e in ext]
    lens = [l if e > 1 else 1.0 for l, e in zip(lens, ext)]
    kp = orc.r2c(psi.astype(dt).astype(np.float64))
    for s in range(steps):
        r = orc.c2r(kp, shp) / N
        kn = orc.r2c(r ** 3 - r)
        xi = orc.kspace_noise(shp, seed, s)
        kp = orc.update(kp, kn, xi, shp, lens, dtt, eps, dtt * namp * float(np.sqrt(np.float32(N))), 1)
    want = orc.c2r(kp, shp) / N
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


def test_peaks_and_bond_order_flow(tmp_path, orc):
    n = 48
    x = np.arange(n) * 2 * np.pi / 12
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rng = np.random.default_rng(2)
    f = np.cos(X) * np.cos(Y) + np.cos(Y) * np.cos(Z) + np.cos(Z) * np.cos(X) + 0.01 * rng.standard_normal(X.shape)
    fin = str(tmp_path / "f.npy")
    np.save(fin, f)
    pk, bo = str(tmp_path / "peaks.txt"), str(tmp_path / "bo.txt")
    h = 0.5
    r = run("f64", "peaks", n, n, n, fin, pk, bo, h, 0.5, 5.5)
    idx = orc.findpeaks(f, 0.5)
    assert f"found {idx.size} maxima" in r.stderr
    rows = np.loadtxt(pk)
    i0, i1, i2 = idx % n, (idx // n) % n, idx // (n * n)
    assert np.allclose(rows[:, 0], i0 * h, atol=1e-6) and np.allclose(rows[:, 2], i2 * h, atol=1e-6)
    assert np.allclose(rows[:, 3], f.reshape(-1)[idx], atol=1e-6)
    table = np.loadtxt(bo, comments="#")
    keep = f.reshape(-1)[idx] > 0.5
    xyz = np.column_stack([i0, i1, i2])[keep] * h
    nb, out = orc.bond_order_cutoff(xyz, n * h, 5.5)
    assert np.all(nb == 8)  # bcc: the 8 nearest sites at sqrt(3)/2 * 12 cells
    m = nb > 0
    want = np.column_stack([xyz[m], nb[m], out[m][:, 8], out[m][:, :8]])
    assert table.shape == want.shape
    assert np.array_equal(table[:, 3:5], want[:, 3:5])
    assert np.abs(table - want).max() <= 6e-7


# call_kernel_calc_fen returns float in both builds (eom.hpp:130), so the FP64 flow is exact to float rounding
@pytest.mark.parametrize("prec,dt,tol", [("f32", np.float32, 1e-5), ("f64", np.float64, 1.2e-7)])
def test_free_energy_flow(orc, tmp_path, prec, dt, tol):
    # update_eom_with_fen -> c2r -> norm -> calc_fen (eom.hpp:88-141) gives the free energy functional
    shp, eps, h = (16, 24, 32), 0.25, 0.7
    psi = orc.init_field(shp, -0.2, 0.2, 6).astype(dt)
    fin = str(tmp_path / "in.npy")
    np.save(fin, psi)
    r = run(prec, "fen", shp[2], shp[1], shp[0], fin, eps, h)
    got = float(r.stdout.split()[-1])
    want = orc.free_energy(psi.astype(np.float64), [h] * 3, eps)
    assert abs(got - want) < tol * abs(want)


@pytest.mark.parametrize("prec,dt,tol", [("f32", np.float32, 1e-4), ("f64", np.float64, 1e-11)])
def test_relaxation_flow(orc, tmp_path, prec, dt, tol):
    # c2r -> psi3_split(norm) -> r2c -> update_el: the relaxation of el.hpp:27-45, its convergence measure
    shp, eps, dtau, steps = (24, 32), 0.3, 0.2, 30
    psi = orc.init_field(shp, -0.3, 0.2, 2).astype(dt)
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, psi)
    r = run(prec, "relax", shp[1], shp[0], 1, fin, fout, steps, eps, dtau)
    maxes = np.array([float(v) for v in r.stdout.split()])
    h = np.pi / 4
    lens = [shp[1] * h, shp[0] * h, 1.0]
    p = psi.astype(np.float64)
    want_m = []
    for _ in range(steps):
        kp = np.fft.rfftn(p)
        kp, m = orc.update_el(kp, np.fft.rfftn(p ** 3 - p), shp, lens, dtau, eps)
        p = np.fft.irfftn(kp, s=shp, axes=(0, 1))
        want_m.append(m)
    assert maxes.size == steps
    assert np.abs(maxes - want_m).max() < tol * max(want_m) * 10
    got = np.load(fout)
    assert np.abs(got - p).max() < tol * np.abs(p).max() * 10
    assert maxes[-1] < maxes[0]


def exp_drv(prec):
    p = os.path.join(REPO, "tests", "cpp", "bin", f"kept_api_exp_driver_{prec}")
    if not os.path.exists(p):
        pytest.fail(f"{p} not built (make driver)")
    return p


def run_exp(prec, *args):
    r = subprocess.run([exp_drv(prec)] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


# the ETD tables take expf in both builds (eom-exp.hpp:7, el-exp.hpp:6), and the device's and the host's
# expf may differ by an ulp of float, so the FP64 build agrees with the oracle to float rounding of the tables
@pytest.mark.parametrize("prec,dt,tol", [("f32", np.float32, 2e-5), ("f64", np.float64, 1e-7)])
@pytest.mark.parametrize("shp", [(12, 16, 20), (24, 32)])
def test_etd_flow_through_kept_exp_headers(orc, tmp_path, prec, dt, tol, shp):
    """eom-exp.hpp: init_eom_op once, then c2r -> psi3_split(norm) -> r2c -> update_eom (ETD) per step"""
    psi = orc.init_field(shp, -0.25, 0.1, 4).astype(dt)
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, psi)
    ext = list(shp[::-1]) + [1] * (3 - len(shp))
    steps, eps, dtt, h = 12, 0.3, 0.5, np.pi / 4
    run_exp(prec, "etd", *ext, fin, fout, steps, eps, dtt)
    lens = [e * h if e > 1 else 1.0 for e in ext]
    kop = orc.etd_op(shp, lens, dtt, eps, 0)
    N = psi.size
    kp = orc.r2c(psi.astype(np.float64))
    for _ in range(steps):
        r = orc.c2r(kp, shp) / N
        kp = orc.update_eom_etd(kop, kp, orc.r2c(r ** 3 - r), None, shp, lens, 0.0)
    want = orc.c2r(kp, shp) / N
    got = np.load(fout)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


@pytest.mark.parametrize("prec,dt,tol", [("f32", np.float32, 1e-4), ("f64", np.float64, 1e-7)])
def test_el_exp_flow_through_kept_exp_headers(orc, tmp_path, prec, dt, tol):
    """el-exp.hpp: init_el_op, then the non-conserved ETD relaxation and its max |change| per step"""
    shp = (20, 24, 16)
    psi = orc.init_field(shp, -0.3, 0.2, 9).astype(dt)
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, psi)
    ext = list(shp[::-1])
    steps, eps, dtt, h = 10, 0.3, 0.3, np.pi / 4
    r = run_exp(prec, "elexp", *ext, fin, fout, steps, eps, dtt)
    maxes = np.array([float(v) for v in r.stdout.split()])
    lens = [e * h for e in ext]
    kop = orc.etd_op(shp, lens, dtt, eps, 1)
    N = psi.size
    kp = orc.r2c(psi.astype(np.float64))
    want_m = []
    for _ in range(steps):
        p = orc.c2r(kp, shp) / N
        kp, m = orc.update_el_etd(kop, kp, orc.r2c(p ** 3 - p), shp)
        want_m.append(m)
    assert maxes.size == steps and np.abs(maxes - want_m).max() < tol * max(want_m) * 10
    want = orc.c2r(kp, shp) / N
    assert np.linalg.norm(np.load(fout) - want) / np.linalg.norm(want) < tol


def test_findpeaks_2d_flow(orc, tmp_path):
    """call_kernel_findpeaks_2d (findpeaks_2d.hpp:29-56): 8-neighbour maxima as "%f %f %f %f" lines"""
    n1, n0 = 48, 64
    y = np.arange(n1)[:, None] * 2 * np.pi / 8
    x = np.arange(n0)[None, :] * 2 * np.pi / 8
    f = np.cos(x) + np.cos(y) + 0.01 * np.random.default_rng(3).standard_normal((n1, n0))
    fin, pk = str(tmp_path / "f.npy"), str(tmp_path / "peaks.txt")
    np.save(fin, f)
    h = 0.5
    r = run_exp("f64", "peaks2d", n0, n1, fin, pk, h, 0.5)
    idx = orc.findpeaks(f, 0.5)
    assert idx.size > 0 and f"found {idx.size} maxima" in r.stderr
    rows = np.loadtxt(pk, ndmin=2)
    assert rows.shape == (idx.size, 4)
    assert np.allclose(rows[:, 0], (idx % n0) * h, atol=1e-6) and np.allclose(rows[:, 1], (idx // n0) * h, atol=1e-6)
    assert np.all(rows[:, 2] == 0) and np.allclose(rows[:, 3], f.reshape(-1)[idx], atol=1e-6)


@pytest.mark.parametrize("prec,dt", [("f32", np.float32), ("f64", np.float64)])
def test_potential_into_psi3_split(tmp_path, prec, dt):
    """init_pot_line (potentials.hpp:5-20) uploaded and added by psi3_split (psi3_split.hpp:1-6):
    a <- a^3 - a - mu + pot with mu = 0.1"""
    n1, n0, h, power, sigma, amp = 24, 32, 0.5, 2, 2.0, 1.5
    psi = np.random.default_rng(4).uniform(-0.5, 0.5, (n1, n0)).astype(dt)
    fin, fout = str(tmp_path / "in.npy"), str(tmp_path / "out.npy")
    np.save(fin, psi)
    run_exp(prec, "pot", n0, n1, fin, fout, h, power, sigma, amp)
    got = np.load(fout).astype(np.float64)
    x = np.arange(n0)[None, :] * h - n0 * h / 2
    yy = np.arange(n1)[:, None] * h - n1 * h / 2
    f = 2 * np.pi / sigma
    w = (1 + np.tanh(yy + sigma / 2)) * (1 + np.tanh(sigma / 2 - yy)) / 4
    pot = amp * ((np.cos(f * x) + np.cos(f * yy) + 2) / 4) ** power * w
    p = psi.astype(np.float64)
    want = p ** 3 - p - 0.1 + pot.astype(dt).astype(np.float64)
    assert np.abs(got - want).max() < (2e-6 if dt == np.float32 else 1e-14)
