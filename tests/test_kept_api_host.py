"""The kept C++ host headers compile and their host-only helpers behave as the
reference's (cuda/math.hpp:125-149 size policy, technicalities.hpp shape_tr, npy I/O).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def helper(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("hh") / "host_helpers")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "cpp", "host_helpers.cpp"), "-o", out], check=True)
    return out


def test_size_policy_and_layout(helper, tmp_path):
    npy = str(tmp_path / "a.npy")
    txt = subprocess.run([helper, npy], check=True, capture_output=True, text=True).stdout.splitlines()
    table = {int(l.split()[0]): (int(l.split()[1]), int(l.split()[2])) for l in txt[:13]}
    # factors235 / roundup235: prime factors above 5 are replaced by the next 2,3,5-smooth number
    assert table[1] == (1, 1) and table[5] == (1, 5) and table[7] == (0, 8) and table[11] == (0, 12)
    assert table[13] == (0, 15) and table[14] == (0, 16) and table[26] == (0, 30) and table[49] == (0, 64)
    assert table[97] == (0, 100) and table[120] == (1, 120) and table[625] == (1, 625) and table[1000] == (1, 1000)
    assert table[1023] == (0, 3 * 12 * 32)  # 1023 = 3 * 11 * 31 -> 3 * 12 * 32
    assert txt[13] == "shape_tr 6 5 4 | 5 4 1 | 4 1 1"
    # (1,0,0),(0,1,1): frame matrix {1, 0; 0, sqrt(2)} -> largest singular value sqrt(2)
    assert abs(float(txt[14].split()[1]) - np.sqrt(2)) < 1e-12
    assert txt[15].startswith("closest 1 0 ")
    a = np.load(npy)
    assert a.shape == (2, 3, 4) and a.dtype == np.float64 and a[0, 0, 0] == -3 and a[-1, -1, -1] == 8.5
    rest = [l for l in txt[16:] if not l.startswith('"init_pot_line"')]
    assert rest[1] == "npy 24 2 3 4 -3 8.5" and rest[2] == "npy dtype mismatch rejected"


def test_init_pot_line(helper):
    """cuda/potentials.hpp:5-20 on a 16 x 12 grid (h = 0.5, box 8 x 6, power 2, sigma 2, amp 1.5):
    x = i0 h - 4, y = i1 h - 3.  The square lattice of period sigma times a strip window in y:
    equal to amp w(y) on lattice points (x, y multiples of sigma), zero on the cell centres (both
    coordinates at half periods), (3/4)^power of that at a quarter period in x, periodic in x with
    period sigma, even in y; w(y) = (1 + tanh(y + sigma/2)) (1 + tanh(sigma/2 - y)) / 4."""
    out = subprocess.run([helper], check=True, capture_output=True, text=True).stdout.splitlines()
    assert '"init_pot_line" potential initialized' in out
    pot = np.array([float(v) for v in next(l for l in out if l.startswith("pot ")).split()[1:]]).reshape(12, 16)
    amp, sigma = 1.5, 2.0
    y = np.arange(12) * 0.5 - 3.0

    def w(yy):
        return (1 + np.tanh(yy + sigma / 2)) * (1 + np.tanh(sigma / 2 - yy)) / 4

    for i1 in (2, 6, 10):          # y = -2, 0, 2
        for i0 in (0, 4, 8, 12):   # x = -4, -2, 0, 2
            assert abs(pot[i1, i0] - amp * w(y[i1])) < 1e-14
    for i1 in (4, 8):              # y = -1, 1
        for i0 in (2, 6, 10, 14):  # x = -3, -1, 1, 3
            assert abs(pot[i1, i0]) < 1e-15
    assert abs(pot[6, 9] - amp * w(0.0) * 0.75 ** 2) < 1e-14  # x = 0.5: cos = 0
    assert np.abs(pot[:, :12] - pot[:, 4:]).max() < 1e-14
    assert np.abs(pot[1:12] - pot[1:12][::-1]).max() < 1e-14
    assert pot.min() >= 0 and abs(pot.max() - amp * w(0.0)) < 1e-14


def test_driver_builds_against_kept_headers():
    subprocess.run(["make", "-s", "driver"], cwd=REPO, check=True)
    for p in ("f32", "f64"):
        for d in ("kept_api_driver", "kept_api_exp_driver"):
            assert os.path.exists(os.path.join(REPO, "tests", "cpp", "bin", f"{d}_{p}"))
