"""Writes the golden fixtures under tests/golden/ (run from the repo root).

  fft_small.npz        numpy.fft.rfftn of seeded inputs on small shapes (odd and even x)
  fft_refshapes.npz    the FFT shapes of cufft_unittest/my_unit_test.py:74-77 and 161-163:
                       seed, shape, 64 sampled output entries and the output's norm
  bond_order_ref.npz   outputs of the reference's own bond-order analysis
                       (bond-order-analysis/standalone.cpp + bond-order.cpp, built by
                       `make -C oracle ref` from the mcstar-libs sources) on seeded point sets

numpy (pocketfft) is the library the reference's own FFT test compares against
(cufft_unittest/my_unit_test.py:138).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

SMALL = [(9,), (8,), (7, 5), (6, 10), (5, 4, 9), (4, 6, 10), (3, 5, 7), (6, 6, 8), (2, 3, 2), (1, 1, 1), (12, 1, 6)]
# my_unit_test.py:74-76 passes (l3, l2, l1) as (n0, n1, n2), so the numpy shape is (l1, l2, l3)
L1 = [36, 20, 625, 30, 30, 27, 24, 60, 27, 10]
L2 = [75, 20, 10, 40, 10, 36, 6, 30, 100, 18]
L3 = [60, 45, 9, 16, 25, 27, 150, 150, 6, 75]
REF3D = list(zip(L1, L2, L3))
REF2D = [(1200, 1200), (1280, 1280), (1600, 1600), (1152, 1152)]  # my_unit_test.py:161-163


def fft_small():
    out = {}
    for i, shp in enumerate(SMALL):
        x = np.random.default_rng(100 + i).standard_normal(shp)
        out[f"x{i}"] = x
        out[f"k{i}"] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_small.npz"), **out)


def fft_refshapes():
    out = {}
    shapes = REF3D + REF2D
    for i, shp in enumerate(shapes):
        seed = 1000 + i
        x = np.random.default_rng(seed).standard_normal(shp)
        k = np.fft.rfftn(x)
        sel = np.random.default_rng(seed + 1).integers(0, k.size, 64)
        out[f"shape{i}"] = np.array(shp)
        out[f"seed{i}"] = np.array(seed)
        out[f"sel{i}"] = sel
        out[f"val{i}"] = k.reshape(-1)[sel]
        out[f"norm{i}"] = np.array(np.linalg.norm(k))
    out["count"] = np.array(len(shapes))
    np.savez_compressed(os.path.join(HERE, "fft_refshapes.npz"), **out)


def bond_order_ref():
    exe = os.path.join(REPO, "oracle", "_ref", "bond_order_standalone")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    out = {}
    cases = []
    rng = np.random.default_rng(3)
    # jittered bcc on an integer grid (box 40, rlim 7.5), random gas (box 30, rlim 6), fcc (box 24, rlim 6)
    n = 40
    pts = []
    for i in range(0, n, 8):
        for j in range(0, n, 8):
            for k in range(0, n, 8):
                pts += [(i, j, k), (i + 4, j + 4, k + 4)]
    pts = (np.array(pts) + rng.integers(-1, 2, size=(len(pts), 3))) % n
    cases.append((pts, n, 7.5))
    cases.append((rng.integers(0, 30, size=(300, 3)), 30, 6.0))
    fcc = []
    for i in range(0, 24, 4):
        for j in range(0, 24, 4):
            for k in range(0, 24, 4):
                fcc += [(i, j, k), (i + 2, j + 2, k), (i + 2, j, k + 2), (i, j + 2, k + 2)]
    cases.append((np.array(fcc), 24, 3.0))
    for c, (p, box, rlim) in enumerate(cases):
        txt = "h\nh\n" + "".join("%d %d %d %d\n" % (t, *q) for t, q in enumerate(p.astype(int)))
        r = subprocess.run([exe, str(box), str(rlim)], input=txt, capture_output=True, text=True, check=True)
        rows = [[float(v) for v in l.split()] for l in r.stdout.splitlines() if l and not l.startswith("#")]
        out[f"pts{c}"] = p.astype(np.float64)
        out[f"box{c}"] = np.array(float(box))
        out[f"rlim{c}"] = np.array(float(rlim))
        out[f"table{c}"] = np.array(rows)
    out["count"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "bond_order_ref.npz"), **out)


if __name__ == "__main__":
    fft_small()
    fft_refshapes()
    bond_order_ref()
    print("written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")), file=sys.stderr)
