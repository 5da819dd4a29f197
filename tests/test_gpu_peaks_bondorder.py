"""GPU parity of the peak finder and the bond-order analysis against the oracle
(bit-exact peak lists and neighbour sets) and the reference's own build."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shp", [(7, 9, 11), (64, 64, 64), (12, 10), (256, 256), (2, 2, 2), (1, 4, 4),
                                 (33, 70, 130), (2, 3, 200), (40, 17, 65), (5, 1, 64), (3, 50, 1), (70, 33)])
def test_peaks_match_oracle(mcs, orc, shp, dt):
    v = np.random.default_rng(sum(shp)).standard_normal(shp).astype(dt)
    for ths in (-10.0, 0.5):
        ths_f = float(np.asarray(ths, dt))
        idx, vals = mcs.findpeaks(v, ths_f, with_values=True)
        want = orc.findpeaks(v.astype(np.float64), ths_f)
        assert np.array_equal(idx, want)
        assert np.array_equal(vals, v.reshape(-1)[want].astype(np.float64))


def test_peaks_special_cases(mcs):
    v = np.zeros((6, 6, 6), np.float32)
    assert mcs.findpeaks(v, -1).size == 0
    v[5, 5, 5] = 1.0
    assert list(mcs.findpeaks(v, 1.0)) == [215]
    assert mcs.findpeaks(v, 1.0001).size == 0
    c = np.zeros((8, 8), np.float64)
    c[::2, ::2] = 1.0
    assert mcs.findpeaks(c, 0.5).size == 16
    with pytest.raises(mcs.McsError) as e:
        mcs.findpeaks(c, 0.5, cap=3)
    assert e.value.code == mcs.ERR_CAPACITY


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_peaks_with_nan_neighbours(mcs, orc, dt):
    # a NaN is never a peak and a cell next to one is not either (v > NaN is false), as in the
    # reference's per-neighbour test
    v = np.random.default_rng(5).standard_normal((20, 36, 70)).astype(dt)
    v.reshape(-1)[np.random.default_rng(6).choice(v.size, 300, replace=False)] = np.nan
    got = mcs.findpeaks(v, -10.0)
    assert np.array_equal(got, orc.findpeaks(v.astype(np.float64), -10.0))


def test_peaks_full_size_repeated_calls(mcs, orc):
    # the workspace kept between calls: a large grid, then a smaller one, then the large one again
    big = np.random.default_rng(1).standard_normal((96, 128, 256)).astype(np.float32)
    small = np.random.default_rng(2).standard_normal((9, 40, 70)).astype(np.float32)
    want_big = orc.findpeaks(big.astype(np.float64), 0.3)
    assert np.array_equal(mcs.findpeaks(big, 0.3), want_big)
    assert np.array_equal(mcs.findpeaks(small, 0.3), orc.findpeaks(small.astype(np.float64), 0.3))
    assert np.array_equal(mcs.findpeaks(big, 0.3), want_big)


def test_peaks_on_crystal_field(mcs, orc):
    # a bcc-like density wave: one peak per lattice site, all on the grid
    n = 48
    x = np.arange(n) * 2 * np.pi / 12
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.cos(X) * np.cos(Y) + np.cos(Y) * np.cos(Z) + np.cos(Z) * np.cos(X)
    idx = mcs.findpeaks(f.astype(np.float32), 0.0)
    assert np.array_equal(idx, orc.findpeaks(f.astype(np.float32).astype(np.float64), 0.0))
    assert idx.size > 0


def test_bond_order_cutoff_matches_oracle_and_reference(mcs, orc):
    g = np.load(os.path.join(GOLD, "bond_order_ref.npz"))
    for c in range(int(g["count"])):
        pts, box, rlim, table = g[f"pts{c}"], float(g[f"box{c}"]), float(g[f"rlim{c}"]), g[f"table{c}"]
        nb, out = mcs.bond_order_cutoff(pts, box, rlim)
        onb, oout = orc.bond_order_cutoff(pts, box, rlim)
        assert np.array_equal(nb, onb)
        assert np.array_equal(np.isnan(out), np.isnan(oout))
        assert np.nanmax(np.abs(out - oout)) < 1e-12
        m = nb > 0
        mine = np.column_stack([pts[m], nb[m], out[m][:, 8], out[m][:, :8]])
        assert np.array_equal(mine[:, :5], table[:, :5])
        assert np.nanmax(np.abs(mine[:, 5:] - table[:, 5:])) <= 6e-7


def test_bond_order_knn_bit_exact_neighbours(mcs, orc):
    rng = np.random.default_rng(8)
    n = 40
    cells = rng.choice(n ** 3, 2000, replace=False)
    ijk = np.column_stack([cells % n, (cells // n) % n, cells // (n * n)]).astype(np.int32)
    nbr, out, h4, h6 = mcs.bond_order_knn(ijk, n, np.pi / 4, 12)
    onbr, oout = orc.bond_order_knn(ijk, n, np.pi / 4, 12)
    assert np.array_equal(nbr, onbr)
    assert np.abs(out - oout).max() < 1e-12
    want4 = np.bincount(np.minimum((oout[:, 0] * 64).astype(int), 63), minlength=64)
    want6 = np.bincount(np.minimum((oout[:, 1] * 64).astype(int), 63), minlength=64)
    assert np.array_equal(h4, want4) and np.array_equal(h6, want6)


def _ijk(cells, n):
    return np.column_stack([cells % n, (cells // n) % n, cells // (n * n)]).astype(np.int32)


def _knn_matches_oracle(mcs, orc, ijk, n, h, k):
    nbr, out, h4, h6 = mcs.bond_order_knn(ijk, n, h, k)
    onbr, oout = orc.bond_order_knn(ijk, n, h, k)
    assert np.array_equal(nbr, onbr)
    assert np.array_equal(np.isnan(out), np.isnan(oout))
    if out.size and not np.isnan(out).all():
        assert np.nanmax(np.abs(out - oout)) < 1e-12
    m = ~np.isnan(oout[:, 0])
    assert h4.sum() == m.sum() and h6.sum() == m.sum()
    return nbr, out


@pytest.mark.parametrize("k", [1, 8, 12, 13, 16, 17, 32])
def test_bond_order_knn_cell_list_every_capacity(mcs, orc, k):
    # every register capacity of the cell-list search (8, 12, 16, 32 slots) against the all-pairs oracle
    rng = np.random.default_rng(100 + k)
    n = 48
    _knn_matches_oracle(mcs, orc, _ijk(rng.choice(n ** 3, 3000, replace=False), n), n, 0.7, k)


def test_bond_order_knn_cell_list_edge_cases(mcs, orc):
    rng = np.random.default_rng(3)
    # one point: no neighbours (-1 and NaN); fewer points than k: every other point, in key order
    nbr, out = _knn_matches_oracle(mcs, orc, np.array([[3, 4, 5]], np.int32), 16, 1.0, 12)
    assert (nbr == -1).all() and np.isnan(out).all()
    nbr, _ = _knn_matches_oracle(mcs, orc, _ijk(rng.choice(16 ** 3, 6, replace=False), 16), 16, 1.0, 12)
    assert (nbr[:, :5] >= 0).all() and (nbr[:, 5:] == -1).all()
    # a dense cluster straddling the periodic corner of a large, otherwise empty box
    n = 256
    c = rng.integers(-4, 4, size=(400, 3))
    c = np.unique(c, axis=0)
    _knn_matches_oracle(mcs, orc, (c % n).astype(np.int32), n, 0.5, 12)
    # the same points given outside [0, n): coordinates are taken modulo n
    _knn_matches_oracle(mcs, orc, c.astype(np.int32) + np.array([n, -n, 3 * n], np.int32), n, 0.5, 12)
    # uneven density: a dense slab and a sparse remainder
    n = 64
    dense = rng.choice(n * n * 8, 2500, replace=False)
    sparse = n * n * 8 + rng.choice(n * n * (n - 8), 300, replace=False)
    _knn_matches_oracle(mcs, orc, _ijk(np.concatenate([sparse, dense]), n), n, 1.0, 12)
    # a simple-cubic lattice: 6 nearest at a, 12 at a*sqrt(2), so the 12-NN set is decided by index ties
    a, m = 5, 8
    g = np.arange(m) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(np.int32)
    perm = rng.permutation(len(lat))
    _knn_matches_oracle(mcs, orc, lat[perm], a * m, 1.0, 12)
    # a box smaller than a cell in every direction and shells that wrap around it
    _knn_matches_oracle(mcs, orc, _ijk(rng.choice(5 ** 3, 60, replace=False), 5), 5, 1.0, 12)


def test_bond_order_fcc_lattice(mcs):
    pts = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                for (x, y, z) in [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]:
                    pts.append(((2 * i + x) % 8, (2 * j + y) % 8, (2 * k + z) % 8))
    nbr, out, h4, h6 = mcs.bond_order_knn(np.array(pts), 8, 1.0, 12)
    assert np.allclose(out[:, 0], 0.19094, atol=5e-5) and np.allclose(out[:, 1], 0.57452, atol=5e-5)
    assert h4[int(0.19094 * 64)] == len(pts) and h6[int(0.57452 * 64)] == len(pts)


def _cutoff_matches_oracle(mcs, orc, pts, box, rlim):
    nb, out = mcs.bond_order_cutoff(pts, box, rlim)
    onb, oout = orc.bond_order_cutoff(pts, box, rlim)
    assert np.array_equal(nb, onb)
    assert np.array_equal(np.isnan(out), np.isnan(oout))
    if not np.isnan(oout).all():
        assert np.nanmax(np.abs(out - oout)) < 1e-12
    return nb, out


def test_bond_order_cutoff_cell_list(mcs, orc):
    # inputs that take the cell list (>= 256 points in the box, >= 3 cells of width rlim per axis)
    rng = np.random.default_rng(21)
    box = 40.0
    _cutoff_matches_oracle(mcs, orc, rng.uniform(0, box, size=(3000, 3)), box, 3.0)
    # a simple cubic lattice with rlim equal to the spacing: neighbours at exactly rlim count (r <= rlim)
    a, m = 2.5, 10
    g = np.arange(m) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])[rng.permutation(m ** 3)]
    nb, _ = _cutoff_matches_oracle(mcs, orc, lat, a * m, a)
    assert (nb == 6).all()
    # fcc with rlim between the first and second shells: 12 neighbours, q6 of the closed form
    fcc = []
    for i in range(7):
        for j in range(7):
            for k in range(7):
                for (x, y, z) in [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]:
                    fcc.append((i + x, j + y, k + z))
    fcc = np.array(fcc) * 2.0
    nb, out = _cutoff_matches_oracle(mcs, orc, fcc, 14.0, 1.6)
    assert (nb == 12).all() and np.allclose(out[:, 1], 0.57452, atol=5e-5)
    # points on cell faces and at the periodic boundary
    edge = rng.uniform(0, box, size=(600, 3))
    edge[:200, 0] = np.round(edge[:200, 0] / (box / 13)) * (box / 13) % box
    edge[200:300, 1] = 0.0
    edge[300:400, 2] = np.nextafter(box, 0)
    _cutoff_matches_oracle(mcs, orc, edge, box, 3.0)
    # a dense cluster (more than 64 neighbours) falls back to the all-pairs scan
    dense = np.concatenate([rng.uniform(0, box, size=(400, 3)), 20 + rng.uniform(0, 1.5, size=(200, 3))])
    nb, _ = _cutoff_matches_oracle(mcs, orc, dense, box, 3.0)
    assert nb.max() > 64
