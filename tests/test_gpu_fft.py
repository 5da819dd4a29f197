"""GPU FFT parity: the library's own R2C/C2R against numpy.fft and the oracle,
on the shapes and tolerances of cufft_unittest/my_unit_test.py."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")

# my_unit_test.py:74-76 (numpy order) and 161-163
L1 = [36, 20, 625, 30, 30, 27, 24, 60, 27, 10]
L2 = [75, 20, 10, 40, 10, 36, 6, 30, 100, 18]
L3 = [60, 45, 9, 16, 25, 27, 150, 150, 6, 75]
REF3D = list(zip(L1, L2, L3))
REF2D = [(1200, 1200), (1280, 1280), (1600, 1600), (1152, 1152)]
TOL = {np.float32: 1e-5, np.float64: 1e-14}


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shp", REF3D)
def test_reference_shapes_forward_and_round_trip(mcs, shp, dt):
    x = np.random.default_rng(sum(shp)).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    want = np.fft.rfftn(x.astype(np.float64))
    n = x.size
    # TestB metric: max |rfftn(x)/sqrt(N) - y/sqrt(N)| (my_unit_test.py:138-142)
    assert np.abs(want - k).max() / np.sqrt(n) < TOL[dt] * (2 if dt == np.float64 else 1)
    y = mcs.irfftn(k, shp)
    # TestA metric: max |x - y/N| (my_unit_test.py:95-99)
    assert np.abs(x.astype(np.float64) - y.astype(np.float64) / n).max() < TOL[dt] * (2 if dt == np.float64 else 1)


@pytest.mark.parametrize("dt,tol", [(np.float32, 1e-5), (np.float64, 1e-13)])
@pytest.mark.parametrize("shp", REF2D)
def test_reference_large_2d(mcs, shp, dt, tol):
    x = np.random.default_rng(shp[0]).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    assert np.abs(np.fft.rfftn(x.astype(np.float64)) - k).max() / np.sqrt(x.size) < tol  # TestC


def test_small_golden(mcs):
    g = np.load(os.path.join(GOLD, "fft_small.npz"))
    i = 0
    while f"x{i}" in g:
        x, k = g[f"x{i}"], g[f"k{i}"]
        assert np.abs(mcs.rfftn(x) - k).max() / np.sqrt(x.size) < 2e-14, x.shape
        assert np.abs(mcs.irfftn(k, x.shape) / x.size - x).max() < 2e-14, x.shape
        i += 1


@pytest.mark.parametrize("shp", [(512,), (96, 160), (64, 48, 40), (32, 64, 128), (8, 1024), (3, 7, 10)])
def test_pfc_lengths_against_oracle(mcs, orc, shp):
    x = np.random.default_rng(7).standard_normal(shp)
    assert np.abs(mcs.rfftn(x) - orc.r2c(x)).max() / np.sqrt(x.size) < 2e-14
    x32 = x.astype(np.float32)
    assert np.abs(mcs.rfftn(x32) - orc.r2c(x32)).max() / np.sqrt(x.size) < 1e-5


@pytest.mark.parametrize("dt,tol", [(np.float32, 1e-5), (np.float64, 2e-14)])
@pytest.mark.parametrize("shp", [(24,), (40,), (4096,), (8192,), (6, 96), (5, 3, 160), (2, 2048), (3, 12)])
def test_register_row_pass_lengths(mcs, shp, dt, tol):
    # even nx whose half length is 2^a, 3*2^a or 5*2^a takes the register-engine x pass
    x = np.random.default_rng(len(shp) + shp[-1]).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    want = np.fft.rfftn(x.astype(np.float64))
    assert np.abs(want - k).max() / np.sqrt(x.size) < tol
    y = mcs.irfftn(want.astype(k.dtype), shp)
    assert np.abs(y.astype(np.float64) / x.size - x).max() < tol * 4


def test_c2r_non_hermitian_follows_numpy(mcs):
    rng = np.random.default_rng(4)
    for shp in [(4, 6, 8), (5, 9), (6, 7)]:
        hs = shp[:-1] + (shp[-1] // 2 + 1,)
        k = rng.standard_normal(hs) + 1j * rng.standard_normal(hs)
        got = mcs.irfftn(k, shp) / np.prod(shp)
        assert np.abs(got - np.fft.irfftn(k, s=shp, axes=tuple(range(len(shp))))).max() < 1e-14


def test_c2r_leaves_input_untouched(mcs):
    shp = (16, 12, 10)
    k = np.fft.rfftn(np.random.default_rng(1).standard_normal(shp))
    nb = mcs.buffer_bytes(shp, np.float64)
    din = mcs.to_device(k, nb)
    dout = mcs.DeviceBuffer(nb)
    mcs.FFTPlan(shp, np.float64, mcs.C2R).execute(din, dout)
    assert np.array_equal(din.download(k.shape, k.dtype), k)


def test_unsupported_length_is_an_error(mcs):
    with pytest.raises(mcs.McsError) as e:
        mcs.FFTPlan((17, 34), np.float32, mcs.R2C)
    assert e.value.code == mcs.ERR_UNSUPPORTED


@pytest.mark.parametrize("shp,dt,tol", [((24, 100, 256), np.float32, 1e-5), ((12, 70, 128), np.float64, 2e-14),
                                        ((10, 130, 256), np.float32, 1e-5)])
def test_long_stride_columns_in_row_interleaved_order(mcs, shp, dt, tol):
    # z columns over rows of nx/2+1 values with a stride above 64 KB run in the row-interleaved
    # XCD order, whose last ny % 8 rows of tiles take plain order (ny = 100, 70, 130 here; the
    # 10-point z columns run on the generic engine)
    x = np.random.default_rng(shp[1]).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    want = np.fft.rfftn(x.astype(np.float64))
    assert np.abs(want - k).max() / np.sqrt(x.size) < tol
    y = mcs.irfftn(k, shp)
    assert np.abs(y.astype(np.float64) / x.size - x).max() < tol * 2


@pytest.mark.parametrize("shp,dt,tol", [((8192, 6), np.float32, 1e-5), ((6144, 4), np.float64, 2e-14),
                                        ((10240, 4), np.float32, 1e-5), ((12288, 2), np.float32, 1e-5),
                                        ((5120, 4), np.float64, 2e-14), ((8192, 4), np.float64, 2e-14)])
def test_long_column_lines(mcs, shp, dt, tol):
    """columns whose default plan would exceed the register kernels' launch bound take the next plan
    within it (32 points for 8192 in FP32, 6 points for 6144 in FP64, 10 for 10240 in FP32)"""
    x = np.random.default_rng(shp[0]).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    assert np.abs(np.fft.rfftn(x.astype(np.float64)) - k).max() / np.sqrt(x.size) < tol
    y = mcs.irfftn(k, shp)
    assert np.abs(x.astype(np.float64) - y.astype(np.float64) / x.size).max() < tol


def test_in_place_r2c_is_rejected(mcs):
    shp = (4, 6, 8)
    x = np.random.default_rng(1).standard_normal(shp).astype(np.float32)
    buf = mcs.to_device(x, mcs.buffer_bytes(shp, np.float32))
    r2c = mcs.FFTPlan(shp, np.float32, mcs.R2C)
    with pytest.raises(mcs.McsError):
        r2c.execute(buf, buf)
    # C2R of a 3D spectrum runs through the plan's scratch, so in place is fine there
    ref = mcs.rfftn(x)
    buf.upload(ref)
    c2r = mcs.FFTPlan(shp, np.float32, mcs.C2R)
    c2r.execute(buf, buf)
    assert np.abs(buf.download(shp, np.float32) / x.size - x).max() < 1e-5


@pytest.mark.parametrize("shp", [(360, 36, 24), (90, 150, 72), (15, 18, 50), (1200, 30)])
@pytest.mark.parametrize("dt,tol", [(np.float32, 1e-5), (np.float64, 2e-14)])
def test_mixed_radix_columns(mcs, shp, dt, tol):
    """columns of lengths 3^b 5^c 2^a (30-, 12- and 20-point register plans) against numpy"""
    x = np.random.default_rng(sum(shp)).standard_normal(shp).astype(dt)
    k = mcs.rfftn(x)
    assert np.abs(np.fft.rfftn(x.astype(np.float64)) - k).max() / np.sqrt(x.size) < tol
    y = mcs.irfftn(k, shp)
    assert np.abs(x.astype(np.float64) - y.astype(np.float64) / x.size).max() < tol
