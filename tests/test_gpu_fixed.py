"""The fixed-plan kernels of the fused step compute exactly what the runtime-plan
kernels compute: the same state bit for bit after several steps, on plans the
fixed table holds (pinned) and on the tuned plans, with and without noise, in
both splits, FP32 and FP64, 2D and 3D.  Device code is built without implicit
multiply-add contraction (every fused operation is written out), so kernels
built from the same expressions round alike whatever their control flow."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = np.pi / 4

CASES = [
    # shape, dtype, temp, split, plan (None: tuned)
    ((512, 512, 512), np.float32, 0.0, 0, [8, 2, 8, 16, 32, 16]),
    ((512, 512, 512), np.float32, 0.0, 0, [8, 2, 16, 8, 16, 8]),
    ((512, 512, 512), np.float64, 0.0, 0, [8, 2, 8, 8, 16, 8]),
    ((256, 256, 256), np.float32, 1e-3, 0, [8, 4, 16, 16, 8, 16]),
    ((256, 256, 256), np.float32, 1e-3, 1, [16, 8, 8, 32, 16, 8]),
    ((256, 256, 256), np.float64, 1e-3, 0, [8, 4, 8, 8, 16, 8]),
    ((256, 256), np.float64, 0.0, 0, [8, 4, 8, 8]),
    # 4-point x plans (a row pair per wavefront)
    ((512, 512, 512), np.float32, 0.0, 0, [4, 2, 32, 16, 32, 16]),
    ((256, 256, 256), np.float64, 1e-3, 1, [4, 4, 8, 8, 16, 8]),
    ((256, 256), np.float32, 1e-3, 0, [4, 2, 8, 8]),
    ((256, 256, 256), np.float32, 1e-3, 0, None),
    ((512, 512, 512), np.float64, 0.0, 0, None),
]


def _run(mcs, psi, shp, dtype, temp, split, plan, mask, steps, form=None):
    mcs.set_fixed_kernels(mask)
    try:
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, temp=temp, seed=7, split=split, dtype=dtype, plan=plan,
                         form=form)
        st.set_psi(psi)
        st.step(steps)
        spec, mass, step = st.get_state()
        info = st.plan_info()
        del st
    finally:
        mcs.set_fixed_kernels(mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)
    return spec, mass, step, info


@pytest.mark.parametrize("form", [0, 1])  # FORM_SPECTRAL, FORM_MIXED: both forms' x passes
@pytest.mark.parametrize("shp,dtype,temp,split,plan", CASES)
def test_fixed_kernels_bit_identical_to_runtime_plans(mcs, shp, dtype, temp, split, plan, form):
    psi = mcs.init_field(shp, -0.25, 0.05, 11, dtype)
    a = _run(mcs, psi, shp, dtype, temp, split, plan, mcs.FIXED_COLUMNS | mcs.FIXED_XPASS, 12, form)
    b = _run(mcs, psi, shp, dtype, temp, split, plan, 0, 12, form)
    assert a[3] == b[3]  # the same plans ran
    assert a[1] == b[1] and a[2] == b[2] == 12
    assert np.isfinite(a[0]).all()
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))


def test_fixed_kernel_mask(mcs):
    assert mcs.get_fixed_kernels() == mcs.FIXED_COLUMNS | mcs.FIXED_XPASS
    with pytest.raises(mcs.McsError):
        mcs.set_fixed_kernels(4)
    mcs.set_fixed_kernels(mcs.FIXED_COLUMNS)
    assert mcs.get_fixed_kernels() == mcs.FIXED_COLUMNS
    mcs.set_fixed_kernels(mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)


def _fft_pair(mcs, x, mask):
    mcs.set_fixed_kernels(mask)
    try:
        k = mcs.rfftn(x)
        y = mcs.irfftn(k, x.shape)
    finally:
        mcs.set_fixed_kernels(mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)
    return k, y


# boundary transforms whose row and column plans the fixed tables hold: row lengths nx/2 of
# 128, 256, 512 and column lengths 256, 512, FP32 and FP64
FFT_CASES = [
    ((512, 512, 512), np.float32),
    ((256, 256, 256), np.float64),
    ((256, 256, 256), np.float32),
    ((64, 1024), np.float32),
    ((32, 512), np.float64),
    ((4, 16, 1024), np.float64),
    ((3, 5, 256), np.float32),
]


@pytest.mark.parametrize("shp,dtype", FFT_CASES)
def test_fixed_boundary_fft_bit_identical_to_runtime_plans(mcs, shp, dtype):
    x = np.random.default_rng(len(shp) + shp[-1]).standard_normal(shp).astype(dtype)
    k1, y1 = _fft_pair(mcs, x, mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)
    k0, y0 = _fft_pair(mcs, x, 0)
    assert np.array_equal(k1.view(np.uint8), k0.view(np.uint8))
    assert np.array_equal(y1.view(np.uint8), y0.view(np.uint8))
    # and both are the transform, within the reference test's tolerances: the round trip at every
    # size, numpy's forward transform below 2^24 points
    tol = 1e-5 if dtype == np.float32 else 2e-14
    assert np.abs(x.astype(np.float64) - y1.astype(np.float64) / x.size).max() < tol
    if x.size <= 1 << 24:
        assert np.abs(np.fft.rfftn(x.astype(np.float64)) - k1).max() / np.sqrt(x.size) < tol


SHUFFLE_CASES = [
    # shape, dtype, temp, split, plan: x plans whose last two passes have radix P (4*8*8, 16*16 and
    # 4*4*4*4 for nx/2 = 256, 2*8*8 and 2*4*4*4 for 128), so the fixed x kernels have a lane-shuffle form
    ((512, 512, 512), np.float32, 0.0, 0, [8, 2, 8, 16, 32, 16]),
    ((512, 512, 512), np.float32, 0.0, 1, [16, 4, 8, 16, 32, 16]),
    ((256, 256, 256), np.float32, 1e-3, 0, [8, 4, 16, 16, 8, 16]),
    ((512, 512, 512), np.float64, 0.0, 0, [8, 2, 8, 8, 16, 8]),
    ((256, 256), np.float64, 1e-3, 1, [8, 4, 8, 8]),
    # 4*4*4*4 and 2*4*4*4: a 4 x 4 transpose over lanes 16 and 32 apart (v_permlane16/32_swap), 8 and 16
    ((512, 512, 512), np.float32, 0.0, 0, [4, 2, 32, 16, 32, 16]),
    ((256, 256, 256), np.float32, 1e-3, 0, [4, 4, 16, 16, 8, 16]),
]


@pytest.mark.parametrize("form", [0, 1])  # FORM_SPECTRAL, FORM_MIXED
@pytest.mark.parametrize("shp,dtype,temp,split,plan", SHUFFLE_CASES)
def test_lane_shuffle_exchange_bit_identical(mcs, shp, dtype, temp, split, plan, form):
    # the register/lane transpose (DPP, v_permlane16/32_swap) moves the same values into the same
    # registers as the LDS exchange it replaces
    psi = mcs.init_field(shp, -0.25, 0.05, 11, dtype)
    out = []
    for mode in (0, 2):
        mcs.set_lane_shuffle(mode)
        try:
            out.append(_run(mcs, psi, shp, dtype, temp, split, plan, mcs.FIXED_COLUMNS | mcs.FIXED_XPASS, 12,
                            form))
        finally:
            mcs.set_lane_shuffle(1)
    a, b = out
    assert a[3] == b[3] and a[2] == b[2] == 12
    assert np.isfinite(a[0]).all()
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))


@pytest.mark.parametrize("form", [0, 1])
def test_lane_shuffle_matches_oracle(mcs, orc, form):
    shp = (256, 256)
    psi = orc.init_field(shp, -0.25, 0.05, 1)
    mcs.set_lane_shuffle(2)
    try:
        st = mcs.Stepper(shp, [H] * 3, 0.25, 0.5, dtype=np.float64, plan=[8, 4, 8, 8], form=form)
        st.set_psi(psi)
        st.step(50)
        got = st.get_psi()
    finally:
        mcs.set_lane_shuffle(1)
    want = orc.pfc_run(psi, [H] * 3, 0.25, 0.5, 0.0, 0, 0, 50, 0)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12
    with pytest.raises(mcs.McsError):
        mcs.set_lane_shuffle(3)

