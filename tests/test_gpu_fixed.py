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
    ((256, 256, 256), np.float32, 1e-3, 0, None),
    ((512, 512, 512), np.float64, 0.0, 0, None),
]


def _run(mcs, psi, shp, dtype, temp, split, plan, mask, steps):
    mcs.set_fixed_kernels(mask)
    try:
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, temp=temp, seed=7, split=split, dtype=dtype, plan=plan)
        st.set_psi(psi)
        st.step(steps)
        spec, mass, step = st.get_state()
        info = st.plan_info()
        del st
    finally:
        mcs.set_fixed_kernels(mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)
    return spec, mass, step, info


@pytest.mark.parametrize("shp,dtype,temp,split,plan", CASES)
def test_fixed_kernels_bit_identical_to_runtime_plans(mcs, shp, dtype, temp, split, plan):
    psi = mcs.init_field(shp, -0.25, 0.05, 11, dtype)
    a = _run(mcs, psi, shp, dtype, temp, split, plan, mcs.FIXED_COLUMNS | mcs.FIXED_XPASS, 12)
    b = _run(mcs, psi, shp, dtype, temp, split, plan, 0, 12)
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
