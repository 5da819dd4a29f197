"""The fixed-plan kernels of the fused step compute what the runtime-plan kernels
compute, on plans the fixed table holds (pinned) and on the tuned plans, with
and without noise, in both splits, FP32 and FP64, 2D and 3D.  The operations
are the same; the compiler fuses multiply-adds per basic block, and the fixed
plans are one straight block where the runtime plans branch per pass, so
individual values may round differently: the states agree to a few units of
the last place (relative L2 within 1e-6 in FP32 and 1e-14 in FP64 after 12
steps), and the mass and step counters exactly."""
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
def test_fixed_kernels_match_runtime_plans(mcs, shp, dtype, temp, split, plan):
    psi = mcs.init_field(shp, -0.25, 0.05, 11, dtype)
    a = _run(mcs, psi, shp, dtype, temp, split, plan, mcs.FIXED_COLUMNS | mcs.FIXED_XPASS, 12)
    b = _run(mcs, psi, shp, dtype, temp, split, plan, 0, 12)
    assert a[3] == b[3]  # the same plans ran
    assert a[1] == b[1] and a[2] == b[2] == 12
    assert np.isfinite(a[0]).all()
    x, y = a[0].astype(np.complex128), b[0].astype(np.complex128)
    rel = np.linalg.norm(x - y) / np.linalg.norm(y)
    print(f"{shp} {np.dtype(dtype).name} T={temp} split={split} plan={plan}: rel-L2 fixed vs runtime {rel:.2e}")
    assert rel < (1e-6 if dtype == np.float32 else 1e-14)


def test_fixed_kernel_mask(mcs):
    assert mcs.get_fixed_kernels() == mcs.FIXED_COLUMNS | mcs.FIXED_XPASS
    with pytest.raises(mcs.McsError):
        mcs.set_fixed_kernels(4)
    mcs.set_fixed_kernels(mcs.FIXED_COLUMNS)
    assert mcs.get_fixed_kernels() == mcs.FIXED_COLUMNS
    mcs.set_fixed_kernels(mcs.FIXED_COLUMNS | mcs.FIXED_XPASS)
