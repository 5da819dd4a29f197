"""Every plan the autotuner may pick is correct: each register plan (points per thread) and
tile shape of each pass, pinned explicitly, reproduces the FP64 oracle; the tuned stepper
reports plans from the candidate lists and also matches the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = np.pi / 4


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def default_plan(mcs, shp, dtype):
    nx = shp[-1]
    axes = [(nx // 2, True), (shp[-2], False)] + ([(shp[0], False)] if len(shp) == 3 else [(1, False)])
    plan = []
    for L, xpass in axes:
        if L == 1:
            plan += [0, 0]
            continue
        P = mcs.plan_options(L, dtype)[0]
        plan += [P, mcs.tile_options(L, dtype, P, xpass)[0]]
    return plan


def variants(mcs, shp, dtype):
    """the default plan with one pass changed at a time to every other (P, tile) candidate"""
    base = default_plan(mcs, shp, dtype)
    nx = shp[-1]
    lens = [nx // 2, shp[-2]] + ([shp[0]] if len(shp) == 3 else [])
    out = []
    for i, L in enumerate(lens):
        for P in mcs.plan_options(L, dtype):
            for G in mcs.tile_options(L, dtype, P, i == 0):
                v = list(base)
                v[2 * i], v[2 * i + 1] = P, G
                if v != base:
                    out.append(v)
    return base, out


# (z, y, x) shapes; lengths 128, 256 and 512 also run the fixed-plan kernels of their plans
CASES = [((8, 512, 16), np.float32), ((8, 32, 1024), np.float32), ((32, 32, 32), np.float32),
         ((8, 16, 256), np.float64), ((48, 80), np.float32), ((64, 256), np.float64),
         ((4, 256, 512), np.float32), ((12, 96, 160), np.float32)]


@pytest.mark.parametrize("shp,dtype", CASES)
def test_every_candidate_plan_matches_oracle(mcs, orc, shp, dtype):
    psi = orc.init_field(shp, -0.2, 0.1, 5)
    want = orc.pfc_run(psi.astype(dtype).astype(np.float64), [H] * 3, 0.3, 0.25, 1e-3, 9, 0, 6)
    tol = 1e-11 if dtype == np.float64 else 5e-5
    base, vs = variants(mcs, shp, dtype)
    assert vs, "no alternative plans to check"
    for plan in [base] + vs:
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, temp=1e-3, seed=9, dtype=dtype, plan=plan)
        info = st.plan_info()
        assert [info["x"]["P"], info["x"]["tile"], info["y"]["P"], info["y"]["tile"]] == plan[:4]
        st.set_psi(psi)
        st.step(6)
        assert rel_l2(st.get_psi(), want) < tol, plan


def test_tuned_plans_are_candidates_and_correct(mcs, orc):
    assert mcs.get_autotune()
    for shp, dtype in CASES[:3]:
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, dtype=dtype)
        info = st.plan_info()
        lens = {"x": shp[-1] // 2, "y": shp[-2], "z": shp[0]}
        for ax, d in info.items():
            assert d["P"] in mcs.plan_options(lens[ax], dtype)
            assert d["tile"] in mcs.tile_options(lens[ax], dtype, d["P"], ax == "x")
        psi = orc.init_field(shp, -0.2, 0.1, 5)
        st.set_psi(psi)
        st.step(6)
        want = orc.pfc_run(psi.astype(dtype).astype(np.float64), [H] * 3, 0.3, 0.25, 0.0, 0, 0, 6)
        assert rel_l2(st.get_psi(), want) < 5e-5


def test_autotune_off_uses_default_rule_and_rejects_bad_plans(mcs):
    shp = (32, 32, 32)
    mcs.set_autotune(False)
    try:
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25)
        info = st.plan_info()
        assert info["x"]["P"] == 8 and info["y"]["P"] == 8 and info["z"]["P"] == 8
    finally:
        mcs.set_autotune(True)
    with pytest.raises(mcs.McsError) as e:
        mcs.Stepper(shp, [H] * 3, 0.3, 0.25, plan=[8, 8, 8, 512, 8, 16])  # 512 lines x 4 threads > 1024
    assert e.value.code == mcs.ERR_UNSUPPORTED
    # any tile within the launch bound and LDS is accepted, also one the tuner would not try, and the
    # tile shape does not change the arithmetic
    psi = mcs.init_field(shp, -0.25, 0.05, 3, np.float32)
    out = []
    for plan in ([8, 8, 8, 3, 8, 16], [8, 8, 8, 8, 8, 8]):
        st = mcs.Stepper(shp, [H] * 3, 0.3, 0.25, plan=plan)
        st.set_psi(psi)
        st.step(5)
        out.append(st.get_psi())
    assert np.array_equal(out[0], out[1])
    with pytest.raises(mcs.McsError):
        mcs.Stepper(shp, [H] * 3, 0.3, 0.25, plan=[12, 16, 8, 16, 8, 16])  # no 12-point plan for 16
