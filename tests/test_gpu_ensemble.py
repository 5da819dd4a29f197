"""Batched ensembles: the batched and graph-replayed launches reproduce the looping
single-replica launches and the single-grid stepper bit for bit, and the oracle within tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = np.pi / 4


def small_2d():
    reps = []
    for i, (s, eps) in enumerate([(64, 0.25), (96, 0.3), (160, 0.2), (64, 0.35), (96, 0.15)]):
        reps.append(dict(shape=(s, s), eps=eps, psi0=-0.25 + 0.01 * i, amp=0.05, temp=5e-4 * (i % 2), dt=0.5,
                         init_seed=10 + i, noise_seed=20 + i))
    return reps


def small_3d():
    reps = []
    for i, (s, eps) in enumerate([(32, 0.3), (48, 0.25), (40, 0.2), (32, 0.22)]):
        reps.append(dict(shape=(s, s, s), eps=eps, psi0=-0.2 - 0.02 * i, amp=0.05, temp=1e-3, dt=0.25,
                         init_seed=30 + i, noise_seed=40 + i))
    return reps


@pytest.mark.parametrize("make", [small_2d, small_3d])
def test_modes_agree_bit_for_bit(mcs, make):
    reps = make()
    outs = []
    for mode in (mcs.ENS_LOOP, mcs.ENS_BATCHED, mcs.ENS_GRAPH):
        e = mcs.Ensemble(reps, dtype=np.float32)
        e.init()
        e.step(19, mode)  # two graph replays of 8 steps + 3 eager steps
        outs.append([e.get_psi(r) for r in range(len(reps))])
    for r in range(len(reps)):
        assert np.array_equal(outs[0][r], outs[1][r]) and np.array_equal(outs[0][r], outs[2][r])


def test_graph_replays_between_eager_steps(mcs):
    # the captured graph holds one branch per size group; replays interleaved with eager steps and host
    # synchronisations must continue the same noise sequence as the looping launches
    reps = small_3d() + [dict(shape=(16, 64), eps=0.3, psi0=-0.25, amp=0.05, temp=5e-4, dt=0.5, init_seed=5,
                              noise_seed=6)]
    a = mcs.Ensemble(reps, dtype=np.float32)
    a.init()
    a.step(27, mcs.ENS_LOOP)
    b = mcs.Ensemble(reps, dtype=np.float32)
    b.init()
    b.step(8, mcs.ENS_GRAPH)
    b.step(3, mcs.ENS_BATCHED)
    b.step(16, mcs.ENS_GRAPH)
    for r in range(len(reps)):
        assert np.array_equal(a.get_psi(r), b.get_psi(r)), r


def test_replicas_match_single_stepper(mcs):
    # the same plans (the default rule: steppers and ensembles tune separately, since the
    # ensemble's work-table kernels run the runtime plans) give the same fields bit for bit,
    # with the stepper's fixed-plan kernels where its shapes have them
    reps = small_2d()
    mcs.set_autotune(False)
    try:
        e = mcs.Ensemble(reps, dtype=np.float32)
        e.init()
        e.step(12, mcs.ENS_GRAPH)
        for r, p in enumerate(reps):
            st = mcs.Stepper(p["shape"], [H] * 3, p["eps"], p["dt"], temp=p["temp"], seed=p["noise_seed"],
                             dtype=np.float32)
            st.set_psi(mcs.init_field(p["shape"], p["psi0"], p["amp"], p["init_seed"], np.float32))
            st.step(12)
            assert np.array_equal(st.get_psi(), e.get_psi(r))
    finally:
        mcs.set_autotune(True)


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-11), (np.float32, 5e-5)])
def test_replicas_match_oracle(mcs, orc, dt, tol):
    reps = small_3d()[:2] + small_2d()[:2]
    e = mcs.Ensemble(reps, dtype=dt)
    e.init()
    e.step(10, mcs.ENS_BATCHED)
    for r, p in enumerate(reps):
        psi = orc.init_field(p["shape"], p["psi0"], p["amp"], p["init_seed"])
        want = orc.pfc_run(psi.astype(dt).astype(np.float64), [H] * 3, p["eps"], p["dt"], p["temp"],
                           p["noise_seed"], 0, 10, 0)
        got = e.get_psi(r).astype(np.float64)
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


def test_config_ensembles_build_and_step(mcs):
    """The two ensemble workloads at full size: 64 replicas of {128..1024}^2 and 32 of {64..192}^3."""
    for kind, n in (("2d", 10), ("3d", 4)):
        reps = mcs.ensemble_config(kind)
        e = mcs.Ensemble(reps, dtype=np.float32)
        e.init()
        e.step(n, mcs.ENS_GRAPH)
        for r in (0, len(reps) - 1):
            p = e.get_psi(r).astype(np.float64)
            assert np.isfinite(p).all()
            assert abs(p.mean() - reps[r]["psi0"]) < 2e-3


def test_analysis_histograms_match_oracle(mcs, orc):
    reps = [dict(shape=(48, 48, 48), eps=0.3, psi0=-0.2, amp=0.05, temp=0.0, dt=0.5, init_seed=7, noise_seed=1),
            dict(shape=(32, 32, 32), eps=0.35, psi0=-0.18, amp=0.05, temp=0.0, dt=0.5, init_seed=8, noise_seed=2)]
    e = mcs.Ensemble(reps, dtype=np.float64)
    e.init()
    e.step(200, mcs.ENS_GRAPH)
    counts, h4, h6 = e.analyze(-10.0, k=12, nbins=64)
    for r, p in enumerate(reps):
        f = e.get_psi(r)
        idx = orc.findpeaks(f, -10.0)
        assert counts[r] == idx.size and idx.size > 12
        n = p["shape"][0]
        ijk = np.column_stack([idx % n, (idx // n) % n, idx // (n * n)]).astype(np.int32)
        _, out = orc.bond_order_knn(ijk, n, H, 12)
        assert np.array_equal(h4[r], np.bincount(np.minimum((out[:, 0] * 64).astype(int), 63), minlength=64))
        assert np.array_equal(h6[r], np.bincount(np.minimum((out[:, 1] * 64).astype(int), 63), minlength=64))


def test_run_with_in_graph_analysis_matches_eager(mcs):
    # mcs_ensemble_run: steps replayed from graphs, the last one also holding the analysis;
    # it must equal eager stepping followed by mcs_ensemble_analyze, on repeated calls too
    reps = small_3d()
    a = mcs.Ensemble(reps, dtype=np.float32)
    a.init()
    b = mcs.Ensemble(reps, dtype=np.float32)
    b.init()
    for nsteps in (19, 16, 8):
        got = a.run(nsteps, -10.0, k=12, nbins=64)
        b.step(nsteps, mcs.ENS_LOOP)
        want = b.analyze(-10.0, k=12, nbins=64)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        assert got[0].min() > 12
    for r in range(len(reps)):
        assert np.array_equal(a.get_psi(r), b.get_psi(r))
    with pytest.raises(mcs.McsError):
        a.run(5, -10.0)
