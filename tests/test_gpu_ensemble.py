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


@pytest.mark.parametrize("form", [(0, 0), (1, 1), (1, 0), (0, 1)])
def test_launch_forms_agree_bit_for_bit(mcs, form):
    # plan-class launches (size groups of one register tier in one persistent launch per pass) and
    # per-group launches, eager and graph-replayed, against the looping launches; 2D and 3D replicas
    # of power-of-two, 3*2^a and 5*2^a sizes, noisy and not
    reps = small_3d() + small_2d()
    a = mcs.Ensemble(reps, dtype=np.float32)
    a.init()
    a.step(19, mcs.ENS_LOOP)
    b = mcs.Ensemble(reps, dtype=np.float32)
    b.set_launch_form(*form)
    b.init()
    b.step(3, mcs.ENS_BATCHED)
    b.step(16, mcs.ENS_GRAPH)
    info = b.launch_info()
    assert info["batched_uses_classes"] == bool(form[0]) and info["graph_replays_classes"] == bool(form[1])
    assert info["class_launches_per_step"] < 2 * info["size_groups"] + 3 * 4
    for r in range(len(reps)):
        assert np.array_equal(a.get_psi(r), b.get_psi(r)), r


@pytest.mark.parametrize("branches", [1, 2, 3])
def test_packed_graph_branches_agree_bit_for_bit(mcs, branches):
    # the per-group step graph with the size groups packed onto fewer branches (each running its
    # groups one after the other) against the looping launches
    reps = small_3d() + small_2d()
    a = mcs.Ensemble(reps, dtype=np.float32)
    a.init()
    a.step(19, mcs.ENS_LOOP)
    b = mcs.Ensemble(reps, dtype=np.float32)
    b.set_launch_form(0, 0)
    b.set_graph_branches(branches)
    b.init()
    b.step(3, mcs.ENS_BATCHED)
    b.step(16, mcs.ENS_GRAPH)
    info = b.launch_info()
    assert not info["graph_replays_classes"]
    assert info["graph_branches"] == min(branches, info["size_groups"])
    for r in range(len(reps)):
        assert np.array_equal(a.get_psi(r), b.get_psi(r)), r


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
    # with the stepper's fixed-plan kernels where its shapes have them; replicas keep the
    # spectral form's state
    reps = small_2d()
    mcs.set_autotune(False)
    try:
        e = mcs.Ensemble(reps, dtype=np.float32)
        e.init()
        e.step(12, mcs.ENS_GRAPH)
        for r, p in enumerate(reps):
            st = mcs.Stepper(p["shape"], [H] * 3, p["eps"], p["dt"], temp=p["temp"], seed=p["noise_seed"],
                             dtype=np.float32, form=mcs.FORM_SPECTRAL)
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


def _size_groups(reps):
    """one replica index per grid size"""
    first = {}
    for r, p in enumerate(reps):
        first.setdefault(tuple(p["shape"]), r)
    return sorted(first.values())


def _graph_equals_loop(mcs, reps, nsteps):
    a = mcs.Ensemble(reps, dtype=np.float32)
    a.init()
    a.step(nsteps, mcs.ENS_LOOP)
    b = mcs.Ensemble(reps, dtype=np.float32)
    b.init()
    b.step(nsteps, mcs.ENS_GRAPH)
    for r in _size_groups(reps):
        assert np.array_equal(a.get_psi(r), b.get_psi(r)), (r, reps[r]["shape"])


def _initial_means(mcs, reps):
    """the exact mean of each replica's seeded FP32 initial field (the FP64 mass the ensemble carries)"""
    return [float(mcs.init_field(p["shape"], p["psi0"], p["amp"], p["init_seed"], np.float32).astype(np.float64)
                  .mean()) for p in reps]


def _assert_mass(f, m0, r):
    """the mass is carried in FP64; the exported FP32 field keeps the mean to the FP32 roundings of the
    packed k = 0 value (which also holds the Nyquist mode) and of the export scale: two ulps of the
    mean (3e-8 at |mean| = 0.25; a replica that loses mass is off by orders of magnitude more)"""
    assert np.isfinite(f).all(), r
    d = abs(f.mean() - m0)
    assert d <= 2 * np.spacing(np.float32(abs(m0))), (r, d)


def test_config_2d_ensemble_1000_steps(mcs, orc):
    """64 replicas of {128,192,320,384,640,768,1024}^2, T = 5e-4, dt = 0.5, 1000 steps replayed from the
    captured graph: every replica finite with its mean density carried exactly (FP64 k = 0; the rest is
    the rounding of the exported FP32 field); graph and looping launches bit for bit over the first 24
    steps for one replica of every size; the peak lists of two replicas equal the oracle's on their
    exported fields (cuda/kernels/findpeaks_2d.hpp:7-27)"""
    reps = mcs.ensemble_config("2d")
    _graph_equals_loop(mcs, reps, 24)
    e = mcs.Ensemble(reps, dtype=np.float32)
    e.init()
    m0 = _initial_means(mcs, reps)
    e.step(1000, mcs.ENS_GRAPH)
    for r in range(len(reps)):
        _assert_mass(e.get_psi(r).astype(np.float64), m0[r], r)
    for r in (_size_groups(reps)[0], _size_groups(reps)[-1]):
        f = e.get_psi(r)
        thr = float(f.mean())
        got = mcs.findpeaks(f, thr)
        assert got.size > 10 and np.array_equal(got, orc.findpeaks(f.astype(np.float64), thr))


def test_config_3d_ensemble_2000_steps_with_in_graph_analysis(mcs, orc):
    """32 replicas of {64,96,128,160,192}^3, T = 1e-3, dt = 0.25, 2000 steps through mcs_ensemble_run:
    the last captured graph also holds every replica's peak finding and 12-NN q4/q6 64-bin
    histograms (cuda/kernels/findpeaks.hpp:31-63 -> bond-order-analysis/bond-order.cpp:289-349).
    Every replica's mean density is carried exactly, every histogram totals its peak count, two
    replicas' peak counts and histograms equal the oracle's on their exported fields, and graph and
    looping launches agree bit for bit over the first 24 steps for one replica of every size."""
    reps = mcs.ensemble_config("3d")
    _graph_equals_loop(mcs, reps, 24)
    e = mcs.Ensemble(reps, dtype=np.float32)
    e.init()
    m0 = _initial_means(mcs, reps)
    counts, h4, h6 = e.run(2000, -10.0, k=12, nbins=64)
    assert (counts > 12).all()
    assert np.array_equal(h4.sum(axis=1), counts) and np.array_equal(h6.sum(axis=1), counts)
    for r in range(len(reps)):
        _assert_mass(e.get_psi(r).astype(np.float64), m0[r], r)
    groups = _size_groups(reps)
    for r in (groups[0], groups[len(groups) // 2]):
        f = e.get_psi(r).astype(np.float64)
        n = reps[r]["shape"][0]
        idx = orc.findpeaks(f, -10.0)
        assert counts[r] == idx.size
        ijk = np.column_stack([idx % n, (idx // n) % n, idx // (n * n)]).astype(np.int32)
        _, out = orc.bond_order_knn(ijk, n, H, 12)
        assert np.array_equal(h4[r], np.bincount(np.minimum((out[:, 0] * 64).astype(int), 63), minlength=64))
        assert np.array_equal(h6[r], np.bincount(np.minimum((out[:, 1] * 64).astype(int), 63), minlength=64))


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
