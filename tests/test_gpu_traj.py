"""The random pass order (the reference's random_sweeps) and the heatbath
trajectories of its HB_TRAJECTORY build, HIP path against the oracle and
against the reference compiled from source, on the same Philox draws."""
import ctypes as C
import os

import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import TrajectoryConfig, make_params, ptr
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu

MODELS = {
    "h": dict(nhiggs=1),
    "hu": dict(nhiggs=1, u1=1),
    "t": dict(nhiggs=0, triplet=1),
    "d": dict(nhiggs=2),
}


def present(f):
    return {k: v for k, v in f.items() if v is not None}


def maxdiff(a, b):
    return max(np.abs(a[k] - b[k]).max() for k in present(b))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("tag", list(MODELS))
def test_random_order_matches_oracle(tag, graph):
    p = make_params(L=(4, 6, 8), link_overrelax=1, seed=21, **MODELS[tag])
    o = orc.Oracle(p)
    o.hot_start()
    b = Batch(p)
    b.set_fields(0, o.get_fields())
    o.lib.orc_set_random_order(o.h, 1)
    b.set_random_order(True)
    b.use_graph(graph)
    o.update(3)
    b.update(3)
    tol = 1e-10 if p.nhiggs < 2 else 1e-8
    assert maxdiff(o.get_fields(), b.get_fields(0)) < tol
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]


def test_random_order_with_multicanonical_steps():
    """Random parities of the scalar rounds inside multicanonical parity
    sweeps (checks, restores) on the device."""
    from test_gpu_parity import muca_weight
    p = make_params(L=(4, 4, 6), nhiggs=1, seed=22)
    o = orc.Oracle(p)
    o.hot_start()
    b = Batch(p)
    b.set_fields(0, o.get_fields())
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, abi.PHISQ, abi.READONLY, 3
    m.bins, m.min, m.max, m.wrk_min, m.wrk_max, m.delta = 40, 0.0, 2.0, 0.0, 2.0, 0.3
    m.do_acceptance = 1
    pos, W = muca_weight(slope=400.0)
    o.muca_set(m, pos, W)
    b.muca_set(0, m, pos, W)
    o.lib.orc_set_random_order(o.h, 1)
    b.set_random_order(True)
    o.update(3)
    b.update(3)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert 0 < co["accepted_muca"] < co["total_muca"]


def test_random_order_is_per_lattice_in_a_batch():
    """Each lattice of a batch draws its own order from its own key: the same
    bits in the mix as alone."""
    ps = [make_params(L=L, nhiggs=1, seed=300 + i) for i, L in enumerate([(4, 4, 4), (6, 4, 8), (4, 6, 6)])]
    mix = Batch(ps)
    for i in range(len(ps)):
        mix.hot_start(i)
    mix.set_random_order(True)
    mix.update(2)
    for i, p in enumerate(ps):
        one = Batch(p)
        one.hot_start(0)
        one.set_random_order(True)
        one.update(2)
        a, g = one.get_fields(0), mix.get_fields(i)
        for k in present(a):
            assert np.array_equal(a[k], g[k]), (i, k)


def write_weight(path, bins, lo, hi, seed=3):
    rng = np.random.default_rng(seed)
    dbin = (hi - lo) / bins
    W = [0.0]
    for s in rng.uniform(-5, 5, bins):
        W.append(W[-1] + s * dbin)
    with open(path, "w") as f:
        f.write("%d %.12f %d %f %f\n" % (bins, 0.3, 0, lo, hi))
        for i in range(bins + 1):
            f.write("%f %f\n" % (lo + i * dbin, W[i]))


@pytest.mark.parametrize("tag,kw", [("hr", dict(nhiggs=1)), ("tr", dict(nhiggs=0, triplet=1))])
def test_trajectories_match_reference(tmp_path, tag, kw):
    """Two sets of heatbath trajectories (make_realtime_trajectories of the
    HB_TRAJECTORY build) from a multicanonical lattice: the trajectory file
    line for line, the state left behind (fields, counters, Philox pass,
    the order parameter's parity values) as the reference leaves it."""
    if not orc.ref_available(tag):
        pytest.skip("reference build absent")
    p = make_params(L=(4, 4, 6), seed=2031, **kw)
    op = abi.SIGMASQ if p.triplet else abi.PHISQ
    o = orc.Oracle(p)
    o.hot_start()
    o.update(30)  # near equilibrium, so that trajectories last several sweeps
    f = o.get_fields()
    os.chdir(tmp_path)
    write_weight("weight", 40, 0.0, 4.0)
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, op, abi.READONLY, 1
    m.bins, m.min, m.max, m.delta = 40, 0.0, 4.0, 0.3
    r = orc.Ref(p, trajectory=True)
    r.set_fields(f)
    r.rng_set(p.seed, 0)
    assert r.lib.ref_muca_load(b"weight", C.byref(m)) == 0
    mr = abi.Muca()
    pos, W = np.zeros(42), np.zeros(42)
    r.lib.ref_muca_get(C.byref(mr), ptr(pos), ptr(W), None)
    v0 = mr.param_value[0] + mr.param_value[1]
    t = TrajectoryConfig()
    t.n_traj, t.interval, t.mode_interval = 3, 1, 1
    t.min, t.max = v0 - 0.1, v0 + 0.1
    # the reference: two sets, one sweep of the main chain in between
    r.lib.ref_trajectories(t.n_traj, t.interval, t.min, t.max, 1)
    r.update(1)
    r.lib.ref_trajectories(t.n_traj, t.interval, t.min, t.max, 2)
    os.rename("trajectory", "trajectory_ref")
    lib = abi.load()
    b = Batch(p)
    b.set_fields(0, f)
    b.muca_set(0, mr, pos, W)
    assert lib.su2hip_trajectories(b.h, 0, C.byref(t), 1, b"trajectory") == 0
    b.update(1)
    assert lib.su2hip_trajectories(b.h, 0, C.byref(t), 2, b"trajectory") == 0
    ours = open("trajectory").read().splitlines()
    ref = open("trajectory_ref").read().splitlines()
    assert len(ours) == len(ref) > 30
    for a, c in zip(ours, ref):
        if a == c:
            continue
        # an order-parameter line whose 8th decimal sits on a rounding edge
        x, y = a.split(), c.split()
        assert len(x) == len(y) == 2 and x[0] == y[0], (a, c)
        assert abs(float(x[1]) - float(y[1])) <= 1.01e-8, (a, c)
    assert maxdiff(r.get_fields(), b.get_fields(0)) < 1e-10
    co, cg = r.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert b.rng_get(0)[1] == r.lib.ref_rng_pass()
    # the parity values the reference leaves (its trajectories' last ones)
    r.lib.ref_muca_get(C.byref(mr), None, None, None)
    pv = b.muca_orderparam(0, recompute=False)
    assert abs(pv[0] - mr.param_value[0]) < 1e-12 and abs(pv[1] - mr.param_value[1]) < 1e-12
    # multicanonical acceptance is back on, the order back to the default
    mg = b.muca_get(0)[0]
    assert mg.do_acceptance == 1
    en = C.c_int()
    assert lib.su2hip_get_random_order(b.h, C.byref(en)) == 0 and en.value == 0


def test_trajectories_leave_other_lattices_alone(tmp_path):
    """Trajectories of one slot of a batch: the other lattice neither sweeps
    nor changes, its budget comes back, and the slot ends as a run alone."""
    from test_gpu_parity import muca_weight
    os.chdir(tmp_path)
    ps = [make_params(L=(4, 4, 6), nhiggs=1, seed=41), make_params(L=(4, 6, 6), nhiggs=1, seed=42)]
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, abi.PHISQ, abi.READONLY, 1
    m.bins, m.min, m.max, m.wrk_min, m.wrk_max, m.delta = 40, 0.0, 4.0, 0.0, 4.0, 0.3
    m.do_acceptance = 1
    pos, W = muca_weight(hi=4.0)
    t = TrajectoryConfig()
    t.n_traj, t.interval, t.mode_interval = 2, 1, 1
    lib = abi.load()
    mix = Batch(ps)
    for i in range(2):
        mix.hot_start(i)
        mix.muca_set(i, m, pos, W)
    mix.update(3)
    v = mix.muca_orderparam(0).sum()
    t.min, t.max = v - 0.1, v + 0.1
    before = mix.get_fields(1)
    c1 = mix.counters(1)
    mix.set_budget(1, 7)
    assert lib.su2hip_trajectories(mix.h, 0, C.byref(t), 1, b"traj_mix") == 0
    after = mix.get_fields(1)
    for k in present(before):
        assert np.array_equal(before[k], after[k]), k
    assert mix.counters(1) == c1
    assert mix.budgets()[1] == 7
    one = Batch(ps[0])
    one.hot_start(0)
    one.muca_set(0, m, pos, W)
    one.update(3)
    assert lib.su2hip_trajectories(one.h, 0, C.byref(t), 1, b"traj_one") == 0
    assert open("traj_mix").read() == open("traj_one").read()
    a, g = one.get_fields(0), mix.get_fields(0)
    for k in present(a):
        assert np.array_equal(a[k], g[k]), k
