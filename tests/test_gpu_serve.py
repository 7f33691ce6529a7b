"""The request loop on the GPU: results independent of the mix, admission
under an HBM cap, chain continuation, flow requests, and the FP64 result of
one request against the oracle."""
import numpy as np
import pytest

import orc
from su2higgs_amd import abi, serve
from su2higgs_amd.abi import make_params

pytestmark = pytest.mark.gpu


def mix(n=9, seed=5, flow_every=0, L=(8, 12, 16), sweeps=(3, 17)):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        rows.append(dict(L=int(rng.choice(L)), sweeps=int(rng.integers(*sweeps)),
                         betasu2=float(rng.uniform(6, 12)), msq_phi=float(rng.uniform(-0.3, -0.1)),
                         lambda_phi=float(rng.uniform(0.005, 0.03)), seed=int(rng.integers(1, 2**62)),
                         flow=int(flow_every > 0 and (i + 1) % flow_every == 0)))
    return rows


def cfg(precision=abi.FP64, **kw):
    base = dict(realtime=0, tick_sweeps=4, flow_dt=0.05, flow_t_max=0.2,
                tmpl_precision=precision, tmpl_reunitarize=1 if precision == abi.FP32 else 0)
    base.update(kw)
    return serve.config(**base)


def same(a, b):
    return np.array_equal(np.array(a.sums), np.array(b.sums)) and \
        np.array_equal(np.array(a.flow_sums), np.array(b.flow_sums)) and a.pass_ == b.pass_


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
def test_result_is_the_same_alone_and_in_the_mix(precision):
    rows = mix(flow_every=4)
    res, st = serve.serve(cfg(precision), serve.requests(rows))
    assert st["completed"] == len(rows) and st["failed"] == 0
    assert st["peak_resident"] > 1
    for i in (0, 3, 7):
        alone, _ = serve.serve(cfg(precision), serve.requests([rows[i]]))
        assert alone[0].status == 0
        assert same(alone[0], res[i]), i
    # and with a different tick length (different interleaving of the chains)
    res2, _ = serve.serve(cfg(precision, tick_sweeps=1), serve.requests(rows))
    assert all(same(a, b) for a, b in zip(res, res2))


def test_fp64_result_matches_the_oracle():
    rows = [dict(L=8, sweeps=4, betasu2=7.5, msq_phi=-0.2, lambda_phi=0.01, seed=77, flow=1)]
    c = cfg(abi.FP64, flow_scheme=abi.FLOW_RK3)
    res, _ = serve.serve(c, serve.requests(rows))
    r = res[0]
    assert r.status == 0
    p = make_params(L=(8, 8, 8), betasu2=7.5, msq_phi=-0.2, lambda_phi=0.01, seed=77,
                    link_overrelax=c.tmpl.link_overrelax, scalar_sweeps=c.tmpl.scalar_sweeps)
    o = orc.Oracle(p)
    o.cold_start(phi0=c.phi0, sigma0=c.sigma0, singlet0=c.singlet0)
    o.rng_set(77, 0)
    o.update(4)
    s = o.measure()
    assert np.abs(np.array(r.sums) - s).max() <= 1e-10 * max(1.0, np.abs(s).max())
    assert r.pass_ == o.rng_get()[1]
    rows_o = o.flow(abi.FLOW_RK3, c.flow_dt, c.flow_t_max, 1 << 30)
    assert abs(r.flow_t - rows_o[-1][0]) < 1e-15
    assert np.abs(np.array(r.flow_sums) - rows_o[-1][1:]).max() <= 1e-10 * max(1.0, np.abs(s).max())


def test_hbm_cap_queues_but_does_not_change_results():
    rows = mix(n=8, seed=9, L=(16,))
    free, st_free = serve.serve(cfg(), serve.requests(rows))
    lib = abi.load()
    import ctypes as C
    fp = C.c_uint64()
    p = make_params(L=(16, 16, 16))
    abi.check(lib.su2hip_footprint(C.byref(p), 0, C.byref(fp)))
    cap = 2 * fp.value
    capped, st = serve.serve(cfg(hbm_budget=cap), serve.requests(rows))
    assert st["peak_bytes"] <= cap < st_free["peak_bytes"]
    assert st["peak_resident"] <= 2
    assert st["max_wait"] > 0
    assert all(same(a, b) for a, b in zip(free, capped))
    # a request larger than the cap fails alone, the rest complete
    big = rows + [dict(L=32, sweeps=2, betasu2=8, msq_phi=-0.2, lambda_phi=0.02, seed=3)]
    res, st = serve.serve(cfg(hbm_budget=cap), serve.requests(big))
    assert res[-1].status == abi.ENOMEM and st["failed"] == 1 and st["completed"] == len(rows)


@pytest.mark.parametrize("backfill", [0, 32], ids=["fifo", "backfill"])
def test_backfill_admits_small_requests_past_a_blocked_head(backfill):
    """Under an HBM cap that holds A with C but not A with B: with FIFO
    admission C waits behind B, which waits for A to finish; with backfill
    C starts beside A at once.  Every result is the same either way."""
    import ctypes as C
    lib = abi.load()

    def fp(L):
        v = C.c_uint64()
        abi.check(lib.su2hip_footprint(C.byref(make_params(L=(L,) * 3)), 0, C.byref(v)))
        return v.value

    rows = [dict(L=16, sweeps=60, betasu2=8, msq_phi=-0.2, lambda_phi=0.02, seed=1),
            dict(L=20, sweeps=10, betasu2=8, msq_phi=-0.2, lambda_phi=0.02, seed=2),
            dict(L=8, sweeps=10, betasu2=8, msq_phi=-0.2, lambda_phi=0.02, seed=3)]
    cap = fp(16) + fp(20) - 1
    assert fp(16) + fp(8) <= cap
    res, st = serve.serve(cfg(hbm_budget=cap, backfill=backfill, max_head_wait=60.0,
                              tick_sweeps=2), serve.requests(rows))
    assert st["completed"] == 3 and st["failed"] == 0
    a, b, c = res
    if backfill:
        assert c.t_admit < b.t_admit and c.t_admit <= a.t_done
    else:
        assert c.t_admit >= b.t_admit >= a.t_done
    ref, _ = serve.serve(cfg(), serve.requests(rows))  # no cap: all at once
    assert all(same(x, y) for x, y in zip(res, ref))


def test_chain_continuation_equals_one_long_request():
    one = [dict(L=12, sweeps=10, betasu2=9, msq_phi=-0.2, lambda_phi=0.02, seed=11)]
    long_, _ = serve.serve(cfg(), serve.requests(one))
    split = [dict(id=100, L=12, sweeps=4, betasu2=9, msq_phi=-0.2, lambda_phi=0.02, seed=11),
             dict(id=101, L=16, sweeps=5, betasu2=7, msq_phi=-0.1, lambda_phi=0.01, seed=12),
             dict(id=102, chain=100, sweeps=6)]
    res, st = serve.serve(cfg(keep_chains=1), serve.requests(split))
    assert [r.status for r in res] == [0, 0, 0]
    assert same(res[2], long_[0])
    # without keep_chains the chain is gone
    res, st = serve.serve(cfg(keep_chains=0), serve.requests(split))
    assert res[2].status == abi.ESTATE and st["failed"] == 1


def test_realtime_stream_reports_latency():
    rows = mix(n=12, seed=3, flow_every=6, L=(8, 12), sweeps=(20, 60))
    rng = np.random.default_rng(1)
    t = np.cumsum(rng.exponential(1 / 40.0, len(rows)))
    for r, a in zip(rows, t):
        r["arrival"] = float(a)
    res, st = serve.serve(cfg(abi.FP32, realtime=1), serve.requests(rows))
    assert st["completed"] == len(rows)
    lat = sorted(r.t_done - r.t_arrival for r in res)
    assert 0 < st["p50"] <= st["p99"] <= lat[-1] + 1e-12
    assert all(r.t_arrival <= r.t_admit <= r.t_done for r in res)
    assert all(abs(r.t_arrival - a) < 1e-12 for r, a in zip(res, t))
    assert st["makespan"] >= t[-1]
    assert st["flows"] == 2


def replay_alone(c, q):
    """The chain of request q alone in its own handle, from the same cold
    start and Philox key."""
    from su2higgs_amd.lattice import Batch
    p = make_params(L=(q.L,) * 3, precision=c.tmpl.precision, reunitarize=c.tmpl.reunitarize,
                    link_overrelax=c.tmpl.link_overrelax, scalar_sweeps=c.tmpl.scalar_sweeps,
                    betasu2=q.betasu2, msq_phi=q.msq_phi, lambda_phi=q.lambda_phi, seed=q.seed,
                    link_format=c.tmpl.link_format)  # the server's link storage
    b = Batch(p)
    b.cold_start(0, phi0=c.phi0, sigma0=c.sigma0, singlet0=c.singlet0)
    b.use_graph(True)
    b.update(int(q.sweeps))
    return b


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
def test_scan_serving_at_its_own_sizes(precision):
    """The serving workload at its own sizes: the seeded Poisson stream at
    20 requests/s of L in {24..64}, betaG in [6,12], msq in [-0.3,-0.1],
    lambda in [0.005,0.03] and 100..1000 sweeps, whose 10th request is 64^3
    with the RK3 flow dt = 0.01 to t = 4.0 (400 steps, 1200 stages).  Every
    request completes.  The flowed request's result equals its chain run
    alone in its own handle and flowed there, bit for bit (measure sums
    after the sweeps and the flowed sums at t = 4)."""
    reqs = serve.stream(2026, 20.0, 10, flow_every=10)
    Ls = {int(q.L) for q in reqs}
    assert reqs[9].flow == 1 and reqs[9].L == 64 and min(Ls) < 64
    c = serve.config(realtime=1, tmpl_precision=precision,
                     tmpl_reunitarize=1 if precision == abi.FP32 else 0)
    assert (c.flow_scheme, c.flow_dt, c.flow_t_max) == (abi.FLOW_RK3, 0.01, 4.0)
    res, st = serve.serve(c, reqs)
    assert st["completed"] == 10 and st["failed"] == 0 and st["flows"] == 1
    assert 0 < st["p50"] <= st["p99"]
    r, q = res[9], reqs[9]
    assert abs(r.flow_t - 4.0) < 1e-12
    b = replay_alone(c, q)
    rows = b.flow(0, c.flow_scheme, c.flow_dt, c.flow_t_max, 100)
    assert np.array_equal(np.array(r.sums), b.measure(0))
    assert abs(rows[-1][0] - 4.0) < 1e-12
    assert np.array_equal(np.array(r.flow_sums), rows[-1][1:])
    # the flow smooths: the Wilson action falls from each whole flow time to
    # the next (the loop clips its last step at t_max, src/gradflow.c:94-96,
    # so t = 4 may be measured twice, the second time after a step of ~1e-13)
    t = rows[:, 0]
    first = np.concatenate([[True], np.diff(t) > 1e-6])
    w = rows[first, 1 + abi.SUM["wilson"]]
    assert len(w) == 5 and (np.diff(w) < 0).all()
    b.close()


def test_flow_at_64_cubed_matches_the_oracle():
    """The FP64 RK3 flow of a 64^3 configuration against the oracle's on the
    same fields: every row of the first 5 steps (measured after each step)
    to 1e-12 of the action's scale (the oracle's 400 steps would take
    minutes; the kernel is the same at every step)."""
    q = serve.stream(2026, 20.0, 10, flow_every=10)[9]
    c = serve.config(tmpl_precision=abi.FP64, tmpl_reunitarize=0)
    b = replay_alone(c, q)
    rows = b.flow(0, abi.FLOW_RK3, 0.01, 0.05, 1)
    p = make_params(L=(64,) * 3, betasu2=q.betasu2, msq_phi=q.msq_phi, lambda_phi=q.lambda_phi,
                    seed=q.seed)
    o = orc.Oracle(p)
    o.set_fields(b.get_fields(0))
    ro = o.flow(abi.FLOW_RK3, 0.01, 0.05, 1)
    assert ro.shape == rows.shape == (6, 1 + abi.NSUMS)
    scale = np.abs(ro[:, 1:]).max()
    assert np.abs(ro - rows).max() <= 1e-12 * scale
    b.close()
