"""HIP path against the oracle on the same inputs and Philox draws (FP64).

Tolerances: deterministic quantities 1e-12 relative (the bar for staples,
action density, overrelaxation and flow steps); after stochastic
passes the fields agree to rounding (1e-11 absolute on O(1) values) because
every accept/reject decision is the same; counters are exact."""
import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import (FP64, HEATBATH, METROPOLIS, OVERRELAX, K_DOUBLET_METROPOLIS,
                              K_DOUBLET_OVERRELAX, K_LINK_HEATBATH, K_LINK_METROPOLIS,
                              K_LINK_OVERRELAX, K_SINGLET_METROPOLIS, K_SINGLET_OVERRELAX,
                              K_TRIPLET_METROPOLIS, K_TRIPLET_OVERRELAX, K_U1_METROPOLIS,
                              make_params)
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu

MODELS = {
    "g": dict(nhiggs=0),
    "h": dict(nhiggs=1),
    "hu": dict(nhiggs=1, u1=1),
    "hs": dict(nhiggs=1, singlet=1),
    "hsu": dict(nhiggs=1, u1=1, singlet=1),
    "t": dict(nhiggs=0, triplet=1),
    "ht": dict(nhiggs=1, triplet=1),
    "d": dict(nhiggs=2),
    "du": dict(nhiggs=2, u1=1),
}
FIELDS = ("su2link", "u1link", "doublet", "triplet", "singlet")


def pair(tag, L=(4, 6, 8), **kw):
    p = make_params(L=L, **MODELS[tag], **kw)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    b = Batch(p)
    b.set_fields(0, f)
    return p, o, b


def maxdiff(fo, fg):
    d = 0.0
    for k in FIELDS:
        if fg.get(k) is not None:
            d = max(d, np.abs(fo[k] - fg[k]).max())
    return d


@pytest.mark.parametrize("tag", list(MODELS))
def test_roundtrip_and_hot_start(tag):
    p, o, b = pair(tag)
    f = o.get_fields()
    g = b.get_fields(0)
    for k in FIELDS:
        if g.get(k) is not None:
            assert np.array_equal(g[k], f[k]), k
    b.hot_start(0)
    g = b.get_fields(0)
    assert maxdiff(f, g) < 1e-13


@pytest.mark.parametrize("tag", list(MODELS))
def test_measure_sums(tag):
    p, o, b = pair(tag)
    a, g = o.measure(), b.measure(0)
    scale = np.maximum(np.abs(a), 1.0)
    assert (np.abs(a - g) / scale).max() < 1e-12


def kinds(p):
    ks = [(K_LINK_HEATBATH, 3), (K_LINK_METROPOLIS, 3), (K_LINK_OVERRELAX, 3)]
    if p.u1:
        ks.append((K_U1_METROPOLIS, 3))
    if p.nhiggs:
        ks += [(K_DOUBLET_OVERRELAX, p.nhiggs), (K_DOUBLET_METROPOLIS, p.nhiggs)]
    if p.triplet:
        ks += [(K_TRIPLET_OVERRELAX, 1), (K_TRIPLET_METROPOLIS, 1)]
    if p.singlet:
        ks += [(K_SINGLET_OVERRELAX, 1), (K_SINGLET_METROPOLIS, 1)]
    return ks


@pytest.mark.parametrize("tag", list(MODELS))
def test_passes_site_by_site(tag):
    """Every pass kind, both parities: same accepts, same fields."""
    p, o, b = pair(tag)
    for kind, ndir in kinds(p):
        for d in range(ndir):
            for par in (0, 1):
                o.pass_(kind, par, d)
                b.pass_(kind, par, d)
        fo, fg = o.get_fields(), b.get_fields(0)
        assert maxdiff(fo, fg) < 1e-11, kind
        # resynchronise rounding before the next kind
        b.set_fields(0, fo)
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]


@pytest.mark.parametrize("L", [(4, 6, 8), (8, 8, 8), (16, 8, 12)])
@pytest.mark.parametrize("beta", [1.0, 8.0])
def test_link_heatbath_wavefront_rounds(L, beta):
    """FP64 link heatbath: full wavefronts share the direction's rejection
    tries across lanes, partial ones (the 32-site tail at 4x6x8) loop per
    lane; both give the oracle's sites, accepts and stream positions."""
    p, o, b = pair("hu", L=L, betasu2=beta)
    for sweep in range(2):
        for d in range(3):
            for par in (0, 1):
                o.pass_(K_LINK_HEATBATH, par, d)
                b.pass_(K_LINK_HEATBATH, par, d)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-12
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]


@pytest.mark.parametrize("L", [(8, 8, 8), (16, 8, 12)])
def test_link_heatbath_loop_limit(L):
    """Pure gauge at beta = 1e-6: the staples nearly vanish (a = 2|V| ~
    1e-5), so every Kennedy-Pendleton try has d ~ 1e5 > 2 and is rejected,
    and the reference's loop stops after 200 tries with a0 = 1e-9
    (src/heatbath.c:64-68).  Sites and accepts are the oracle's, and so is
    the stream position the loop leaves (the direction's draws start
    there)."""
    p, o, b = pair("g", L=L, betasu2=1e-6)
    for d in range(3):
        for par in (0, 1):
            o.pass_(K_LINK_HEATBATH, par, d)
            b.pass_(K_LINK_HEATBATH, par, d)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-12
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]


@pytest.mark.parametrize("tag", list(MODELS))
def test_compound_sweeps(tag):
    p, o, b = pair(tag, link_overrelax=2)
    o.update(2)
    b.update(2)
    # two doublets: the two-doublet overrelaxation amplifies rounding 10-30x
    # per sweep (test_oracle_ref.test_higgs2_rounding_growth)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < (1e-10 if p.nhiggs < 2 else 1e-8)
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]


@pytest.mark.parametrize("tag", ["h", "hu", "ht", "hsu"])
@pytest.mark.parametrize("random_order", [0, 1])
@pytest.mark.parametrize("cache", [True, False], ids=["cache", "inline"])
def test_two_link_rounds_per_sweep(tag, random_order, cache):
    """update_links = 2: the U(1) passes of the first round change the
    hopping term that the second round's link staples read (the link passes
    take it from a cache filled before each round), and the doublet passes
    of one sweep change it for the next.  Without the cache the passes form
    the term themselves."""
    p, o, b = pair(tag, link_overrelax=1, update_links=2)
    assert b.hop_cache()
    b.set_hop_cache(cache)
    assert b.hop_cache() == cache
    if random_order:
        o.lib.orc_set_random_order(o.h, 1)
        b.set_random_order(True)
    o.update(3)
    b.update(3)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k


def test_hop_cache_switched_between_sweeps():
    """The cache dropped and kept again between sweeps (graph replay on): the
    chain is the oracle's throughout, pass counters and budgets included."""
    p, o, b = pair("hu", link_overrelax=1)
    b.use_graph(True)
    b.set_budget(0, 7)
    for on in (True, False, True, False):
        b.set_hop_cache(on)
        o.update(1)
        b.update(1)
        assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10, on
    assert o.rng_get()[1] == b.rng_get(0)[1]
    assert b.budgets()[0] == 3


def test_hop_cache_needs_one_doublet_and_fits_the_budget():
    import ctypes as C
    for tag in ("g", "t", "d"):
        b = Batch(make_params(L=(4, 4, 4), **MODELS[tag]))
        assert not b.hop_cache()
        with pytest.raises(abi.Su2hipError, match="one doublet"):
            b.set_hop_cache(True)
        b.close()
    p = make_params(L=(8, 8, 8), **MODELS["hu"])
    lib = abi.load()
    fp = C.c_uint64()
    assert lib.su2hip_footprint(C.byref(p), 0, C.byref(fp)) == 0
    # the footprint counts the fields, the cache (12 words per site) and a
    # doublet backup for a weight (4 words per site): a budget one byte short
    # of fields and cache
    V, w = abi.volume(p), 8
    b = Batch(p, hbm_budget=fp.value - 4 * V * w - 1)
    assert not b.hop_cache()
    with pytest.raises(abi.Su2hipError, match="budget"):
        b.set_hop_cache(True)
    assert not b.hop_cache()
    b.update(1)
    b.close()


def test_hop_cache_default_follows_the_infinity_cache():
    """Kept for the 96^3 workload and for lattices beyond the Infinity Cache;
    dropped for the FP32 96-lattice scan, whose fields fit it (about 200 MB)
    and whose fields and caches (about 390 MB) do not."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    for prec, want in ((abi.FP32, False), (FP64, True)):
        ps = bench.scan_params(prec)
        b = Batch(ps)
        assert b.hop_cache() == want, prec
        b.close()
    b = Batch(bench.workload_params(abi.FP32))
    assert b.hop_cache()
    b.close()


@pytest.mark.parametrize("tag", ["h", "hu"])
def test_link_pass_after_scalar_and_u1_passes(tag):
    """Single passes interleaved: each link pass sees the doublet and U(1)
    links that the passes before it left."""
    p, o, b = pair(tag)
    seq = [(K_LINK_OVERRELAX, 0, 1), (K_DOUBLET_OVERRELAX, 1, 0), (K_LINK_HEATBATH, 1, 1),
           (K_DOUBLET_METROPOLIS, 0, 0), (K_LINK_OVERRELAX, 0, 2)]
    if p.u1:
        seq += [(K_U1_METROPOLIS, 1, 2), (K_LINK_OVERRELAX, 1, 2), (K_U1_METROPOLIS, 0, 0),
                (K_LINK_HEATBATH, 0, 0)]
    for kind, par, d in seq:
        o.pass_(kind, par, d)
        b.pass_(kind, par, d)
        assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-11, (kind, par, d)


def test_graph_replay_matches_eager():
    p = make_params(L=(8, 8, 8), nhiggs=1, link_overrelax=2)
    b1, b2 = Batch(p), Batch(p)
    b1.hot_start(0)
    b2.hot_start(0)
    b2.use_graph(True)
    for _ in range(3):
        b1.update(2)
        b2.update(2)
        b1.sync()
        b2.sync()
        f1, f2 = b1.get_fields(0), b2.get_fields(0)
        for k in FIELDS:
            if f1.get(k) is not None:
                assert np.array_equal(f1[k], f2[k])
    assert b1.counters(0) == b2.counters(0)


def test_graph_replay_after_every_reallocation_matches_eager():
    """A captured sweep graph holds only kernel launches whose arguments are
    device pointers and integers; every buffer they name (block tables and
    their counters, partial sums, random-order tables, weight arrays) is
    reallocated only together with dropping the graph.  Replays across
    each such path are bitwise the eager launches: lattices added until
    the block table grows past its capacity, a slot freed and refilled, a
    chunk change (the other block table), random order switched on (its
    tables) and a weight installed (its arrays)."""
    base = dict(nhiggs=1, link_overrelax=1)
    hs = [Batch(make_params(L=(8, 8, 8), seed=1, **base), max_slots=6) for _ in range(2)]
    hs[1].use_graph(True)

    def step(n=2):
        for h in hs:
            h.update(n)
            h.sync()
        for slot in hs[0].params:
            f0, f1 = hs[0].get_fields(slot), hs[1].get_fields(slot)
            for k in FIELDS:
                if f0.get(k) is not None:
                    assert np.array_equal(f0[k], f1[k]), (slot, k)
            assert hs[0].counters(slot) == hs[1].counters(slot), slot

    for h in hs:
        h.hot_start(0)
    step()
    for L, seed in (((16, 16, 16), 2), ((24, 16, 32), 3), ((32, 32, 32), 4)):
        for h in hs:
            h.hot_start(h.add(make_params(L=L, seed=seed, **base)))
        step()
    for h in hs:
        h.remove(1)
        h.hot_start(h.add(make_params(L=(16, 8, 24), seed=5, **base)))
    step()
    for h in hs:
        h.set_chunk(abi.K_LINK_OVERRELAX, 64)
    step()
    for h in hs:
        h.set_random_order(True)
    step()
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep, m.bins = 1, abi.PHISQ, abi.FAST, 2, 40
    m.min, m.max, m.wrk_min, m.wrk_max, m.do_acceptance, m.delta = 0.0, 2.0, 0.0, 2.0, 1, 0.3
    m.update_interval = 5
    pos, W = muca_weight()
    for h in hs:
        h.muca_set(0, m, pos, W)
    step(3)
    for h in hs:
        h.close()


def test_batch_same_alone_and_in_mix():
    """Lattices of different volume swept in one launch give bitwise the
    results each gives alone."""
    ps = [make_params(L=L, nhiggs=1, seed=100 + i, link_overrelax=1, msq_phi=-0.3 + 0.05 * i)
          for i, L in enumerate([(4, 4, 4), (6, 4, 8), (8, 8, 8), (4, 6, 6)])]
    mix = Batch(ps)
    for i in range(len(ps)):
        mix.hot_start(i)
    mix.update(3)
    for i, p in enumerate(ps):
        alone = Batch(p)
        alone.hot_start(0)
        alone.update(3)
        fa, fm = alone.get_fields(0), mix.get_fields(i)
        for k in FIELDS:
            if fa.get(k) is not None:
                assert np.array_equal(fa[k], fm[k]), (i, k)
        assert np.array_equal(alone.measure(0), mix.measure(i))


def test_batch_matches_oracle():
    ps = [make_params(L=L, nhiggs=1, seed=7 + i) for i, L in enumerate([(4, 4, 4), (6, 4, 8)])]
    mix = Batch(ps)
    for i, p in enumerate(ps):
        o = orc.Oracle(p)
        o.hot_start()
        mix.set_fields(i, o.get_fields())
        o.update(2)
        ps[i] = (p, o)
    mix.update(2)
    for i, (p, o) in enumerate(ps):
        assert maxdiff(o.get_fields(), mix.get_fields(i)) < 1e-10


def test_budgets():
    ps = [make_params(L=(4, 4, 4), nhiggs=1, seed=i) for i in range(3)]
    b = Batch(ps)
    for i in range(3):
        b.hot_start(i)
    f1 = b.get_fields(1)
    b.set_budget(0, 2)
    b.set_budget(1, 0)
    b.update(3)
    bud = b.budgets()
    assert bud[0] == 0 and bud[1] == 0 and bud[2] == -1
    # lattice 0 advanced exactly 2 sweeps: compare with the oracle
    o = orc.Oracle(ps[0])
    o.hot_start()
    o.update(2)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10
    # lattice 1 (budget 0) was not touched at all
    assert maxdiff(f1, b.get_fields(1)) == 0.0


@pytest.mark.parametrize("model", [dict(nhiggs=1), dict(nhiggs=2, u1=1)])
def test_flow_force_and_euler(model):
    p = make_params(L=(4, 6, 8), **model)
    o = orc.Oracle(p)
    o.hot_start()
    b = Batch(p)
    b.set_fields(0, o.get_fields())
    F, _ = o.flow_force()
    G = b.flow_force(0)
    assert np.abs(F - G).max() < 1e-12 * max(1.0, np.abs(F).max())
    ro = o.flow(abi.FLOW_EULER, 0.01, 0.05, 2)
    rg = b.flow(0, abi.FLOW_EULER, 0.01, 0.05, 2)
    assert ro.shape == rg.shape
    assert np.abs(ro - rg).max() / np.abs(ro).max() < 1e-12
    ro = o.flow(abi.FLOW_RK3, 0.02, 0.1, 1)
    rg = b.flow(0, abi.FLOW_RK3, 0.02, 0.1, 1)
    assert np.abs(ro - rg).max() / np.abs(ro).max() < 1e-12


def muca_weight(bins=40, lo=0.0, hi=2.0, seed=3, slope=5.0):
    rng = np.random.default_rng(seed)
    pos = np.zeros(bins + 2)
    W = np.zeros(bins + 2)
    dbin = (hi - lo) / bins
    for i in range(bins + 1):
        pos[i + 1] = lo + i * dbin
    slopes = rng.uniform(-slope, slope, bins)
    for i in range(1, bins + 1):
        W[i + 1] = W[i] + slopes[i - 1] * dbin
    pos[0] = lo - 1000.0 * (hi - lo)
    W[0] = W[1]
    W[bins + 1] = W[bins]
    return pos, W


@pytest.mark.parametrize("mode", [abi.READONLY, abi.FAST, abi.SLOW])
@pytest.mark.parametrize("tag,op,cps,slope", [("h", abi.PHISQ, 1, 5.0), ("t", abi.SIGMASQ, 1, 5.0),
                                               ("h", abi.PHISQ, 3, 5.0),
                                               ("ht", abi.PHI2MINUSSIGMA2, 2, 5.0),
                                               ("h", abi.PHISQ, 3, 400.0),
                                               ("t", abi.SIGMASQ, 2, 400.0),
                                               ("d", abi.PHI2SQ, 2, 5.0),
                                               ("du", abi.PHI2SQ, 3, 400.0)])
def test_muca_sweeps(tag, op, cps, slope, mode):
    p, o, b = pair(tag, L=(4, 4, 6))
    m = abi.Muca()
    m.enabled = 1
    m.orderparam = op
    m.mode = mode
    m.checks_per_sweep = cps
    m.bins = 40
    m.min, m.max = 0.0, 2.0
    m.wrk_min, m.wrk_max = 0.0, 2.0
    m.do_acceptance = 1
    m.delta = 0.3
    m.update_interval = 5
    lo, hi = (-1.0, 1.0) if op == abi.PHI2MINUSSIGMA2 else (0.0, 2.0)
    m.min, m.max, m.wrk_min, m.wrk_max = lo, hi, lo, hi
    pos, W = muca_weight(lo=lo, hi=hi, slope=slope)
    o.muca_set(m, pos, W)
    b.muca_set(0, m, pos, W)
    mo, _, _, _ = o.muca_get()
    mg, _, _, _ = b.muca_get(0)
    assert abs(mo.param_value[0] - mg.param_value[0]) < 1e-13
    nsw = 3 if p.nhiggs < 2 else 2
    o.update(nsw)
    b.update(nsw)
    tol = 1e-10 if p.nhiggs < 2 else 1e-8
    assert maxdiff(o.get_fields(), b.get_fields(0)) < tol
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    if slope > 100:
        assert 0 < co["accepted_muca"] < co["total_muca"]  # the restore path ran
    mo, po, Wo, ho = o.muca_get()
    mg, pg, Wg, hg = b.muca_get(0)
    assert np.abs(Wo - Wg).max() < (1e-12 if p.nhiggs < 2 else 1e-9)
    assert (ho == hg).all()
    assert abs(mo.param_value[1] - mg.param_value[1]) < (1e-13 if p.nhiggs < 2 else 1e-9)
    if mode == abi.SLOW:
        go, no = o.muca_sums()
        gg, ng = b.muca_sums(0)
        assert (go[1:-1] == gg[1:-1]).all() and (no[1:-1] == ng[1:-1]).all()


def test_muca_set_failure_leaves_the_slot_as_it_was():
    """A weight whose backup does not fit the handle's HBM budget is
    refused with SU2HIP_ENOMEM before anything is freed or installed: the
    slot stays without a weight, and its sweeps are bitwise those of a
    handle that was never asked."""
    import ctypes as C
    p = make_params(L=(4, 6, 8), **MODELS["h"])
    lib = abi.load()
    fp = C.c_uint64()
    assert lib.su2hip_footprint(C.byref(p), 0, C.byref(fp)) == 0
    b = Batch(p, hbm_budget=fp.value - 1)  # the fields fit, fields + backup do not
    ref = Batch(p)
    for h in (b, ref):
        h.hot_start(0)
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep, m.bins = 1, abi.PHISQ, abi.READONLY, 1, 40
    m.min, m.max, m.wrk_min, m.wrk_max, m.do_acceptance = 0.0, 2.0, 0.0, 2.0, 1
    pos, W = muca_weight()
    with pytest.raises(abi.Su2hipError, match="budget"):
        b.muca_set(0, m, pos, W)
    assert b.muca_get(0)[0].enabled == 0
    for h in (b, ref):
        h.update(2)
    fb, fr = b.get_fields(0), ref.get_fields(0)
    for k in FIELDS:
        if fr.get(k) is not None:
            assert np.array_equal(fb[k], fr[k]), k
    assert b.counters(0) == ref.counters(0)
    b.close()
    ref.close()


def test_muca_slow_recursion_from_loaded_sums():
    """SLOW weight updates (src/multicanonical.c:561-666) that change W, from
    gsum/nsum loaded as a weight_params file would give them."""
    p, o, b = pair("t", L=(4, 4, 6), seed=2027)
    o.rng_set(99, 0)
    b.rng_set(0, 99, 0)
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, abi.SIGMASQ, abi.SLOW, 2
    m.bins, m.delta, m.do_acceptance, m.update_interval = 40, 0.3, 1, 30
    m.min, m.max, m.wrk_min, m.wrk_max = 0.5, 1.5, 0.5, 1.5
    pos, W = muca_weight(lo=0.5, hi=1.5, slope=5.0)
    rng = np.random.default_rng(4)
    gs = rng.integers(5, 60, 42).astype(np.int64)
    ns = rng.integers(1, 40, 42).astype(np.int64)
    o.muca_set(m, pos, W)
    o.muca_set_sums(gs, ns)
    b.muca_set(0, m, pos, W)
    b.muca_set_sums(0, gs, ns)
    o.update(6)
    b.update(6)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10
    mo, _, Wo, _ = o.muca_get()
    mg, _, Wg, _ = b.muca_get(0)
    assert np.abs(Wo - Wg).max() < 1e-12 and np.abs(Wg - W).max() > 0  # W was updated
    go, no = o.muca_sums()
    gg, ng = b.muca_sums(0)
    assert (go[1:-1] == gg[1:-1]).all() and (no[1:-1] == ng[1:-1]).all()
    assert (gg != gs).any()


def test_plain_16cubed_trajectory():
    """The plain 16^3 case: SU(2)+Higgs, betaG 8, msq -0.25, lambda 0.02, hot
    start from seed 1234, each sweep 1 link heatbath + 4 link overrelaxation
    + 4 Higgs overrelaxation + 1 Higgs Metropolis, plaquette and <phi^+phi>
    measured every sweep: the GPU's FP64 series follows the oracle's with
    the same accept/reject decisions (exact counters).  The two differ by
    rounding (fused multiply-adds, libm ulps), which the overrelaxation
    dynamics amplify about 4x per sweep, so the series is compared over the
    first 5 sweeps (1e-10 relative); the measure of the same fields agrees
    to 1e-12."""
    p = make_params(L=(16, 16, 16), nhiggs=1, betasu2=8.0, msq_phi=-0.25, lambda_phi=0.02,
                    link_overrelax=4, scalar_sweeps=4, seed=1234)
    o = orc.Oracle(p)
    o.hot_start()
    b = Batch(p)
    b.hot_start(0)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-14
    b.set_fields(0, o.get_fields())  # the same bits (the hot starts differ in libm ulps)
    b.use_graph(True)
    V = abi.volume(p)
    for it in range(5):
        so, sg = o.measure(), b.measure(0)
        for name in ("wilson", "phi2"):
            k = abi.SUM[name]
            assert abs(so[k] - sg[k]) <= 1e-10 * abs(so[k]), (it, name)
        o.update(1)
        b.update(1)
    assert o.counters() == b.counters(0)
    c = Batch(p)
    c.set_fields(0, o.get_fields())
    so, sc = o.measure(), c.measure(0)
    assert np.abs(so - sc).max() <= 1e-12 * np.abs(so).max()
    plaq = 1 - sg[abi.SUM["wilson"]] / (V * p.betasu2) / 3
    assert 0.5 < plaq < 1.0


def test_async_flow_equals_sync_flow_while_others_sweep():
    """su2hip_flow_start/poll/finish on the second stream gives the rows of
    su2hip_flow bit for bit, while another lattice of the handle keeps
    sweeping, and that lattice ends exactly as it would alone."""
    ps = [make_params(L=(8, 8, 8), nhiggs=1, seed=7), make_params(L=(12, 8, 8), nhiggs=1, seed=8)]
    sync = Batch(ps)
    for i in range(2):
        sync.hot_start(i)
    rows_sync = sync.flow(0, abi.FLOW_RK3, 0.05, 1.0, 4)
    b = Batch(ps)
    for i in range(2):
        b.hot_start(i)
    b.set_budget(0, 0)
    b.flow_start(0, abi.FLOW_RK3, 0.05, 1.0, 4)
    with pytest.raises(abi.Su2hipError):
        b.set_budget(0, 3)          # the flowing lattice may not sweep
    with pytest.raises(abi.Su2hipError):
        b.flow_start(0, abi.FLOW_RK3, 0.05, 1.0, 4)
    b.update(5)                      # lattice 1 sweeps meanwhile
    rows = b.flow_finish(0)
    assert rows.shape == rows_sync.shape and np.array_equal(rows, rows_sync)
    alone = Batch(ps[1])
    alone.hot_start(0)
    alone.update(5)
    assert np.array_equal(alone.get_fields(0)["su2link"], b.get_fields(1)["su2link"])
    assert not np.array_equal(sync.get_fields(1)["su2link"], b.get_fields(1)["su2link"])
    # a lattice with a budget cannot start an asynchronous flow
    with pytest.raises(abi.Su2hipError):
        b.flow_start(1, abi.FLOW_RK3, 0.05, 1.0, 4)


@pytest.mark.parametrize("precision", [FP64, abi.FP32])
@pytest.mark.parametrize("tag", ["h", "t", "ht"])
@pytest.mark.parametrize("L", [(4, 6, 8), (6, 6, 6)])
def test_correlators(tag, L, precision):
    """Plane-sum correlators (the CORRELATORS build's measure_correlators)
    against the oracle on the same fields: FP64 to 1e-12 of each column's
    scale, FP32 to float rounding; every direction and the longest one."""
    p = make_params(L=L, **MODELS[tag], precision=precision, seed=5)
    o = orc.Oracle(make_params(L=L, **MODELS[tag], seed=5))
    o.hot_start()
    f = o.get_fields()
    if precision == abi.FP32:
        f = {k: (v.astype(np.float32).astype(np.float64) if v is not None else None) for k, v in f.items()}
        o.set_fields(f)
    b = Batch(p)
    b.set_fields(0, f)
    tol = 1e-12 if precision == FP64 else 2e-5
    for d in (0, 1, 2, -1):
        dd = o.lib.orc_longest_dir(o.h) if d < 0 else d
        a, g = o.correlators(dd), b.correlators(0, d)
        assert g.shape == a.shape
        scale = np.abs(a).max(0) + 1e-300
        assert (np.abs(a - g) <= tol * scale).all(), (d, np.abs(a - g).max(0) / scale)


@pytest.mark.parametrize("tag", ["h", "t", "hsu"])
@pytest.mark.parametrize("precision", [FP64, abi.FP32])
def test_wall_start_and_zprofile(tag, precision):
    """su2hip_wall_start equals the oracle's two-phase start (same Philox
    draws; FP32 rounds the same values), and su2hip_zprofile its plane
    averages, along the longest direction and a given one."""
    L = (4, 8, 6)
    p = make_params(L=L, **MODELS[tag], precision=precision, seed=21)
    o = orc.Oracle(make_params(L=L, **MODELS[tag], seed=21))
    o.cold_start()
    o.wall_start()
    b = Batch(p)
    b.wall_start(0)
    fo, fg = o.get_fields(), b.get_fields(0)
    # FP32: sqrt(1 - u^2) with a fused multiply-add; links also through the
    # packed 12-byte form (renormalised, lowest mantissa bits carry the
    # position and sign of the rebuilt component: a few units in the last place)
    tol = 1e-13 if precision == FP64 else 1e-7
    for k in FIELDS:
        if fg.get(k) is not None and k != "u1link":
            t = 4e-7 if (k == "su2link" and precision == abi.FP32) else tol
            assert np.abs(fo[k] - fg[k]).max() <= t, k
    if precision == abi.FP32:
        o.set_fields({k: v for k, v in fg.items() if v is not None})
    for d in (-1, 0, 2):
        a = o.zprofile(d)
        g = b.zprofile(0, d)
        assert g.shape == a.shape == (L[1 if d < 0 else d], 4)
        assert np.abs(a - g).max() <= 1e-12 * max(1.0, np.abs(a).max())
    # the two phases: doublet/singlet low then high, triplet the other way
    z = b.zprofile(0)
    if p.nhiggs:
        assert z[0, 0] < 0.1 and z[-1, 0] > 0.5
    if p.triplet:
        assert z[0, 1] > 1.0 and z[-1, 1] < 0.1


@pytest.mark.parametrize("precision", [FP64, abi.FP32])
@pytest.mark.parametrize("L", [(8, 8, 12), (16, 8, 8)])
def test_blocked_correlators(L, precision):
    """Smeared, blocked lattices and their correlators against the oracle;
    models with a doublet measure nothing, as in the reference."""
    p = make_params(L=L, nhiggs=0, triplet=1, precision=precision, seed=33)
    o = orc.Oracle(make_params(L=L, nhiggs=0, triplet=1, seed=33))
    o.hot_start()
    f = o.get_fields()
    if precision == abi.FP32:
        f = {k: (v.astype(np.float32).astype(np.float64) if v is not None else None) for k, v in f.items()}
        o.set_fields(f)
    b = Batch(p)
    b.set_fields(0, f)
    a, g = o.blocked_correlators(5), b.blocked_correlators(0, 5)
    assert a.shape == g.shape and a.shape[0] == 2
    tol = 1e-12 if precision == FP64 else 1e-4
    scale = np.abs(a).max(1, keepdims=True) + 1e-300
    assert (np.abs(a - g) <= tol * scale).all(), np.abs(a - g).max((0, 1))
    h = Batch(make_params(L=L, nhiggs=1, triplet=1, precision=precision))
    assert h.blocked_correlators(0, 3).shape[0] == 0
