"""The oracle restatement against the reference compiled from its own sources
(oracle/_ref, built by `make -C oracle ref`) on the same inputs and the same
Philox draws.  CPU only; skipped where the reference build is absent."""
import numpy as np
import pytest

import orc
from su2higgs_amd.abi import (FP64, HEATBATH, METROPOLIS, OVERRELAX, K_DOUBLET_METROPOLIS,
                              K_DOUBLET_OVERRELAX, K_LINK_HEATBATH, K_LINK_METROPOLIS,
                              K_SINGLET_METROPOLIS, K_SINGLET_OVERRELAX, K_TRIPLET_METROPOLIS,
                              K_TRIPLET_OVERRELAX, K_U1_METROPOLIS, make_params)

MODELS = {
    "h": dict(nhiggs=1),
    "hu": dict(nhiggs=1, u1=1),
    "t": dict(nhiggs=0, triplet=1),
    "ht": dict(nhiggs=1, triplet=1),
    "hsu": dict(nhiggs=1, u1=1, singlet=1),
    "d": dict(nhiggs=2),
    "du": dict(nhiggs=2, u1=1),
}


def need_ref(tag):
    if not orc.ref_available(tag):
        pytest.skip("reference build oracle/_ref/libref_%s.so absent" % tag)


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert orc.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_mapping():
    # draw k uses block k>>1, words (0,1) for even k and (2,3) for odd k,
    # top 53 bits of the 64-bit word
    w = orc.philox([3, 17, 5, 0], [1234, 0])
    u0 = orc.uniform(1234, 5, 17, 6)
    u1 = orc.uniform(1234, 5, 17, 7)
    assert u0 == ((w[1] << 32 | w[0]) >> 11) * 2.0 ** -53
    assert u1 == ((w[3] << 32 | w[2]) >> 11) * 2.0 ** -53


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(tag, L=(4, 4, 6), **kw):
    p = make_params(L=L, **MODELS[tag], **kw)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p)
    r.set_fields(f)
    return p, o, r, f


@pytest.mark.parametrize("tag", list(MODELS))
def test_layout_tables(tag):
    need_ref(tag)
    p, o, r, _ = setup(tag, L=(4, 6, 8))
    nx, pv, nat = o.tables()
    rnx, rpv = r.tables()
    assert (nx == rnx).all() and (pv == rpv).all()
    # even sites first, natural order preserved inside a parity
    x = nat % 4
    y = (nat // 4) % 6
    z = nat // 24
    par = (x + y + z) % 2
    V = o.V
    assert (par[:V // 2] == 0).all() and (par[V // 2:] == 1).all()
    assert (np.diff(nat[:V // 2]) > 0).all() and (np.diff(nat[V // 2:]) > 0).all()


@pytest.mark.parametrize("tag", list(MODELS))
def test_deterministic_pieces(tag):
    need_ref(tag)
    p, o, r, _ = setup(tag)
    for i in range(o.V):
        for d in range(3):
            a = o.vec("orc_link_staple", 4, i, d)
            b = r.vec("ref_link_staple", 4, i, d)
            assert rel(a, b) < 1e-12
            a = o.vec("orc_wilson_staple", 4, i, d)
            b = r.vec("ref_wilson_staple", 4, i, d)
            assert rel(a, b) < 1e-12
            assert rel(o.lib.orc_localact_su2link(o.h, i, d), r.lib.ref_localact_su2link(i, d)) < 1e-12
            if p.u1:
                assert rel(o.lib.orc_localact_u1link(o.h, i, d), r.lib.ref_localact_u1link(i, d)) < 1e-12
        assert rel(o.lib.orc_action_local(o.h, i), r.lib.ref_action_local(i)) < 1e-12
        if p.nhiggs:
            a = o.vec("orc_doublet_staple", 4, i)
            b = r.vec("ref_doublet_staple", 4, i)
            assert rel(a, b) < 1e-12
            assert rel(o.lib.orc_localact_doublet(o.h, i), r.lib.ref_localact_doublet(i)) < 1e-12
        for hh in range(1, p.nhiggs):
            a = o.vec("orc_doublet_staple_h", 4, i, hh)
            b = r.vec("ref_doublet_staple_h", 4, i, hh)
            assert rel(a, b) < 1e-12
            assert rel(o.lib.orc_localact_doublet_h(o.h, i, hh),
                       r.lib.ref_localact_doublet_h(i, hh)) < 1e-12
        if p.triplet:
            assert rel(o.lib.orc_localact_triplet(o.h, i), r.lib.ref_localact_triplet(i)) < 1e-12
        if p.singlet:
            assert rel(o.lib.orc_localact_singlet(o.h, i), r.lib.ref_localact_singlet(i)) < 1e-12


def kinds_for(p):
    """(kind, number of values of its dir argument): directions for links,
    doublets for the doublet kinds"""
    ks = [(K_LINK_HEATBATH, 3), (K_LINK_METROPOLIS, 3)]
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
def test_site_updates(tag):
    """Every per-site update, site by site, same draws, same result."""
    need_ref(tag)
    p, o, r, f = setup(tag)
    pas = 11
    for kind, ndir in kinds_for(p):
        for i in range(0, o.V, 5):
            d = (i // 5) % ndir
            a = o.site(kind, i, d, pas)
            b = r.site(kind, i, d, pas)
            assert a == b, (kind, i)
            pas += 1
        fo, fr = o.get_fields(), r.get_fields()
        for k in orc.FIELDS:
            assert np.abs(fo[k] - fr[k]).max() < 1e-12, (kind, k)
        # resynchronise rounding before the next kind
        r.set_fields(fo)


@pytest.mark.parametrize("tag", list(MODELS))
@pytest.mark.parametrize("variant", ["or", "metro"])
def test_full_sweeps(tag, variant):
    """The reference's own update_lattice against orc_update."""
    need_ref(tag)
    kw = {}
    if variant == "metro":
        kw = dict(algorithm_su2link=METROPOLIS, algorithm_su2doublet=METROPOLIS,
                  algorithm_su2triplet=METROPOLIS, algorithm_singlet=METROPOLIS,
                  scalar_sweeps=2, update_links=2)
    p, o, r, f = setup(tag, **kw)
    o.rng_set(99, 0)
    r.rng_set(99, 0)
    o.update(3)
    r.update(3)
    fo, fr = o.get_fields(), r.get_fields()
    # the two-doublet overrelaxation keeps Y and moves X to the other root
    # (src/overrelax.c:295-417); that map amplifies a rounding difference
    # 10-30x per sweep (test_higgs2_rounding_growth), so 3 sweeps from 1e-15
    # per-site agreement end near 1e-10
    tol = 1e-11 if p.nhiggs < 2 else 1e-9
    for k in orc.FIELDS:
        assert np.abs(fo[k] - fr[k]).max() < tol, k
    co, cr = o.counters(), r.counters()
    for k in cr:
        if k in ("total_link_or", "accepted_link_or"):
            continue
        assert co[k] == cr[k], k
    assert o.rng_get()[1] == r.lib.ref_rng_pass()


@pytest.mark.parametrize("tag", ["h", "hu", "t", "d"])
def test_random_order_sweeps(tag):
    """The reference's random_sweeps (shuffled link passes, src/update.c:
    292-300, a random first parity per scalar round, :341-345) against the
    oracle's random order: same shuffles from the sweep's schedule draws,
    same passes, same fields."""
    need_ref(tag)
    p, o, r, f = setup(tag, link_overrelax=0)
    o.rng_set(7, 0)
    r.rng_set(7, 0)
    o.lib.orc_set_random_order(o.h, 1)
    r.lib.ref_set_random_sweeps(1)
    o.update(3)
    r.update(3)
    fo, fr = o.get_fields(), r.get_fields()
    tol = 1e-11 if p.nhiggs < 2 else 1e-9
    for k in orc.FIELDS:
        if fr[k] is not None and np.abs(fr[k]).max() > 0:
            assert np.abs(fo[k] - fr[k]).max() < tol, k
    co, cr = o.counters(), r.counters()
    for k in cr:
        if k not in ("total_link_or", "accepted_link_or"):
            assert co[k] == cr[k], k
    # the order really changed: the same start without it ends elsewhere
    o2 = orc.Oracle(p)
    o2.set_fields(f)
    o2.rng_set(7, 0)
    o2.update(3)
    assert np.abs(o2.get_fields()["su2link"] - fo["su2link"]).max() > 1e-3


def test_shuffle_draws_are_the_iran_words():
    """iran() is the full 64-bit word of a draw and dran() its top 53 bits,
    as genrand64_real2 derives from genrand64_int64."""
    lib = orc.oracle_lib()
    for k in range(8):
        w = lib.orc_word(11, 5, 0xFFFFFFFE, k)
        assert orc.uniform(11, 5, 0xFFFFFFFE, k) == (w >> 11) * 2.0 ** -53


def test_higgs2_rounding_growth():
    """The tolerance of the two-doublet sweep comparison: a 1e-14 relative
    change of the doublets grows by more than 10x per sweep under the
    oracle's own two-doublet dynamics, but stays near 1e-13 with one
    doublet."""
    def growth(**kw):
        p = make_params(L=(4, 4, 6), **kw)
        a = orc.Oracle(p)
        a.hot_start()
        f = a.get_fields()
        b = orc.Oracle(p)
        g = dict(f)
        g["doublet"] = f["doublet"] * (1 + 1e-14)
        b.set_fields(g)
        a.rng_set(99, 0)
        b.rng_set(99, 0)
        out = []
        for _ in range(4):
            a.update(1)
            b.update(1)
            fa, fb = a.get_fields(), b.get_fields()
            out.append(max(np.abs(fa[k] - fb[k]).max() for k in orc.FIELDS if fa[k] is not None))
        return out
    one = growth(nhiggs=1)
    two = growth(nhiggs=2)
    assert max(one) < 1e-12
    assert two[3] > 100 * two[1] > 0


@pytest.mark.parametrize("tag", list(MODELS))
def test_measure_sums(tag):
    need_ref(tag)
    p, o, r, _ = setup(tag, L=(6, 4, 4))
    a, b = o.measure(), r.measure()
    scale = np.maximum(np.abs(b), 1.0)
    assert (np.abs(a - b) / scale).max() < 1e-12


def test_polysolve3_against_reference():
    need_ref("h")
    lib = orc.ref_lib("h")
    olib = orc.oracle_lib()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        a = rng.uniform(0.001, 0.05)
        b = rng.normal() * a
        c = rng.uniform(0.5, 5.0)
        d = rng.normal() * 5
        x = olib.orc_polysolve3(a, b, c, d)
        y = lib.ref_polysolve3(a, b, c, d)
        assert x == y or abs(x - y) < 1e-12 * max(1.0, abs(y))
        # and it is a root
        assert abs(((a * x + b) * x + c) * x + d) < 1e-8 * (abs(d) + c * abs(x) + 1)


CORR_MODELS = {"hc": dict(nhiggs=1), "tc": dict(nhiggs=0, triplet=1),
               "htc": dict(nhiggs=1, triplet=1)}


@pytest.mark.parametrize("tag", list(CORR_MODELS))
@pytest.mark.parametrize("L", [(4, 6, 8), (6, 6, 6), (8, 4, 6)])
def test_correlators(tag, L):
    """measure_correlators of the CORRELATORS build against the oracle on a
    hot start, along every direction (the photon operators carry momentum
    only when the direction is not 0)."""
    if not orc.ref_available(tag):
        pytest.skip("reference build absent")
    p = make_params(L=L, **CORR_MODELS[tag], seed=31)
    o = orc.Oracle(p)
    o.hot_start()
    r = orc.Ref(p, correlators=True)
    r.set_fields(o.get_fields())
    assert o.lib.orc_longest_dir(o.h) == r.lib.ref_longest_dir()
    for d in range(3):
        a, b = o.correlators(d), r.correlators(d)
        assert a.shape == b.shape == (L[d] // 2, 10)
        scale = np.abs(b).max(0) + 1e-300
        assert (np.abs(a - b) <= 1e-12 * scale).all(), (d, np.abs(a - b).max(0) / scale)
        if p.triplet:
            assert np.abs(b[:, 2:]).max() > 0


@pytest.mark.parametrize("L", [(8, 8, 12), (8, 12, 8), (16, 8, 8)])
def test_blocked_correlators(L):
    """Correlators on the blocked lattices of the BLOCKING build (smeared
    links and triplet, every direction but the longest halved per level)
    against the oracle."""
    if not orc.ref_available("tcb"):
        pytest.skip("reference build absent")
    p = make_params(L=L, nhiggs=0, triplet=1, seed=33)
    o = orc.Oracle(p)
    o.hot_start()
    r = orc.Ref(p, blocking=True)
    r.set_fields(o.get_fields())
    levels = r.lib.ref_blocked_setup(5)
    d = r.lib.ref_longest_dir()
    ref = np.zeros((levels, L[d] // 2, 10))
    assert r.lib.ref_blocked_values(orc.ptr(ref)) == levels >= 1
    ours = o.blocked_correlators(5)
    assert ours.shape == ref.shape
    scale = np.abs(ref).max(1, keepdims=True) + 1e-300
    assert (np.abs(ours - ref) <= 1e-12 * scale).all(), np.abs(ours - ref).max((0, 1))


def test_flow_with_two_doublets(tmp_path):
    """grad_flow of the -DNHIGGS=2 -DU1 -DGRADFLOW build: the links flow,
    both doublets stay, measure_flow rows through the two-doublet columns."""
    need_ref("dug")
    import ctypes as C
    import os
    from su2higgs_amd import abi
    p = make_params(L=(4, 4, 6), nhiggs=2, u1=1, seed=77)
    o = orc.Oracle(p)
    o.hot_start()
    r = orc.Ref(p, gradflow=True)
    r.set_fields(o.get_fields())
    F, _ = o.flow_force()
    G = np.zeros((o.V, 3, 4))
    r.lib.ref_flow_force(orc.ptr(G), None)
    assert rel(F, G) < 1e-12
    os.chdir(tmp_path)
    r.lib.ref_grad_flow(0.25, 0.05, 2, 1)
    rows = o.flow(abi.FLOW_EULER, 0.05, 0.25, 2)
    text = open("measure_flow").read().strip().splitlines()[1:]
    assert len(text) == len(rows)
    lib = abi.load()
    for line, row in zip(text, rows):
        vals = [float(x) for x in line.split()]
        cols = (C.c_double * 64)()
        n = C.c_int()
        lib.su2hip_measure_columns(C.byref(p), orc.ptr(np.ascontiguousarray(row[1:])), -0.0, 0.0,
                                   cols, C.byref(n))
        assert len(vals) == 1 + n.value
        for a, b in zip(vals[1:], list(cols)[: n.value]):
            assert abs(a - b) <= 1e-5 * max(1.0, abs(b)), (a, b)
