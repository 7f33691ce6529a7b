"""The benchmark's 96^3 SU(2)xU(1)+Higgs lattice (beta 9, beta_Y 27,
m^2 -0.20, lambda 0.015) at its full size.

The oracle finishes a pass over 96^3 in well under a second, so the HIP path
is compared with it here directly (FP64, same Philox draws): every pass kind
in every direction and parity, and one whole compound sweep.  Beside that,
properties that hold at any size: link overrelaxation keeps the total action
and undoes itself, graph replay and eager launches give the same bits, and
the links stay on SU(2).  The batched scan's 96 lattices (16^3..48^3) at
their full shape: each sampled lattice bitwise the same alone and in the
mix (FP32), and its smallest and largest lattices against the oracle (FP64).

Tolerances as in test_gpu_parity: fields to rounding (1e-11 absolute on O(1)
values after single passes, 1e-10 after a sweep), sums 1e-12 relative,
counters exact."""
import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import (FP32, FP64, K_DOUBLET_METROPOLIS, K_DOUBLET_OVERRELAX,
                              K_LINK_HEATBATH, K_LINK_OVERRELAX, K_U1_METROPOLIS, make_params)
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu

FIELDS = ("su2link", "u1link", "doublet")


def params(precision=FP64, **kw):
    return make_params(L=(96, 96, 96), nhiggs=1, u1=1, precision=precision, betasu2=9.0,
                       betau1=27.0, msq_phi=-0.2, lambda_phi=0.015, link_overrelax=4,
                       scalar_sweeps=4, seed=96, **kw)


def maxdiff(a, b):
    return max(np.abs(a[k] - b[k]).max() for k in FIELDS)


def hot_pair(p):
    o = orc.Oracle(p)
    o.hot_start()
    b = Batch(p)
    b.set_fields(0, o.get_fields())
    return o, b


def test_every_pass_kind_matches_oracle_at_96cubed():
    p = params()
    o, b = hot_pair(p)
    kinds = [(K_LINK_HEATBATH, 3), (K_LINK_OVERRELAX, 3), (K_U1_METROPOLIS, 3),
             (K_DOUBLET_OVERRELAX, 1), (K_DOUBLET_METROPOLIS, 1)]
    for kind, ndir in kinds:
        for d in range(ndir):
            for par in (0, 1):
                o.pass_(kind, par, d)
                b.pass_(kind, par, d)
        fo = o.get_fields()
        assert maxdiff(fo, b.get_fields(0)) < 1e-11, abi.KIND_NAMES[kind]
        # resynchronise rounding before the next kind
        b.set_fields(0, fo)
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    assert o.rng_get()[1] == b.rng_get(0)[1]
    so, sg = o.measure(), b.measure(0)
    assert (np.abs(so - sg) / np.maximum(np.abs(so), 1.0)).max() < 1e-12
    b.close()


def test_fp64_compound_sweep_matches_oracle_at_96cubed():
    """1 heatbath round + 4 link-overrelaxation rounds + U(1) Metropolis +
    4 doublet-overrelaxation rounds + 1 doublet-Metropolis round: 46 passes."""
    p = params()
    o, b = hot_pair(p)
    o.update(1)
    b.update(1)
    assert maxdiff(o.get_fields(), b.get_fields(0)) < 1e-10
    co, cg = o.counters(), b.counters(0)
    for k in co:
        assert co[k] == cg[k], k
    b.close()


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (FP32, 2e-5)])
def test_link_overrelaxation_keeps_action_and_undoes_itself(precision, tol):
    """U -> W^+ U^+ W^+ keeps each link's local action, so the total action
    is unchanged; the staple W does not contain the link, so a second
    application restores it."""
    p = params(precision, reunitarize=0)
    o = orc.Oracle(params())
    o.hot_start()
    b = Batch(p)
    b.set_fields(0, o.get_fields())
    b.update(2)  # leave the hot start: the action is O(V), links near their staples
    s0 = b.measure(0)[abi.SUM["action"]]
    for d in range(3):
        f0 = b.get_fields(0)
        b.pass_(K_LINK_OVERRELAX, 0, d)
        s1 = b.measure(0)[abi.SUM["action"]]
        f1 = b.get_fields(0)
        assert np.abs(f1["su2link"] - f0["su2link"]).max() > 0.1  # the links moved
        assert abs(s1 - s0) < tol * abs(s0), (d, s0, s1)
        b.pass_(K_LINK_OVERRELAX, 0, d)
        f2 = b.get_fields(0)
        assert np.abs(f2["su2link"] - f0["su2link"]).max() < (1e-12 if precision == FP64 else 1e-5)
    b.close()


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_graph_replay_equals_eager_and_links_stay_unitary(precision):
    p = params(precision)
    res = []
    for graph in (True, False):
        b = Batch(p)
        b.cold_start(0, phi0=1.0)
        b.use_graph(graph)
        b.update(3)
        b.sync()
        res.append(b.get_fields(0))
        b.close()
    for k in FIELDS:
        assert np.array_equal(res[0][k], res[1][k]), k
    u = res[0]["su2link"].reshape(-1, 4)
    norm = (u * u).sum(1)
    assert np.abs(norm - 1).max() < (1e-12 if precision == FP64 else 1e-5)
    assert np.isfinite(res[0]["doublet"]).all()
    assert np.abs(res[0]["u1link"]).max() <= np.pi * (1 + 1e-6) or precision == FP64


def scan_params(precision):
    """The batched scan's 96 lattices: volumes drawn from {16^3, 24^3, 32^3,
    48^3}, beta 8, lambda 0.02, m^2 spaced across [-0.30, -0.15], a Philox key
    per lattice."""
    rng = np.random.default_rng(96)
    Ls = rng.choice([16, 24, 32, 48], size=96)
    return [make_params(L=(int(L),) * 3, nhiggs=1, precision=precision, betasu2=8.0,
                        lambda_phi=0.02, msq_phi=-0.30 + 0.15 * k / 95, link_overrelax=4,
                        scalar_sweeps=4, seed=1000 + k) for k, L in enumerate(Ls)]


@pytest.mark.parametrize("links", [abi.LINKS_PACKED, abi.LINKS_FULL], ids=["packed", "full"])
def test_batched_scan_full_shape_same_alone_and_in_mix(links):
    """FP32, the whole 96-lattice batch in one launch per pass: a lattice's
    chain is bitwise the one it follows alone in the same link storage (the
    batch alone would take the packed form, a small lattice alone the full
    one: both are forced here, each in turn)."""
    ps = scan_params(FP32)
    for p in ps:
        p.link_format = links
    b = Batch(ps)
    for i in range(len(ps)):
        b.hot_start(i)
    b.use_graph(True)
    b.update(3)
    vols = [abi.volume(p) for p in ps]
    picks = sorted({0, 95, int(np.argmax(vols)), int(np.argmin(vols))})
    mixed = {k: b.get_fields(k) for k in picks}
    b.close()
    for k in picks:
        a = Batch(ps[k])
        a.hot_start(0)
        a.update(3)
        f = a.get_fields(0)
        a.close()
        for name in ("su2link", "doublet"):
            assert np.array_equal(f[name], mixed[k][name]), (k, name)


def test_batched_scan_full_shape_fp64_matches_oracle():
    """FP64, the whole batch swept twice; its smallest and largest lattices
    against the oracle run alone from the same fields."""
    ps = scan_params(FP64)
    vols = [abi.volume(p) for p in ps]
    picks = [int(np.argmin(vols)), int(np.argmax(vols))]
    b = Batch(ps)
    oracles = {}
    for i in range(len(ps)):
        b.hot_start(i)
    for k in picks:
        o = orc.Oracle(ps[k])
        o.hot_start()
        b.set_fields(k, o.get_fields())
        oracles[k] = o
    b.update(2)
    for k in picks:
        o = oracles[k]
        o.update(2)
        fo, fg = o.get_fields(), b.get_fields(k)
        assert max(np.abs(fo[n] - fg[n]).max() for n in ("su2link", "doublet")) < 1e-10, k
        co, cg = o.counters(), b.counters(k)
        for c in co:
            assert co[c] == cg[c], (k, c)
    b.close()
