"""Invariants that pin the parts of the oracle with no reference
counterpart: SU(2) link overrelaxation, Luscher's RK3 flow and the hot
start.  CPU only."""
import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import K_LINK_OVERRELAX, make_params


@pytest.mark.parametrize("model", [dict(nhiggs=0), dict(nhiggs=1), dict(nhiggs=1, u1=1)])
def test_link_overrelax_preserves_action_and_is_an_involution(model):
    p = make_params(L=(4, 4, 6), **model)
    o = orc.Oracle(p)
    o.hot_start()
    f0 = o.get_fields()
    s0 = o.measure()[abi.SUM["action"]]
    for par in (0, 1):
        for d in range(3):
            # local action of each link unchanged
            before = [o.lib.orc_localact_su2link(o.h, i, d) for i in range(o.V)]
            o.pass_(K_LINK_OVERRELAX, par, d)
            after = [o.lib.orc_localact_su2link(o.h, i, d) for i in range(o.V)]
            off = 0 if par == 0 else o.V // 2
            for i in range(off, off + o.V // 2):
                assert abs(before[i] - after[i]) < 1e-12 * max(1.0, abs(before[i]))
    f1 = o.get_fields()
    # unit determinant
    assert np.abs((f1["su2link"] ** 2).sum(-1) - 1).max() < 1e-13
    # total action preserved
    assert abs(o.measure()[abi.SUM["action"]] - s0) < 1e-10 * abs(s0)
    # applying the same passes in reverse order undoes them
    for par in (1, 0):
        for d in (2, 1, 0):
            o.pass_(K_LINK_OVERRELAX, par, d)
    assert np.abs(o.get_fields()["su2link"] - f0["su2link"]).max() < 1e-12


def test_link_overrelax_changes_the_links():
    p = make_params(L=(4, 4, 6), nhiggs=1)
    o = orc.Oracle(p)
    o.hot_start()
    f0 = o.get_fields()["su2link"]
    o.pass_(K_LINK_OVERRELAX, 0, 0)
    assert np.abs(o.get_fields()["su2link"] - f0).max() > 0.1


def flow_links(p, f, scheme, dt, t):
    o = orc.Oracle(p)
    o.set_fields(f)
    n = int(round(t / dt))
    for _ in range(n):
        o.lib.orc_flow_step(o.h, scheme, dt)
    return o.get_fields()["su2link"]


def test_rk3_is_third_order_and_agrees_with_euler():
    p = make_params(L=(4, 4, 4), nhiggs=0, betasu2=8.0)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    t = 0.2
    exact = flow_links(p, f, abi.FLOW_RK3, 0.0025, t)
    errs = [np.abs(flow_links(p, f, abi.FLOW_RK3, dt, t) - exact).max() for dt in (0.05, 0.025)]
    order = np.log2(errs[0] / errs[1])
    assert 2.6 < order < 3.4, (errs, order)
    # Euler is first order and converges to the same flow
    e1 = np.abs(flow_links(p, f, abi.FLOW_EULER, 0.01, t) - exact).max()
    e2 = np.abs(flow_links(p, f, abi.FLOW_EULER, 0.005, t) - exact).max()
    assert 1.7 < e1 / e2 < 2.3
    # flow lowers the Wilson action
    o.set_fields(f)
    w0 = o.measure()[abi.SUM["wilson"]]
    o.lib.orc_flow_step(o.h, abi.FLOW_RK3, 0.02)
    assert o.measure()[abi.SUM["wilson"]] < w0


def test_hot_start_moments():
    p = make_params(L=(16, 16, 16), nhiggs=1, u1=1, triplet=1, singlet=1)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    # N(0,1) scalar components
    for k in ("doublet", "triplet", "singlet"):
        x = f[k].ravel()
        assert abs(x.mean()) < 5 / np.sqrt(len(x))
        assert abs(x.var() - 1) < 5 * np.sqrt(2 / len(x))
    # links uniform on SU(2): unit norm, <u_a^2> = 1/4
    u = f["su2link"].reshape(-1, 4)
    assert np.abs((u ** 2).sum(1) - 1).max() < 1e-14
    assert np.abs((u ** 2).mean(0) - 0.25).max() < 5 * np.sqrt(0.05 / len(u))
    a = f["u1link"].ravel()
    assert a.min() >= -np.pi and a.max() < np.pi
    assert abs(a.mean()) < 5 * np.pi / np.sqrt(3 * len(a))
