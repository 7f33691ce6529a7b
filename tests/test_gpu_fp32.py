"""FP32-mixed sweeps: FP32 fields and staples with FP64 reductions.

There is no per-site oracle for FP32 (different draw mapping and rounding);
these tests pin it by invariants and by agreement with the FP64 path within
statistical errors."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from su2higgs_amd import abi
from su2higgs_amd.abi import FP32, FP64, make_params
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("links", [abi.LINKS_FULL, abi.LINKS_PACKED], ids=["full", "packed"])
def test_fp32_roundtrip_rounds_to_float(links):
    """Scalar fields and angles come back as their float rounding, and so
    do links in the full form.  Packed links come back as the unit
    quaternion U/|U| to a few units in the last place."""
    p = make_params(L=(8, 8, 8), nhiggs=1, u1=1, precision=FP32, link_format=links)
    b = Batch(p)
    rng = np.random.default_rng(1)
    f = abi.field_arrays(p)
    f["su2link"][:] = rng.normal(size=f["su2link"].shape)
    f["doublet"][:] = rng.normal(size=f["doublet"].shape)
    f["u1link"][:] = rng.normal(size=f["u1link"].shape)
    b.set_fields(0, f)
    g = b.get_fields(0)
    for k in ("doublet", "u1link") + (("su2link",) if links == abi.LINKS_FULL else ()):
        assert np.array_equal(g[k], f[k].astype(np.float32).astype(np.float64)), k
    if links == abi.LINKS_PACKED:
        unit = f["su2link"] / np.linalg.norm(f["su2link"], axis=-1, keepdims=True)
        assert np.abs(g["su2link"] - unit).max() < 4e-7
        assert np.abs(np.linalg.norm(g["su2link"], axis=-1) - 1).max() < 4e-7


def slot_links(b, slot=0):
    p = abi.Params()
    abi.check(b.lib.su2hip_slot_params(b.h, slot, C.byref(p)))
    return p.link_format


def test_link_format_resolution():
    """AUTO takes the full form below SU2HIP_LINKS_PACKED_ABOVE bytes of FP32
    fields (in the full form) and the packed form above it, and packed for a
    handle created empty; FP64 is always full; an explicit choice is kept."""
    small = make_params(L=(32, 32, 32), nhiggs=1, u1=1, precision=FP32)   # 2.5 MB
    mid = make_params(L=(64, 64, 64), nhiggs=1, precision=FP32)            # 16.8 MB
    big = make_params(L=(96, 96, 96), nhiggs=1, u1=1, precision=FP32)      # 67 MB
    assert slot_links(Batch(small)) == abi.LINKS_FULL
    assert slot_links(Batch(mid)) == abi.LINKS_FULL
    assert slot_links(Batch(big)) == abi.LINKS_PACKED
    assert slot_links(Batch([small] * 12)) == abi.LINKS_FULL              # 29.9 MB together
    assert slot_links(Batch([small] * 14)) == abi.LINKS_PACKED            # 34.9 MB together
    assert slot_links(Batch(make_params(L=(128,) * 3, precision=FP64))) == abi.LINKS_FULL
    forced = make_params(L=(8, 8, 8), precision=FP32, link_format=abi.LINKS_PACKED)
    assert slot_links(Batch(forced)) == abi.LINKS_PACKED
    lib = abi.load()
    h = C.c_void_p()
    abi.check(lib.su2hip_create(0, C.byref(small), 2, 0, C.byref(h)))
    slot = C.c_int()
    abi.check(lib.su2hip_add(h, C.byref(small), C.byref(slot)))
    p = abi.Params()
    abi.check(lib.su2hip_slot_params(h, slot.value, C.byref(p)))
    assert p.link_format == abi.LINKS_PACKED
    lib.su2hip_destroy(h)


@pytest.mark.parametrize("kind", [abi.K_LINK_OVERRELAX, abi.K_DOUBLET_OVERRELAX],
                         ids=["link_or", "doublet_or"])
def test_packed_and_full_links_make_the_same_move(kind):
    """The same deterministic overrelaxation pass on the same fields in the
    two FP32 link forms agrees to FP32 rounding (the packed form rebuilds
    every link from its projection, to a few units in the last place)."""
    import orc
    p64 = make_params(L=(8, 8, 12), nhiggs=1, u1=1, betau1=9.0, seed=5)
    o = orc.Oracle(p64)
    o.hot_start()
    f = o.get_fields()
    out = {}
    for links in (abi.LINKS_FULL, abi.LINKS_PACKED):
        b = Batch(make_params(L=(8, 8, 12), nhiggs=1, u1=1, betau1=9.0, seed=5, precision=FP32,
                              link_format=links))
        b.set_fields(0, f)
        b.pass_(kind, 1, 2)
        out[links] = b.get_fields(0)
    for k in ("su2link", "doublet"):
        d = np.abs(out[abi.LINKS_FULL][k] - out[abi.LINKS_PACKED][k])
        assert d.max() < 3e-6, (k, d.max())


def test_fp32_packed_links_every_position_and_sign():
    """The packed FP32 link is the stereographic projection of the unit
    quaternion from the pole opposite its real part's sign, that sign in the
    lowest mantissa bit: unit quaternions with either sign of the real part
    and each component dominant in turn, ties among equal magnitudes, the
    poles (a = +-1), the equator (a = 0) and the axes come back to a few
    units in the last place with the sign of every component kept."""
    p = make_params(L=(4, 4, 4), nhiggs=1, precision=FP32, link_format=abi.LINKS_PACKED)
    V = abi.volume(p)
    rng = np.random.default_rng(7)
    q = rng.normal(size=(V, 3, 4))
    fixed = np.array([[1, 0, 0, 0], [-1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1],
                      [0, 0, 0, -1], [.5, .5, .5, .5], [-.5, .5, -.5, .5], [.5, -.5, .5, -.5],
                      [0, -.5, -.5, -.5 * np.sqrt(2)], [np.sqrt(.5), -np.sqrt(.5), 0, 0],
                      [1e-30, 1, -1e-30, 0]], dtype=float)
    q.reshape(-1, 4)[: len(fixed)] = fixed
    # each component dominant in turn, with each sign, many times over
    big = rng.integers(0, 4, size=(V, 3))
    sign = rng.choice([-1.0, 1.0], size=(V, 3))
    rest = q.reshape(-1, 4)[len(fixed):]
    b_, s_ = big.reshape(-1)[len(fixed):], sign.reshape(-1)[len(fixed):]
    rest[np.arange(len(rest)), b_] = s_ * (np.abs(rest).max(1) + 0.1)
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    f = abi.field_arrays(p)
    f["su2link"][:] = q
    b = Batch(p)
    b.set_fields(0, f)
    g = b.get_fields(0)["su2link"]
    assert np.abs(g - q).max() < 4e-7
    nz = np.abs(q) > 1e-6
    assert (np.sign(g[nz]) == np.sign(q[nz])).all()
    # a second round trip moves nothing beyond rounding
    f["su2link"][:] = g
    b.set_fields(0, f)
    assert np.abs(b.get_fields(0)["su2link"] - g).max() < 2.5e-7


@pytest.mark.parametrize("links", [abi.LINKS_FULL, abi.LINKS_PACKED], ids=["full", "packed"])
def test_fp32_links_stay_unitary_and_angles_wrapped(links):
    p = make_params(L=(16, 16, 16), nhiggs=1, u1=1, precision=FP32, link_overrelax=4,
                    betasu2=9.0, betau1=27.0, msq_phi=-0.2, lambda_phi=0.015, link_format=links)
    b = Batch(p)
    b.hot_start(0)
    b.use_graph(True)
    b.update(200)
    f = b.get_fields(0)
    n = np.sqrt((f["su2link"] ** 2).sum(-1))
    assert np.abs(n - 1).max() < 5e-6
    assert np.abs(f["u1link"]).max() <= np.pi * (1 + 1e-6)


# FP64 values of the 48^3 SU(2)xU(1)+Higgs theory at betaG = 9, betaY = 27,
# r = 1, msq = -0.20, lambda = 0.015 (the single-lattice workload's
# couplings): 512 independent FP64 chains of the LEAN sweep mix below, cold
# start, 16000 sweeps each measured every 10, the first 4000 discarded (the
# protocol of the FP32 chains here, so both precisions are compared at the
# same Monte Carlo time from the same start); each value is the mean over
# chains of the chain's time average, its error the chain-to-chain spread /
# sqrt(512).  Made with tests/fp64_reference_48.py.
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp64_reference_48.json")) as _f:
    FP64_REF_48 = {k: tuple(v) for k, v in json.load(_f).items()}


def lean48(precision, seed):
    """The workload's couplings at 48^3 with a cheap sweep mix (heatbath,
    U(1) Metropolis, one doublet overrelaxation round and the Metropolis
    round): the same equilibrium distribution as the full mix at a third of
    the cost per U(1) update."""
    return make_params(L=(48, 48, 48), nhiggs=1, u1=1, precision=precision, betasu2=9.0,
                       betau1=27.0, r_u1=1.0, msq_phi=-0.20, lambda_phi=0.015, link_overrelax=0,
                       scalar_sweeps=1, update_su2doublet=1, update_links=1,
                       reunitarize=1 if precision == FP32 else 0, seed=seed)


def chain_series(precision, chains, therm, sweeps, every, seed0):
    """[samples, chains] series of the plaquette, <phi^+phi> and the U(1)
    plaquette of independent chains swept in one batch from a cold start."""
    ps = [lean48(precision, seed0 + 17 * k) for k in range(chains)]
    b = Batch(ps)
    for k in range(chains):
        b.cold_start(k, phi0=1.0)
    b.use_graph(True)
    b.update(therm)
    V = float(abi.volume(ps[0]))
    rows = []
    for _ in range(sweeps // every):
        b.update(every)
        rows.append(b.measure_all()[:chains])
    b.close()
    s = np.array(rows)
    return {"plaquette": 1 - s[:, :, abi.SUM["wilson"]] / (V * 9.0) / 3,
            "phi2": s[:, :, abi.SUM["phi2"]] / V,
            "u1_plaquette": 1 - s[:, :, abi.SUM["u1wilson"]] / (V * 27.0) / 3}


def chain_mean(x):
    """Mean over chains of each chain's time average and its error from
    the chain-to-chain spread (independent chains: no autocorrelation
    model needed)."""
    m = x.mean(0)
    return m.mean(), m.std(ddof=1) / np.sqrt(len(m))


def agrees(mean, err, ref, ref_err, nsigma=3.5):
    return abs(mean - ref) < nsigma * np.hypot(err, ref_err)


def test_fp32_observables_match_fp64_over_many_chains():
    """128 FP32-mixed chains against the FP64 values above: plaquette,
    <phi^+phi> and the U(1) plaquette within 3.5 combined standard errors,
    each after a plateau check (the two halves of the sampling agree).  The
    U(1) plaquette is measured to ~1e-7: the same check fails on the FP32
    value shifted by 1e-6 either way, so a bias of that size could not pass.
    32 FP64 chains are run beside them to check the stored values."""
    x32 = chain_series(FP32, 128, 4000, 12000, 10, seed0=32001)
    for name, (ref, ref_err) in FP64_REF_48.items():
        m, e = chain_mean(x32[name])
        h = len(x32[name]) // 2
        m1, e1 = chain_mean(x32[name][:h])
        m2, e2 = chain_mean(x32[name][h:])
        assert agrees(m1, e1, m2, e2), (name, "no plateau", m1, e1, m2, e2)
        assert agrees(m, e, ref, ref_err), (name, m, e, ref, ref_err)
        if name == "u1_plaquette":
            assert np.hypot(e, ref_err) < 2.0e-7, e
            for shift in (1e-6, -1e-6):
                assert not agrees(m + shift, e, ref, ref_err), (shift, m, e)
    x64 = chain_series(FP64, 32, 4000, 6000, 10, seed0=640001)
    for name, (ref, ref_err) in FP64_REF_48.items():
        m, e = chain_mean(x64[name])
        assert agrees(m, e, ref, ref_err), (name, "FP64 check of the stored value", m, e)


def adjoint_series(precision, chains, therm, sweeps, every, seed0, L=24):
    """[samples, chains] series of the plaquette, the triplet moments
    <Sigma^2/2> and <(Sigma^2/2)^2> per site and the hopping term per link
    of independent SU(2)+adjoint chains at the
    multicanonical workload's couplings (betaG = 10, msq = -0.12, b4 = 0.01),
    swept in one batch from a cold start with sigma = 0.5."""
    ps = [make_params(L=(L, L, L), nhiggs=0, triplet=1, precision=precision, betasu2=10.0,
                      msq_triplet=-0.12, b4=0.01, a2=0.0, link_overrelax=4, scalar_sweeps=4,
                      reunitarize=1 if precision == FP32 else 0, seed=seed0 + 13 * k)
          for k in range(chains)]
    b = Batch(ps)
    for k in range(chains):
        b.cold_start(k, sigma0=0.5)
    b.use_graph(True)
    b.update(therm)
    V = float(abi.volume(ps[0]))
    rows = []
    for _ in range(sweeps // every):
        b.update(every)
        rows.append(b.measure_all()[:chains])
    b.close()
    s = np.array(rows)
    return {"plaquette": 1 - s[:, :, abi.SUM["wilson"]] / (V * 10.0) / 3,
            "sigma2": s[:, :, abi.SUM["sigma2"]] / V,
            "sigma4": s[:, :, abi.SUM["sigma4"]] / V,
            "hopping": s[:, :, abi.SUM["hop_sigma"]] / (3 * V)}


def test_fp32_adjoint_observables_match_fp64_over_many_chains():
    """SU(2)+adjoint, 24^3: 64 FP32-mixed chains against 64 FP64 chains
    (independent keys), each observable's mean over chains within 3.5
    combined standard errors after a plateau check in both precisions.
    The errors are small enough that a shift of 1e-4 in the plaquette or of
    0.2% in a triplet moment would fail the same check."""
    x32 = adjoint_series(FP32, 64, 1000, 4000, 10, seed0=240001)
    x64 = adjoint_series(FP64, 64, 1000, 4000, 10, seed0=640001)
    for name in ("plaquette", "sigma2", "sigma4", "hopping"):
        for x in (x32[name], x64[name]):
            h = len(x) // 2
            m1, e1 = chain_mean(x[:h])
            m2, e2 = chain_mean(x[h:])
            assert agrees(m1, e1, m2, e2), (name, "no plateau", m1, e1, m2, e2)
        m32, e32 = chain_mean(x32[name])
        m64, e64 = chain_mean(x64[name])
        assert agrees(m32, e32, m64, e64), (name, m32, e32, m64, e64)
        shift = 1e-4 if name == "plaquette" else 2e-3 * abs(m64)
        for sh in (shift, -shift):
            assert not agrees(m32 + sh, e32, m64, e64), (name, "too weak to see", sh, e32, e64)


def test_fp32_heatbath_distribution():
    """a0 = (U W)_0 / |W| of freshly heatbathed links follows
    sqrt(1-a0^2) exp(2 k a0) (Kennedy-Pendleton); checked on a cold lattice
    where every staple has the same norm."""
    from scipy import stats
    p = make_params(L=(16, 16, 16), nhiggs=0, precision=FP32, betasu2=1.5)
    b = Batch(p)
    b.cold_start(0)
    b.pass_(abi.K_LINK_HEATBATH, 0, 0)
    f = b.get_fields(0)
    # even sites, direction 0: staple of a cold lattice = -beta/2 * 4 * 1
    a0 = f["su2link"][: abi.volume(p) // 2, 0, 0]
    k = 0.5 * p.betasu2 * 4  # |staple|; weight exp(+2k a0) after the sign flip
    xs = np.linspace(-1, 1, 20001)
    pdf = np.sqrt(1 - xs ** 2) * np.exp(2 * k * (xs - 1))
    cdf = np.cumsum(pdf)
    cdf /= cdf[-1]
    stat = stats.kstest(a0, lambda x: np.interp(x, xs, cdf))
    assert stat.pvalue > 1e-4, stat


@pytest.mark.parametrize("links", [abi.LINKS_FULL, abi.LINKS_PACKED], ids=["full", "packed"])
def test_fp32_heatbath_direction_is_isotropic(links):
    """The FP32 heatbath draws the direction of a = (a0, a) on S^2 directly
    (cos theta uniform in [-1, 1], phi uniform) instead of by the
    reference's cube rejection: on a cold lattice (every staple the same)
    the directions of the new links pass KS tests for uniform cos theta and
    uniform phi, and their mean vector is zero within its error."""
    from scipy import stats
    p = make_params(L=(32, 32, 32), nhiggs=0, precision=FP32, betasu2=1.5, link_format=links)
    b = Batch(p)
    b.cold_start(0)
    b.pass_(abi.K_LINK_HEATBATH, 0, 0)
    u = b.get_fields(0)["su2link"][: abi.volume(p) // 2, 0, 1:]
    n = u / np.linalg.norm(u, axis=1, keepdims=True)
    assert stats.kstest(n[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-4
    phi = np.arctan2(n[:, 1], n[:, 0])
    assert stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-4
    m = n.mean(0)
    assert (np.abs(m) < 4.5 * np.sqrt(1 / 3 / len(n))).all(), m


MOVES = [
    ("doublet", dict(nhiggs=1), abi.K_DOUBLET_OVERRELAX, 0, "doublet"),
    ("doublet_u1", dict(nhiggs=1, u1=1), abi.K_DOUBLET_OVERRELAX, 1, "doublet"),
    ("triplet", dict(nhiggs=0, triplet=1), abi.K_TRIPLET_OVERRELAX, 0, "triplet"),
    ("link_or_u1", dict(nhiggs=1, u1=1), abi.K_LINK_OVERRELAX, 1, "su2link"),
]


@pytest.mark.parametrize("name,model,kind,parity,field", MOVES, ids=[m[0] for m in MOVES])
def test_fp32_overrelaxation_is_the_fp64_move(name, model, kind, parity, field):
    """One FP32 overrelaxation pass against the FP64 oracle's pass on the
    same (float-rounded) fields: the deterministic move agrees to FP32
    accuracy wherever both accept (the accept draws differ; acceptance is
    close to 1)."""
    import orc
    p64 = make_params(L=(8, 8, 12), precision=FP64, msq_phi=-0.2, lambda_phi=0.02,
                      msq_triplet=-0.3, b4=0.1, betau1=9.0, seed=5, **model)
    o = orc.Oracle(p64)
    o.hot_start()
    f = {k: (v.astype(np.float32).astype(np.float64) if v is not None else None)
         for k, v in o.get_fields().items()}
    o.set_fields(f)
    p32 = make_params(L=(8, 8, 12), precision=FP32, msq_phi=-0.2, lambda_phi=0.02,
                      msq_triplet=-0.3, b4=0.1, betau1=9.0, seed=5, reunitarize=0, **model)
    b = Batch(p32)
    b.set_fields(0, f)
    b.pass_(kind, parity, 1)
    g = b.get_fields(0)[field]
    o.pass_(kind, parity, 1)
    h = o.get_fields()[field]
    V = abi.volume(p64)
    sl = slice(parity * V // 2, (parity + 1) * V // 2)
    g, h, f0 = g[sl].reshape(V // 2, -1), h[sl].reshape(V // 2, -1), f[field][sl].reshape(V // 2, -1)
    scale = np.abs(h).max(1) + 1.0
    same = np.abs(g - h).max(1) <= 2e-4 * scale
    moved = np.abs(h - f0).max(1) > 1e-3 * scale
    assert moved.mean() > 0.9          # the oracle's move changes almost every variable
    assert same.mean() > 0.9, same.mean()
    # where they differ, one of the two rejected and kept the old value
    # (the two accept draws are independent; acceptance is 0.95-1 here)
    diff = ~same
    kept = (np.abs(g - f0).max(1) == 0) | (np.abs(h - f0).max(1) == 0)
    assert kept[diff].all()
