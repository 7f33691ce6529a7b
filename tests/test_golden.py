"""The oracle against the golden fixtures written from the reference build
(tests/golden/make_golden.py): pins the oracle where oracle/_ref is absent."""
import ctypes as C
import glob
import os
import tempfile

import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import Muca, make_params, ptr

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODELS = {
    "h": dict(nhiggs=1),
    "hu": dict(nhiggs=1, u1=1),
    "t": dict(nhiggs=0, triplet=1),
    "ht": dict(nhiggs=1, triplet=1),
    "hsu": dict(nhiggs=1, u1=1, singlet=1),
    "d": dict(nhiggs=2),
    "du": dict(nhiggs=2, u1=1),
}


def load(name):
    path = os.path.join(HERE, name + ".npz")
    if not os.path.exists(path):
        pytest.skip("fixture %s absent" % name)
    return dict(np.load(path, allow_pickle=False))


def fields(d, prefix):
    return {k: d[prefix + k] for k in orc.FIELDS}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("tag", list(MODELS))
def test_model_fixture(tag):
    g = load("model_" + tag)
    p = make_params(L=tuple(g["L"]), **MODELS[tag], seed=2024)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    for k in orc.FIELDS:
        assert np.array_equal(f[k], g["in_" + k]), k  # the hot start and Philox stream
    V = o.V
    st = np.array([[o.vec("orc_link_staple", 4, i, d) for d in range(3)] for i in range(V)])
    assert rel(st, g["link_staple"]) < 1e-12
    la = np.array([[o.lib.orc_localact_su2link(o.h, i, d) for d in range(3)] for i in range(V)])
    assert rel(la, g["localact_su2link"]) < 1e-12
    al = np.array([o.lib.orc_action_local(o.h, i) for i in range(V)])
    assert rel(al, g["action_local"]) < 1e-12
    if p.nhiggs:
        ds = np.array([o.vec("orc_doublet_staple", 4, i) for i in range(V)])
        assert rel(ds, g["doublet_staple"]) < 1e-12
    if p.nhiggs == 2:
        ds = np.array([o.vec("orc_doublet_staple_h", 4, i, 1) for i in range(V)])
        assert rel(ds, g["doublet2_staple"]) < 1e-12
        la = np.array([o.lib.orc_localact_doublet_h(o.h, i, 1) for i in range(V)])
        assert rel(la, g["localact_doublet2"]) < 1e-12
    s = o.measure()
    # fixtures of one-doublet models hold the sums before the second
    # doublet's slots were appended; those slots are 0 there
    sums = np.zeros(abi.NSUMS)
    sums[:len(g["measure_sums"])] = g["measure_sums"]
    assert (np.abs(s - sums) / np.maximum(np.abs(sums), 1)).max() < 1e-12
    # the measure line of the reference from its own sums, through our formatter
    lib = abi.load()
    buf = C.create_string_buffer(4096)
    assert lib.su2hip_measure_format(C.byref(p), ptr(sums), -0.0, 0.0, buf, 4096) == 0
    assert buf.value == g["measure_text"].tobytes()
    o.rng_set(99, 0)
    o.update(2)
    h = o.get_fields()
    # two doublets: rounding grows 10-30x per sweep under the two-doublet
    # overrelaxation (test_oracle_ref.test_higgs2_rounding_growth)
    tol = 1e-11 if p.nhiggs < 2 else 1e-9
    for k in orc.FIELDS:
        assert np.abs(h[k] - g["sweep2_" + k]).max() < tol, k
    c = o.counters()
    for name, v in zip(g["counter_names"], g["sweep2_counters"]):
        assert c[str(name)] == v, name
    assert o.rng_get()[1] == int(g["sweep2_pass"][0])
    # the checkpoint bytes of the reference, rewritten by our writer
    ref = g["lattice_bytes"].tobytes()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "lattice")
        sw = fields(g, "sweep2_")
        have = abi.field_arrays(p)
        args = [ptr(np.ascontiguousarray(sw[k])) if have[k] is not None else None for k in orc.FIELDS]
        assert lib.su2hip_lattice_write(path.encode(), C.byref(p), 3, 0.0, 0.0, *args) == 0
        assert open(path, "rb").read() == ref


def test_flow_fixture():
    g = load("flow_h")
    p = make_params(L=tuple(g["L"]), nhiggs=1, seed=2025)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    F, _ = o.flow_force()
    assert rel(F, g["force"]) < 1e-12
    o.lib.orc_flow_step(o.h, abi.FLOW_EULER, 0.02)
    assert np.abs(o.get_fields()["su2link"] - g["euler_su2link"]).max() < 1e-13
    # measure_flow text: the same rows to the 6 digits it prints
    o.set_fields(fields(g, "in_"))
    rows = o.flow(abi.FLOW_EULER, 0.05, 0.25, 2)
    text = g["measure_flow"].tobytes().decode().strip().splitlines()[1:]
    assert len(text) == len(rows)
    lib = abi.load()
    for line, row in zip(text, rows):
        vals = [float(x) for x in line.split()]
        assert abs(vals[0] - row[0]) < 1e-6
        cols = (C.c_double * 64)()
        n = C.c_int()
        lib.su2hip_measure_columns(C.byref(p), ptr(np.ascontiguousarray(row[1:])), -0.0, 0.0, cols, C.byref(n))
        for a, b in zip(vals[1:], list(cols)[: n.value]):
            assert abs(a - b) <= 1e-5 * max(1.0, abs(b)), (a, b)


def test_triplet_flow_force_fixture():
    g = load("flow_t")
    p = make_params(L=tuple(g["L"]), nhiggs=0, triplet=1, seed=2026)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    F, T = o.flow_force()
    assert rel(F, g["force"]) < 1e-12
    assert rel(T, g["tforce"]) < 1e-12


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "muca_*.npz"))))
def test_muca_fixture(path):
    name = os.path.basename(path)[:-4]
    tag = name.split("_")[1]
    g = load(name)
    p = make_params(L=tuple(g["L"]), **MODELS[tag], seed=2027)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    m = Muca.from_buffer_copy(g["muca0"].tobytes())
    o.muca_set(m, g["pos"], g["W0"])
    m0 = o.muca_get()[0]
    assert abs(m0.param_value[0] - m.param_value[0]) < 1e-14
    assert m0.update_interval == m.update_interval
    lib = abi.load()
    n = m.bins + 2
    P64 = C.POINTER(C.c_int64)
    if m.mode == abi.SLOW:
        # gsum/nsum through our weight_params reader (or its missing-file default)
        gs, ns, found = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64), C.c_int()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "weight_params")
            if "params_file" in g:
                open(path, "wb").write(g["params_file"].tobytes())
            assert lib.su2hip_weight_params_load(path.encode(), m.bins, gs.ctypes.data_as(P64),
                                                 ns.ctypes.data_as(P64), C.byref(found)) == 0
        assert found.value == ("params_file" in g)
        assert (gs[1:-1] == g["gsum0"][1:-1]).all() and (ns[1:-1] == g["nsum0"][1:-1]).all()
        o.muca_set_sums(gs, ns)
    nsw = int(g["nsweeps"][0]) if "nsweeps" in g else 3
    o.rng_set(99, 0)
    o.update(nsw)
    h = o.get_fields()
    tol = 1e-11 if p.nhiggs < 2 else 1e-9
    for k in orc.FIELDS:
        assert np.abs(h[k] - g["sweep3_" + k]).max() < tol, k
    c = o.counters()
    for nm, v in zip(g["counter_names"], g["sweep3_counters"]):
        assert c[str(nm)] == v, nm
    m3, pos, W, hits = o.muca_get()
    ref3 = Muca.from_buffer_copy(g["muca3"].tobytes())
    assert np.abs(W - g["W3"]).max() < 1e-12
    assert (hits == g["hits3"]).all()
    assert abs(m3.param_value[1] - ref3.param_value[1]) < (1e-13 if p.nhiggs < 2 else 1e-10)
    assert m3.delta == ref3.delta and m3.last_max == ref3.last_max
    assert m3.muca_count == ref3.muca_count
    if m.mode == abi.SLOW:
        g3, n3 = o.muca_sums()
        assert (g3[1:-1] == g["gsum3"][1:-1]).all() and (n3[1:-1] == g["nsum3"][1:-1]).all()
    if "weight3_file" in g and m.mode != abi.READONLY:
        # the files save_weight wrote after the last update, from our writers
        with tempfile.TemporaryDirectory() as d:
            wp = os.path.join(d, "weight")
            assert lib.su2hip_weight_save(wp.encode(), C.byref(m3), ptr(pos), ptr(W)) == 0
            assert open(wp, "rb").read() == g["weight3_file"].tobytes()
            if m.mode == abi.SLOW:
                pp = os.path.join(d, "weight_params")
                g3c, n3c = np.ascontiguousarray(g3), np.ascontiguousarray(n3)
                assert lib.su2hip_weight_params_save(pp.encode(), C.byref(m3), ptr(pos),
                                                     g3c.ctypes.data_as(P64), n3c.ctypes.data_as(P64)) == 0
                assert open(pp, "rb").read() == g["params3_file"].tobytes()


@pytest.mark.parametrize("name,kw", [("correl_hc", dict(nhiggs=1)),
                                     ("correl_htc", dict(nhiggs=1, triplet=1))])
def test_correlator_fixture(name, kw):
    g = load(name)
    p = make_params(L=tuple(g["L"]), **kw, seed=31)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    assert o.lib.orc_longest_dir(o.h) == int(g["longest"][0])
    for d in range(3):
        a, b = o.correlators(d), g["correl_%d" % d]
        assert (np.abs(a - b) <= 1e-12 * (np.abs(b).max(0) + 1e-300)).all(), d


def test_blocked_correlator_fixture():
    g = load("blocked_tcb")
    p = make_params(L=tuple(g["L"]), nhiggs=0, triplet=1, seed=33)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    ours, ref = o.blocked_correlators(5), g["blocked"]
    assert ours.shape == ref.shape
    assert (np.abs(ours - ref) <= 1e-12 * (np.abs(ref).max(1, keepdims=True) + 1e-300)).all()


def test_zprofile_fixture():
    g = load("zprofile_hsuz")
    p = make_params(L=tuple(g["L"]), nhiggs=1, u1=1, singlet=1, seed=12)
    o = orc.Oracle(p)
    o.set_fields(fields(g, "in_"))
    prof = np.ascontiguousarray(o.zprofile())
    lib = abi.load()
    with tempfile.TemporaryDirectory() as d:
        lp, mp = os.path.join(d, "labels_z"), os.path.join(d, "measure_z")
        assert lib.su2hip_zprof_labels_write(lp.encode(), C.byref(p)) == 0
        assert open(lp, "rb").read() == g["labels_z"].tobytes()
        assert lib.su2hip_zprof_append(mp.encode(), C.byref(p), ptr(prof), len(prof)) == 0
        ours, ref = open(mp).read().split(), g["measure_z"].tobytes().decode().split()
    assert len(ours) == len(ref)
    for x, y in zip(ours, ref):
        assert x == y or abs(float(x) - float(y)) <= 1e-5 * abs(float(y))
