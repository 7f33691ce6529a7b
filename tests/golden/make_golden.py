"""Write the golden fixtures of tests/golden/ from the reference compiled from
its own sources (oracle/_ref, built by `make -C oracle ref`).

Each fixture is data only: an input configuration (a Philox hot start of a
4x4x6 lattice) and what the reference's own functions compute from it:
per-link staples and local actions, the measure sums and the measure line,
the fields and counters after update_lattice sweeps with the Philox stream,
the gradient-flow force and an Euler step, multicanonical sweeps, the
bytes of a saved lattice file, and the correlators of the CORRELATORS build.

    python tests/golden/make_golden.py [fixture names, default all]
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import orc  # noqa: E402
from su2higgs_amd import abi  # noqa: E402
from su2higgs_amd.abi import Muca, make_params, ptr  # noqa: E402

MODELS = {
    "h": dict(nhiggs=1),
    "hu": dict(nhiggs=1, u1=1),
    "t": dict(nhiggs=0, triplet=1),
    "ht": dict(nhiggs=1, triplet=1),
    "hsu": dict(nhiggs=1, u1=1, singlet=1),
    "d": dict(nhiggs=2),
    "du": dict(nhiggs=2, u1=1),
}
L = (4, 4, 6)


def fields_of(d, prefix):
    return {k: d[prefix + k] for k in orc.FIELDS if prefix + k in d}


def model_fixture(tag):
    p = make_params(L=L, **MODELS[tag], seed=2024)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p)
    r.set_fields(f)
    V = o.V
    out = {"L": np.array(L)}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    out["link_staple"] = np.array([[r.vec("ref_link_staple", 4, i, d) for d in range(3)] for i in range(V)])
    out["localact_su2link"] = np.array([[r.lib.ref_localact_su2link(i, d) for d in range(3)] for i in range(V)])
    out["action_local"] = np.array([r.lib.ref_action_local(i) for i in range(V)])
    if p.nhiggs:
        out["doublet_staple"] = np.array([r.vec("ref_doublet_staple", 4, i) for i in range(V)])
    if p.nhiggs == 2:
        out["doublet2_staple"] = np.array([r.vec("ref_doublet_staple_h", 4, i, 1) for i in range(V)])
        out["localact_doublet2"] = np.array([r.lib.ref_localact_doublet_h(i, 1) for i in range(V)])
    out["measure_sums"] = r.measure()
    out["measure_text"] = np.frombuffer(r.measure_text().encode(), dtype=np.uint8)
    # two compound sweeps of the reference's update_lattice, Philox seed 99
    r.rng_set(99, 0)
    r.update(2)
    g = r.get_fields()
    for k in orc.FIELDS:
        out["sweep2_" + k] = g[k]
    c = r.counters()
    out["sweep2_counters"] = np.array([c[k] for k in sorted(c)], dtype=np.int64)
    out["counter_names"] = np.array(sorted(c))
    out["sweep2_pass"] = np.array([r.lib.ref_rng_pass()], dtype=np.uint64)
    # the lattice file of the swept fields
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "lattice")
        r.lib.ref_save_lattice(path.encode(), 3)
        out["lattice_bytes"] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    return out


def flow_fixture():
    p = make_params(L=L, nhiggs=1, seed=2025)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p, gradflow=True)
    r.set_fields(f)
    out = {"L": np.array(L)}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    F = np.zeros((o.V, 3, 4))
    r.lib.ref_flow_force(ptr(F), None)
    out["force"] = F
    r.lib.ref_flow_euler_step(0.02)
    out["euler_su2link"] = r.get_fields()["su2link"]
    r.set_fields(f)
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            r.lib.ref_grad_flow(0.25, 0.05, 2, 1)
            out["measure_flow"] = np.frombuffer(open("measure_flow", "rb").read(), dtype=np.uint8)
        finally:
            os.chdir(cwd)
    return out


def triplet_flow_fixture():
    p = make_params(L=L, nhiggs=0, triplet=1, seed=2026)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p, gradflow=True)
    r.set_fields(f)
    out = {"L": np.array(L)}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    F = np.zeros((o.V, 3, 4))
    T = np.zeros((o.V, 3))
    r.lib.ref_flow_force(ptr(F), ptr(T))
    out["force"] = F
    out["tforce"] = T
    return out


def write_weight(path, bins, lo, hi, seed=3, slope=5.0):
    rng = np.random.default_rng(seed)
    dbin = (hi - lo) / bins
    W = [0.0]
    for s in rng.uniform(-slope, slope, bins):
        W.append(W[-1] + s * dbin)
    with open(path, "w") as f:
        f.write("%d %.12f %d %f %f\n" % (bins, 0.3, 0, lo, hi))
        for i in range(bins + 1):
            f.write("%f %f\n" % (lo + i * dbin, W[i]))


def write_params(path, bins, lo, hi, seed=4):
    """A weight_params file (src/multicanonical.c:99-109 format) with
    seeded gsum/nsum."""
    rng = np.random.default_rng(seed)
    dbin = (hi - lo) / bins
    with open(path, "w") as f:
        f.write("%d \n" % bins)
        for i in range(1, bins + 1):
            f.write("%f %d %d\n" % (lo + (i - 1) * dbin, rng.integers(5, 60), rng.integers(1, 40)))


def muca_fixture(tag, op, mode, cps, slope, interval=None, params=False, sweeps=3, rng=None):
    p = make_params(L=L, **MODELS[tag], seed=2027)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p)
    r.set_fields(f)
    lo, hi = (-1.0, 1.0) if op == abi.PHI2MINUSSIGMA2 else (0.0, 2.0)
    if rng:
        lo, hi = rng
    out = {"L": np.array(L), "nsweeps": np.array([sweeps]), "range": np.array([lo, hi])}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            write_weight("weight", 40, lo, hi, slope=slope)
            out["weight_file"] = np.frombuffer(open("weight", "rb").read(), dtype=np.uint8)
            if params:
                write_params("weight_params", 40, lo, hi)
                out["params_file"] = np.frombuffer(open("weight_params", "rb").read(), dtype=np.uint8)
            m = Muca()
            m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, op, mode, cps
            m.bins, m.min, m.max, m.delta = 40, lo, hi, 0.3
            r.lib.ref_muca_load(b"weight", C.byref(m))
            if interval:
                r.lib.ref_muca_set_interval(interval)
            n = 42
            g0, n0 = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
            r.lib.ref_muca_sums(g0.ctypes.data_as(C.POINTER(C.c_int64)), n0.ctypes.data_as(C.POINTER(C.c_int64)))
            out["gsum0"], out["nsum0"] = g0, n0
            mr = Muca()
            pos, W, hits = np.zeros(42), np.zeros(42), np.zeros(42, dtype=np.int32)
            r.lib.ref_muca_get(C.byref(mr), ptr(pos), ptr(W), hits.ctypes.data_as(C.POINTER(C.c_int32)))
            out["pos"], out["W0"] = pos.copy(), W.copy()
            out["muca0"] = np.frombuffer(bytes(mr), dtype=np.uint8)
            r.rng_set(99, 0)
            r.update(sweeps)
            r.lib.ref_muca_get(C.byref(mr), ptr(pos), ptr(W), hits.ctypes.data_as(C.POINTER(C.c_int32)))
            g3, n3 = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
            r.lib.ref_muca_sums(g3.ctypes.data_as(C.POINTER(C.c_int64)), n3.ctypes.data_as(C.POINTER(C.c_int64)))
            out["gsum3"], out["nsum3"] = g3, n3
            # the files the reference's save_weight wrote after its last update
            out["weight3_file"] = np.frombuffer(open("weight", "rb").read(), dtype=np.uint8)
            if mode == abi.SLOW:
                out["params3_file"] = np.frombuffer(open("weight_params", "rb").read(), dtype=np.uint8)
        finally:
            os.chdir(cwd)
    g = r.get_fields()
    for k in orc.FIELDS:
        out["sweep3_" + k] = g[k]
    c = r.counters()
    out["sweep3_counters"] = np.array([c[k] for k in sorted(c)], dtype=np.int64)
    out["counter_names"] = np.array(sorted(c))
    out["W3"], out["hits3"] = W, hits
    out["muca3"] = np.frombuffer(bytes(mr), dtype=np.uint8)
    return out


def correl_fixture(tag, kw):
    """Correlators of the reference's CORRELATORS build along every
    direction of a hot start (src/correlation.c:339-391)."""
    p = make_params(L=L, **kw, seed=31)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p, correlators=True)
    r.set_fields(f)
    out = {"L": np.array(L), "longest": np.array([r.lib.ref_longest_dir()])}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    for d in range(3):
        out["correl_%d" % d] = r.correlators(d)
    return out


def blocked_fixture():
    """Correlators on the blocked lattices of the BLOCKING build."""
    Lb = (8, 8, 12)
    p = make_params(L=Lb, nhiggs=0, triplet=1, seed=33)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p, blocking=True)
    r.set_fields(f)
    levels = r.lib.ref_blocked_setup(5)
    vals = np.zeros((levels, Lb[r.lib.ref_longest_dir()] // 2, 10))
    r.lib.ref_blocked_values(ptr(vals))
    out = {"L": np.array(Lb), "blocked": vals}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    return out


def zprofile_fixture():
    """measure_along_z text and labels_z of the MEASURE_Z build on a hot start."""
    Lz = (4, 8, 6)
    p = make_params(L=Lz, nhiggs=1, u1=1, singlet=1, seed=12)
    o = orc.Oracle(p)
    o.hot_start()
    f = o.get_fields()
    r = orc.Ref(p, measure_z=True)
    r.set_fields(f)
    out = {"L": np.array(Lz)}
    for k in orc.FIELDS:
        out["in_" + k] = f[k]
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            r.lib.ref_print_z_labels()
            r.lib.ref_measure_along_z(1)
            out["labels_z"] = np.frombuffer(open("labels_z", "rb").read(), dtype=np.uint8)
            out["measure_z"] = np.frombuffer(open("measure_z", "rb").read(), dtype=np.uint8)
        finally:
            os.chdir(cwd)
    return out


CORREL_CASES = {"correl_hc": ("hc", dict(nhiggs=1)),
                "correl_htc": ("htc", dict(nhiggs=1, triplet=1))}


MUCA_CASES = {
    "muca_h_phisq_ro": ("h", abi.PHISQ, abi.READONLY, 1, 5.0),
    "muca_h_phisq_fast3": ("h", abi.PHISQ, abi.FAST, 3, 400.0),
    "muca_t_sigmasq_ro": ("t", abi.SIGMASQ, abi.READONLY, 1, 400.0),
    "muca_ht_phi2minus_fast2": ("ht", abi.PHI2MINUSSIGMA2, abi.FAST, 2, 5.0),
    # weight recursions that run within the 3 sweeps (update interval set
    # after load_weight); SLOW from a fresh and from a loaded weight_params
    "muca_h_phisq_fast_i7": ("h", abi.PHISQ, abi.FAST, 1, 5.0, 7),
    "muca_h_phisq_slow_i30": ("h", abi.PHISQ, abi.SLOW, 2, 5.0, 30, False, 6, (1.0, 2.0)),
    "muca_t_sigmasq_slow_params": ("t", abi.SIGMASQ, abi.SLOW, 2, 5.0, 30, True, 6, (0.5, 1.5)),
    # two doublets, the second one's phi^2 as order parameter (2 sweeps:
    # the two-doublet sweep amplifies rounding)
    "muca_d_phi2sq_fast2": ("d", abi.PHI2SQ, abi.FAST, 2, 5.0, None, False, 2),
    "muca_du_phi2sq_ro": ("du", abi.PHI2SQ, abi.READONLY, 3, 400.0, None, False, 2),
}


def main():
    only = sys.argv[1:]
    for tag in MODELS:
        if not orc.ref_available(tag):
            sys.exit("reference build absent: make -C oracle ref")
        if not only or "model_" + tag in only:
            np.savez_compressed(os.path.join(HERE, "model_%s.npz" % tag), **model_fixture(tag))
    if not only or "flow_h" in only:
        np.savez_compressed(os.path.join(HERE, "flow_h.npz"), **flow_fixture())
    if not only or "flow_t" in only:
        np.savez_compressed(os.path.join(HERE, "flow_t.npz"), **triplet_flow_fixture())
    for name, args in MUCA_CASES.items():
        if not only or name in only:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **muca_fixture(*args))
    for name, (tag, kw) in CORREL_CASES.items():
        if not only or name in only:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **correl_fixture(tag, kw))
    if not only or "blocked_tcb" in only:
        np.savez_compressed(os.path.join(HERE, "blocked_tcb.npz"), **blocked_fixture())
    if not only or "zprofile_hsuz" in only:
        np.savez_compressed(os.path.join(HERE, "zprofile_hsuz.npz"), **zprofile_fixture())
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
