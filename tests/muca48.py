"""The 48^3 SU(2)+adjoint multicanonical run through the `su2hip` driver,
and its comparison with the reference's own chains recorded in
tests/golden/muca48_reference.npz (order parameter Sigmasq, READONLY weight,
one global check per parity half-sweep, 500 + 2000 sweeps)."""
import ctypes as C
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "su2higgs_amd", "bin", "su2hip")
FIXTURE = os.path.join(HERE, "golden", "muca48_reference.npz")


def blocked(x, nblocks=20):
    """Mean and its error from `nblocks` contiguous blocks of a chain."""
    x = np.asarray(x)
    n = len(x) // nblocks * nblocks
    b = x[:n].reshape(nblocks, -1).mean(1)
    return float(b.mean()), float(b.std(ddof=1) / np.sqrt(nblocks))


def reference():
    f = np.load(FIXTURE, allow_pickle=False)
    return {k: f[k] for k in f.files}


def run(fp32=True, seed=48, timeout=600):
    """Run the driver on the fixture's config and weight with a fixed Philox
    seed (the reference seeds from the clock; a `seed` line is a key it does
    not read); returns the measure columns 2 (order parameter) and 6
    (Sigma^2), the wall time and the acceptance text it printed."""
    ref = reference()
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "config"), "wb") as f:
            f.write(ref["config"].tobytes())
            f.write(b"seed %d\n" % seed)
        with open(os.path.join(d, "weight"), "wb") as f:
            f.write(ref["weight"].tobytes())
        cmd = [EXE, "--model", "triplet"] + (["--fp32"] if fp32 else []) + ["config"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=timeout)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("su2hip failed:\n" + r.stdout[-2000:])
        a = np.loadtxt(os.path.join(d, "measure"), ndmin=2)
        w = open(os.path.join(d, "weight"), "rb").read()
    return {"orderparam": a[:, 1], "sigma2": a[:, 5], "wall_s": wall, "log": r.stdout,
            "weight_unchanged": w == ref["weight"].tobytes()}


def compare(ours, ref):
    """Blocked means of the order parameter (ours: one chain; reference: the
    chains pooled), the z-score of their difference, the ratio of the
    per-sweep spreads, and both histograms on common bins."""
    ro = ref["orderparam"]
    m_ref_chains = [blocked(c) for c in ro]
    means = np.array([m for m, _ in m_ref_chains])
    errs = np.array([e for _, e in m_ref_chains])
    m_ref = float(means.mean())
    e_ref = float(np.sqrt((errs ** 2).sum()) / len(errs))
    m_us, e_us = blocked(ours["orderparam"])
    z = (m_us - m_ref) / np.hypot(e_us, e_ref)
    lo = min(ro.min(), ours["orderparam"].min())
    hi = max(ro.max(), ours["orderparam"].max())
    edges = np.linspace(lo, hi, 41)
    h_ref = np.histogram(ro.ravel(), edges)[0] / ro.size
    h_us = np.histogram(ours["orderparam"], edges)[0] / len(ours["orderparam"])
    return {"mean": m_us, "err": e_us, "ref_mean": m_ref, "ref_err": e_ref, "z": float(z),
            "spread_ratio": float(ours["orderparam"].std() / ro.std(axis=1).mean()),
            "hist_tv_distance": float(0.5 * np.abs(h_ref - h_us).sum()),
            "hist_edges": [float(edges[0]), float(edges[-1]), len(edges) - 1],
            "ref_chains": int(ro.shape[0]), "sweeps": int(len(ours["orderparam"]))}


def run_batch(fp32=True, chains=16, seed0=480000):
    """The fixture's run as `chains` independent chains (Philox keys seed0 +
    k) in one handle, every kernel launch covering all of them: the config
    parsed by su2hip_config_read, the weight by su2hip_weight_load
    (READONLY), cold start, n_thermalize sweeps, then `iterations` sweeps
    with the order parameter measured before each, in the driver's
    measure-then-update order (src/main.c:302-353).  Returns the order
    parameter as [chains, iterations] and the wall time."""
    sys.path.insert(0, ROOT)
    from su2higgs_amd import abi
    from su2higgs_amd.lattice import Batch
    ref = reference()
    lib = abi.load()
    with tempfile.TemporaryDirectory() as d:
        cpath, wpath = os.path.join(d, "config"), os.path.join(d, "weight")
        with open(cpath, "wb") as f:
            f.write(ref["config"].tobytes())
        with open(wpath, "wb") as f:
            f.write(ref["weight"].tobytes())
        cfg = abi.RunConfig()
        abi.check(lib.su2hip_config_read(cpath.encode(), b"triplet", C.byref(cfg)))
        m = abi.Muca()
        C.memmove(C.byref(m), C.byref(cfg.muca), C.sizeof(abi.Muca))
        n = m.bins + 2
        pos, W = np.zeros(n), np.zeros(n)
        abi.check(lib.su2hip_weight_load(wpath.encode(), C.byref(m), abi.ptr(pos), abi.ptr(W),
                                         m.bins))
    p = cfg.p
    p.precision = abi.FP32 if fp32 else abi.FP64
    p.reunitarize = 1 if fp32 else 0
    ps = []
    for k in range(chains):
        q = abi.Params()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(abi.Params))
        q.seed = seed0 + k
        ps.append(q)
    t0 = time.perf_counter()
    b = Batch(ps)
    for k in range(chains):
        b.cold_start(k, phi0=cfg.phi0, sigma0=cfg.sigma0, singlet0=cfg.singlet0)
        b.muca_set(k, m, pos, W)
    b.use_graph(True)
    b.update(int(cfg.n_thermalize))
    V = float(abi.volume(p))
    op = np.zeros((chains, int(cfg.iterations)))
    for it in range(int(cfg.iterations)):
        op[:, it] = b.measure_all()[:chains, abi.SUM["sigma2"]] / V
        b.update(1)
    b.sync()
    wall = time.perf_counter() - t0
    b.close()
    return op, wall


def tv(a, b, edges):
    """Total-variation distance of the histograms of two samples."""
    ha = np.histogram(np.ravel(a), edges)[0] / np.size(a)
    hb = np.histogram(np.ravel(b), edges)[0] / np.size(b)
    return float(0.5 * np.abs(ha - hb).sum())


def compare_chains(ours, ref):
    """Our chains (pooled) against the reference's: the order parameter's
    mean with its error from our chain-to-chain spread, and the reference's
    from its blocked chains; z of the difference; the histogram distance
    between the two pools and, as its noise scale, the distance between each
    reference chain and the pool of the other reference chains."""
    ro = ref["orderparam"]
    means = np.array([blocked(c)[0] for c in ro])
    errs = np.array([blocked(c)[1] for c in ro])
    m_ref = float(means.mean())
    e_ref = float(np.sqrt((errs ** 2).sum()) / len(errs))
    cm = ours.mean(1)
    m_us, e_us = float(cm.mean()), float(cm.std(ddof=1) / np.sqrt(len(cm)))
    lo, hi = min(ro.min(), ours.min()), max(ro.max(), ours.max())
    edges = np.linspace(lo, hi, 41)
    loo = [tv(ro[k], np.delete(ro, k, axis=0), edges) for k in range(len(ro))]
    return {"chains": int(ours.shape[0]), "sweeps": int(ours.shape[1]), "mean": m_us, "err": e_us,
            "ref_mean": m_ref, "ref_err": e_ref, "z": (m_us - m_ref) / float(np.hypot(e_us, e_ref)),
            "hist_tv_distance": tv(ours, ro, edges), "ref_leave_one_out_tv": loo,
            "spread_ratio": float(ours.std(axis=1).mean() / ro.std(axis=1).mean()),
            "ref_chains": int(ro.shape[0])}
