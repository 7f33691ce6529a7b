"""The 48^3 SU(2)+adjoint multicanonical run through the `su2hip` driver,
and its comparison with the reference's own chains recorded in
tests/golden/muca48_reference.npz (order parameter Sigmasq, READONLY weight,
one global check per parity half-sweep, 500 + 2000 sweeps)."""
import os
import subprocess
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
