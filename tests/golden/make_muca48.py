"""Write tests/golden/muca48_reference.npz: the multicanonical SU(2)+adjoint
Higgs run at 48^3 (betaG 10, msq_triplet -0.12, b4 0.01, order parameter
Sigmasq, READONLY weight of 200 bins on [0,2] with seeded slopes in [-5,5],
one global check per parity half-sweep, 500 thermalisation + 2000 measured
sweeps) made by the reference's own binary, compiled from its sources
(oracle/_ref/su2_t, `make -C oracle ref`), in four independent chains (the
reference seeds its Mersenne twister from the clock).

The fixture holds the weight file, the config text and, per chain, the
`measure` columns 2 (order parameter) and 6 (Sigma^2) of every sweep.

    python tests/golden/make_muca48.py            # runs the four chains (~25 min)
    python tests/golden/make_muca48.py --collect DIR1 DIR2 ...   # from finished runs
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "su2_t")

CONFIG = """dim 3
L1 48
L2 48
L3 48
reset 1
iterations 2000
n_thermalize 500
interval 1
checkpoint 100000
run_checks 0
resultsfile measure
latticefile lattice
betasu2 10.0
sigma0 0.5
msq_triplet -0.12
b4 0.01
update_triplet 1
algorithm_su2link heatbath
algorithm_su2triplet overrelax
scalar_sweeps 4
update_links 1
update_doublet 1
multicanonical 1
muca_mode 0
weightfile weight
checks_per_sweep 1
bins 200
min 0.0
max 2.0
muca_delta 0.0
orderparam Sigmasq
measure_local 0
"""


def weight_text(seed=48, bins=200, lo=0.0, hi=2.0, slope=5.0):
    """Piecewise-linear weight in the reference's file format
    (src/multicanonical.c:82-111): header `bins delta last_max min max`, then
    `pos W` for the bins+1 bin edges, W_0 = 0, W_{i+1} = W_i + s_i * dbin."""
    rng = np.random.default_rng(seed)
    d = (hi - lo) / bins
    W = [0.0]
    for s in rng.uniform(-slope, slope, bins):
        W.append(W[-1] + s * d)
    lines = ["%d %.12f %d %f %f" % (bins, 0.0, 0, lo, hi)]
    lines += ["%f %f" % (lo + i * d, W[i]) for i in range(bins + 1)]
    return "\n".join(lines) + "\n"


def read_measure(path):
    a = np.loadtxt(path, ndmin=2)
    return a[:, 1], a[:, 5]


def run_chains(n):
    dirs = []
    procs = []
    for k in range(n):
        d = tempfile.mkdtemp(prefix="muca48_")
        with open(os.path.join(d, "config"), "w") as f:
            f.write(CONFIG)
        with open(os.path.join(d, "weight"), "w") as f:
            f.write(weight_text())
        procs.append(subprocess.Popen([EXE, "config"], cwd=d, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL))
        dirs.append(d)
        time.sleep(2)  # distinct clock seeds
    for p in procs:
        if p.wait() != 0:
            sys.exit("reference run failed")
    return dirs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=4)
    ap.add_argument("--collect", nargs="*", help="finished run directories")
    args = ap.parse_args()
    if args.collect:
        dirs, cleanup = args.collect, False
        for d in dirs:
            if open(os.path.join(d, "config")).read() != CONFIG or \
                    open(os.path.join(d, "weight")).read() != weight_text():
                sys.exit("%s: config or weight differs from this script's" % d)
    else:
        if not os.path.exists(EXE):
            sys.exit("reference build absent: make -C oracle ref")
        dirs, cleanup = run_chains(args.chains), True
    op, s2 = [], []
    for d in dirs:
        a, b = read_measure(os.path.join(d, "measure"))
        if len(a) != 2000:
            sys.exit("%s: %d measure lines" % (d, len(a)))
        op.append(a)
        s2.append(b)
    np.savez_compressed(os.path.join(HERE, "muca48_reference.npz"),
                        orderparam=np.array(op), sigma2=np.array(s2),
                        config=np.frombuffer(CONFIG.encode(), dtype=np.uint8),
                        weight=np.frombuffer(weight_text().encode(), dtype=np.uint8))
    if cleanup:
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)
    print("wrote", os.path.join(HERE, "muca48_reference.npz"))


if __name__ == "__main__":
    main()
