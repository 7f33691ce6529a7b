"""Write the FP64 reference values tests/fp64_reference_48.json, read by
tests/test_gpu_fp32.py and bench.py (needs the GPU, about 7 minutes):

    python tests/fp64_reference_48.py

512 independent FP64 chains (Philox keys 64001 + 17 k) of the 48^3
SU(2)xU(1)+Higgs theory at the single-lattice workload's couplings, the
lean sweep mix of the test, cold start, 4000 sweeps discarded, then 12000
sweeps measured every 10: the protocol of the test's FP32 chains, so that
the two precisions are compared at the same Monte Carlo time from the
same start.  Prints each observable's mean over chains of
the chain's time average and its error from the chain-to-chain spread
(the stored file keeps 9 and 2 significant digits).
The FP64 path is the one pinned site by site to the oracle."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from test_gpu_fp32 import chain_mean, chain_series  # noqa: E402
from su2higgs_amd.abi import FP64  # noqa: E402

if __name__ == "__main__":
    # four batches of 128 chains (chain k of batch j is chain 128 j + k: a
    # chain's results do not depend on the batch it runs in), a line each
    parts = []
    for j in range(4):
        parts.append(chain_series(FP64, 128, 4000, 12000, 10, seed0=64001 + 17 * 128 * j))
        print("batch %d of 4 done" % (j + 1), flush=True)
    x = {name: np.concatenate([p[name] for p in parts], axis=1) for name in parts[0]}
    out = {name: [float(v) for v in chain_mean(series)] for name, series in x.items()}
    print(json.dumps(out))
    with open(os.path.join(HERE, "fp64_reference_48.json"), "w") as f:
        json.dump({k: [float("%.9f" % m), float("%.2g" % e)] for k, (m, e) in out.items()}, f)
        f.write("\n")
