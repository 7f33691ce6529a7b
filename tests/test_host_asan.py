"""The host readers of untrusted files under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

tests/asan_io_harness.c is built with gcc -fsanitize=address,undefined from
su2higgs_amd/host/su2hip_io.c and su2hip_traj.c, the code that parses what
the reference parses at src/parameters.c:30-92 (config), src/checkpoint.c:
133-214 (lattice), src/multicanonical.c:118-333 (weight, weight_params) and
src/hb_trajectory.c:99-163 (realtime_config).  It is fed valid files and
seeded mutations of them: every truncation length of a checkpoint header and
many of its body, corrupted headers (huge, zero and negative sizes and
extents), byte flips, overlong lines and tokens, missing values, binary
garbage and empty files.  Every call must return (any status); a sanitizer
report fails the test."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "su2higgs_amd", "host")
HARNESS = os.path.join(ROOT, "tests", "asan_io_harness.c")

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]

CONFIG = """# sample
dim 3
L1 4
L2 6
L3 8
iterations 10
interval 1
checkpoint 5
n_thermalize 0
reset 1
run_checks 0
resultsfile measure
latticefile lattice
betasu2 8.0
betau1 9.0
r_u1 1
msq -0.25
lambda 0.02
msq_triplet -0.3
a2 0.1
b4 0.1
phi0 1.0
sigma0 0.2
algorithm_su2link heatbath
algorithm_u1link metropolis
algorithm_su2doublet overrelax
algorithm_su2triplet overrelax
scalar_sweeps 3
update_doublet 1
update_triplet 1
update_links 1
multicanonical 1
weightfile weight
bins 10
min 0.0
max 2.0
muca_mode 1
orderparam phisq
muca_delta 0.3
checks_per_sweep 1
measure_local 0
do_flow 1
flow_interval 5
flow_meas_interval 3
flow_dt 0.05
flow_t_max 0.5
"""

REALTIME = "n_traj 4\ninterval 2\nmode_interval 10\nmin 0.1\nmax 0.9\n"
MODELS = ["higgs+u1", "2hdm+u1", "triplet", "higgs+triplet+singlet"]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc absent")
    d = tmp_path_factory.mktemp("asan")
    exe = str(d / "harness")
    cmd = [gcc, "-std=gnu11", "-g", "-O1", "-w"] + SAN + [
        "-I" + os.path.join(ROOT, "include"), HARNESS, os.path.join(HOST, "su2hip_io.c"),
        os.path.join(HOST, "su2hip_traj.c"), "-o", exe, "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0 and "asan" in r.stdout.lower():
        pytest.skip("no sanitizer runtime: " + r.stdout[-300:])
    assert r.returncode == 0, r.stdout
    return exe, d


def run(exe, lines):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    r = subprocess.run([exe], input="".join(lines), stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    out = [l.split() for l in r.stdout.splitlines()]
    assert len(out) == len(lines), r.stdout[-2000:]
    return [int(o[-1]) for o in out]


def mutations(data, rng, n):
    """Seeded corruptions of a byte string."""
    out = []
    for k in range(n):
        b = bytearray(data)
        kind = k % 6
        if kind == 0 and len(b):                      # truncation
            b = b[: int(rng.integers(0, len(b)))]
        elif kind == 1 and len(b):                    # byte flips
            for _ in range(int(rng.integers(1, 8))):
                i = int(rng.integers(0, len(b)))
                b[i] = int(rng.integers(0, 256))
        elif kind == 2:                               # a very long line / token
            i = int(rng.integers(0, len(b) + 1))
            b[i:i] = b"x" * int(rng.integers(100, 20000))
        elif kind == 3 and len(b):                    # a duplicated chunk
            i, j = sorted(int(x) for x in rng.integers(0, len(b), 2))
            b[j:j] = b[i:j]
        elif kind == 4:                               # random garbage
            b = bytearray(rng.integers(0, 256, int(rng.integers(0, 2000)), dtype=np.uint8))
        else:                                         # lines dropped
            ls = bytes(b).split(b"\n")
            keep = rng.random(len(ls)) > 0.3
            b = bytearray(b"\n".join(l for l, k_ in zip(ls, keep) if k_))
        out.append(bytes(b))
    return out


def test_config_reader_survives_malformed_files(harness):
    exe, d = harness
    rng = np.random.default_rng(1)
    data = CONFIG.encode()
    cases = [data, b"", b"\n" * 1000, b"L1\nL2\nL3\n", b"dim 3\nL1 999999999999999999999\n",
             b"dim 3\nL1 -4\nL2 -6\nL3 0\n", data.replace(b"bins 10", b"bins -5"),
             data.replace(b"bins 10", b"bins 2147483647"), data.replace(b"0.3", b"nan"),
             data.replace(b"resultsfile measure", b"resultsfile " + b"m" * 5000)]
    cases += mutations(data, rng, 240)
    lines = []
    for i, c in enumerate(cases):
        p = d / ("config%d" % i)
        p.write_bytes(c)
        lines.append("config %s %s\n" % (p, MODELS[i % len(MODELS)]))
    rc = run(exe, lines)
    assert rc[0] == 0  # the valid file parses


@pytest.mark.parametrize("model", MODELS)
def test_lattice_reader_survives_truncation_and_corruption(harness, model):
    exe, d = harness
    good = d / ("lattice_%s" % model.replace("+", "_"))
    assert run(exe, ["mklattice %s %s\n" % (good, model)]) == [0]
    data = good.read_bytes()
    rng = np.random.default_rng(2)
    hdr = 4 * 5 + 8 * 3  # size, dim, L[3], iter, total, comms
    cases = [data]
    cases += [data[:n] for n in range(0, hdr + 1)]                # every header truncation
    cases += [data[: int(n)] for n in rng.integers(hdr, len(data), 40)]
    cases.append(data + b"\0" * 64)                             # trailing bytes
    for size in (0, -1, 2, 3, 4, 7, 1 << 30, -(1 << 31)):
        cases.append(struct.pack("<i", size) + data[4:])
    for dim in (0, 2, 4, -3, 1 << 20):
        cases.append(data[:4] + struct.pack("<i", dim) + data[8:])
    for L in ((0, 6, 8), (-4, 6, 8), (4, 6, 1 << 30), (1 << 20, 1 << 20, 1 << 20), (8, 6, 4)):
        cases.append(data[:8] + struct.pack("<3i", *L) + data[20:])
    cases += mutations(data, rng, 60)
    lines = []
    for i, c in enumerate(cases):
        p = d / ("lat_%s_%d" % (model.replace("+", "_"), i))
        p.write_bytes(c)
        lines.append("lattice %s %s\n" % (p, model))
    rc = run(exe, lines)
    assert rc[0] == 0
    assert all(r != 0 for r in rc[1: hdr + 2])  # a truncated header is an error, not a read


def test_weight_readers_survive_malformed_files(harness):
    exe, d = harness
    good = d / "weight_good"
    assert run(exe, ["mkweight %s 10\n" % good]) == [0]
    data = good.read_bytes()
    rng = np.random.default_rng(3)
    cases = [data, b"", b"10 0.0 2.0\n", b"-5 0 2\n", b"2147483647 0 2\n" + b"0 0\n" * 5,
             b"1000000 0 2\n", data.replace(b"\n", b" ", 3)]
    cases += mutations(data, rng, 120)
    wp = b"".join(b"%d %d %d\n" % (i, 5 + i, 1 + i) for i in range(1, 11))
    wcases = [wp, b"", b"3 1 1\n", b"-1 5 1\n" * 20, b"999999 5 1\n"] + mutations(wp, rng, 60)
    lines = []
    for i, c in enumerate(cases):
        p = d / ("weight%d" % i)
        p.write_bytes(c)
        for bins in (10, 12, 3):
            lines.append("weight %s %d\n" % (p, bins))
    for i, c in enumerate(wcases):
        p = d / ("wparams%d" % i)
        p.write_bytes(c)
        lines.append("wparams %s 10\n" % p)
    rc = run(exe, lines)
    assert rc[0] == 0


def test_realtime_reader_and_mpi_slicing_survive_bad_input(harness):
    exe, d = harness
    rng = np.random.default_rng(4)
    data = REALTIME.encode()
    cases = [data, b"", b"n_traj\n", b"#n_traj 3\n" * 300, b"min " + b"9" * 400 + b"\n"]
    cases += mutations(data, rng, 80)
    lines = []
    for i, c in enumerate(cases):
        p = d / ("realtime%d" % i)
        p.write_bytes(c)
        lines.append("realtime %s\n" % p)
    lines += ["slices %d\n" % s for s in (1, 2, 3, 4, 6, 8, 12, 16, 24, 48, 96, 0, -1, 5, 7, 97,
                                          1 << 20, 2147483647)]
    rc = run(exe, lines)
    assert rc[0] == 0
