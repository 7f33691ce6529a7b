"""The reference's own driver with its sweep on the GPU.

oracle/_ref/su2_<tag>_gpu is the reference's unmodified main loop
(src/main.c) linked with the reference-side binding
su2higgs_amd/bind/su2hip_bind.c: update_lattice is routed to libsu2hip
(oracle/bind_shim.c).  On the same config it must write the `measure` file
the su2hip driver writes: both run the same GPU sweeps (the Philox key the
binding drew from the reference's generator and printed, pass counter from
0); the reference measures on the CPU with its own measure(), su2hip on the
GPU.  The final `lattice` files hold bitwise the same fields."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import make_params, ptr
from test_gpu_cli import CONFIG, TWO_HIGGS, run_cli, same_lines

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_ref_gpu(tag, d, cfg, mt_seed=None):
    """Run the reference's main loop with the GPU sweep in directory d;
    mt_seed seeds the reference's generator (through the test shim) before
    the binding draws its Philox key from it."""
    exe = os.path.join(ROOT, "oracle", "_ref", "su2_%s_gpu" % tag)
    if not os.path.exists(exe):
        pytest.skip("reference build with the binding absent")
    with open(os.path.join(d, "config"), "w") as f:
        f.write(cfg)
    env = dict(os.environ)
    if mt_seed is not None:
        env["SU2HIP_SHIM_MT_SEED"] = str(mt_seed)
    r = subprocess.run([exe, "config"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def philox_key(stdout):
    """The key (and starting pass) su2hip_bind_init printed."""
    m = re.search(r"su2hip: Philox key (\d+), pass (\d+)", stdout)
    assert m, stdout[-2000:]
    return int(m.group(1)), int(m.group(2))


def with_key(cfg, key):
    return cfg.replace("seed 4242", "seed %d" % key)


def acceptance_lines(stdout):
    """The rate line after each "-- Acceptance rates --" (print_acceptance,
    src/checkpoint.c:17-78, printed at every checkpoint)."""
    lines = stdout.splitlines()
    return [lines[i + 1].strip() for i, l in enumerate(lines) if l.strip() == "-- Acceptance rates --"]


def lattice_fields(path, p):
    lib = abi.load()
    g = abi.field_arrays(p)
    it = C.c_int64()
    assert lib.su2hip_lattice_read(path.encode(), C.byref(p), C.byref(it), None, None,
                                   *[ptr(g[k]) for k in orc.FIELDS]) == 0
    return it.value, g


TRIPLET_KEYS = ("msq_triplet -0.3\nb4 0.1\nsigma0 0.4\nalgorithm_su2triplet overrelax\n"
                "update_triplet 1\n")
SINGLET_KEYS = ("update_singlet 1\nsinglet0 0.3\nalgorithm_singlet overrelax\nmsq_s -0.1\n"
                "b1_s 0.2\nb3_s 0.1\nb4_s 0.3\na1_s 0.3\na2_s 0.25\n")


@pytest.mark.parametrize("tag,model,extra,p_kw", [
    ("hu", "higgs+u1", "", dict(nhiggs=1, u1=1)),
    ("d", "2hdm", "".join("%s %r\n" % kv for kv in TWO_HIGGS.items()), dict(nhiggs=2)),
    ("t", "triplet", TRIPLET_KEYS, dict(nhiggs=0, triplet=1)),
    ("hsu", "higgs+u1+singlet", SINGLET_KEYS, dict(nhiggs=1, u1=1, singlet=1)),
], ids=["higgs_u1", "two_doublets", "adjoint", "higgs_u1_singlet"])
def test_reference_main_with_gpu_sweep_matches_su2hip(tag, model, extra, p_kw, tmp_path):
    """One doublet with U(1), two doublets, the adjoint alone, and a doublet
    with U(1) and a singlet: every field path of the binding."""
    cfg = CONFIG.format(model=model, it=8, do_flow=0).replace("checkpoint 1000", "checkpoint 4") + extra
    dr, dc = tmp_path / "ref", tmp_path / "su2hip"
    dr.mkdir()
    dc.mkdir()
    out_ref = run_ref_gpu(tag, str(dr), cfg)
    key, pass0 = philox_key(out_ref)
    assert pass0 == 0
    out_cli = run_cli(str(dc), with_key(cfg, key))
    ref = open(dr / "measure").read().splitlines(keepends=True)
    ours = open(dc / "measure").read().splitlines(keepends=True)
    assert len(ref) == 8
    same_lines(ours, ref)
    # acceptance rates at the checkpoints of iterations 4 and 8
    acc = acceptance_lines(out_ref)
    assert len(acc) == 2 and acc == acceptance_lines(out_cli), (acc, acceptance_lines(out_cli))
    p = make_params(L=(4, 6, 8), **p_kw)  # the field content is what the reader needs
    itr, fr = lattice_fields(str(dr / "lattice"), p)
    itc, fc = lattice_fields(str(dc / "lattice"), p)
    assert itr == itc
    for k in orc.FIELDS:
        if fr.get(k) is not None:
            assert np.array_equal(fr[k], fc[k]), k


@pytest.mark.parametrize("mode", [0, 1], ids=["readonly", "fast"])
def test_reference_main_multicanonical_with_gpu_sweep_matches_su2hip(mode, tmp_path):
    """A multicanonical run (order parameter phi^2, 40 bins on [0, 2], two
    global checks per parity pass): the weight moves to the device through
    su2hip_bind_muca and its state comes back after every sweep, so the
    reference's measure() prints the device's order parameter and weight.
    In FAST mode the weight is updated on the device as it goes."""
    from test_gpu_traj import write_weight
    cfg = CONFIG.format(model="higgs+u1", it=8, do_flow=0).replace("multicanonical 0", "multicanonical 1")
    cfg = cfg.replace("checkpoint 1000", "checkpoint 4")
    cfg += ("weightfile weight\nbins 40\nmin 0.0\nmax 2.0\nchecks_per_sweep 2\n"
            "orderparam phisq\nmuca_mode %d\nmuca_delta 0.3\n" % mode)
    dr, dc = tmp_path / "ref", tmp_path / "su2hip"
    for d in (dr, dc):
        d.mkdir()
        write_weight(str(d / "weight"), 40, 0.0, 2.0)
    out_ref = run_ref_gpu("hu", str(dr), cfg)
    out_cli = run_cli(str(dc), with_key(cfg, philox_key(out_ref)[0]))
    ref = open(dr / "measure").read().splitlines(keepends=True)
    ours = open(dc / "measure").read().splitlines(keepends=True)
    assert len(ref) == 8
    same_lines(ours, ref)
    acc = acceptance_lines(out_ref)
    assert len(acc) == 2 and acc == acceptance_lines(out_cli), (acc, acceptance_lines(out_cli))


def test_binding_keys_differ_between_runs(tmp_path):
    """Two runs whose reference generators are seeded differently (as two
    runs started at different times are) draw different Philox keys and so
    different chains."""
    cfg = CONFIG.format(model="higgs+u1", it=4, do_flow=0)
    outs = []
    for k, mt in enumerate((11, 12)):
        d = tmp_path / ("run%d" % k)
        d.mkdir()
        outs.append((philox_key(run_ref_gpu("hu", str(d), cfg, mt_seed=mt))[0],
                     open(d / "measure").read().splitlines()))
    (k0, m0), (k1, m1) = outs
    assert k0 != k1
    assert m0[0] == m1[0]  # the cold start, measured before the first sweep
    assert m0[1:] != m1[1:]


def test_binding_resume_continues_the_stream(tmp_path):
    """A run of 4 iterations resumed to iteration 8 (reset 0: the reference
    continues at the stored iteration + 1, src/main.c:280-282, so 7 sweeps
    in all) ends on the same fields, bit for bit, and writes the same
    measure lines as one run of 7 iterations: the binding saved the Philox
    position with the checkpoint and took it back.  Without the .rng file
    the resumed run would replay the stream from pass 0."""
    base = CONFIG.format(model="higgs+u1", it=7, do_flow=0)
    d1, d2, d3 = tmp_path / "straight", tmp_path / "resumed", tmp_path / "replayed"
    for d in (d1, d2, d3):
        d.mkdir()
    run_ref_gpu("hu", str(d1), base, mt_seed=5)
    for d in (d2, d3):
        run_ref_gpu("hu", str(d), base.replace("iterations 7", "iterations 4"), mt_seed=5)
        assert os.path.exists(d / "lattice.rng")
    os.remove(d3 / "lattice.rng")
    resume = base.replace("iterations 7", "iterations 8").replace("reset 1", "reset 0")
    out2 = run_ref_gpu("hu", str(d2), resume, mt_seed=99)
    run_ref_gpu("hu", str(d3), resume, mt_seed=5)
    k1 = open(d1 / "lattice.rng").read().split()
    k2 = open(d2 / "lattice.rng").read().split()
    assert k1 == k2  # same key, same pass count at the end
    assert "pass 0" not in out2
    p = make_params(L=(4, 6, 8), nhiggs=1, u1=1)
    _, f1 = lattice_fields(str(d1 / "lattice"), p)
    _, f2 = lattice_fields(str(d2 / "lattice"), p)
    _, f3 = lattice_fields(str(d3 / "lattice"), p)
    for k in ("su2link", "u1link", "doublet"):
        assert np.array_equal(f1[k], f2[k]), k
    assert not np.array_equal(f1["su2link"], f3["su2link"])
    assert open(d1 / "measure").read() == open(d2 / "measure").read()
