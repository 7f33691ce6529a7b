"""The reference's own driver with its sweep on the GPU.

oracle/_ref/su2_<tag>_gpu is the reference's unmodified main loop
(src/main.c) linked with the reference-side binding
su2higgs_amd/bind/su2hip_bind.c: update_lattice is routed to libsu2hip
(oracle/bind_shim.c).  On the same config it must write the `measure` file
the su2hip driver writes: both run the same GPU sweeps (Philox seed 1234,
pass counter from 0); the reference measures on the CPU with its own
measure(), su2hip on the GPU.  The final `lattice` files hold bitwise the
same fields."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import make_params, ptr
from test_gpu_cli import CONFIG, TWO_HIGGS, run_cli, same_lines

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINDING_SEED = 1234  # the Philox key su2hip_bind.c gives the lattice


def run_ref_gpu(tag, d, cfg):
    exe = os.path.join(ROOT, "oracle", "_ref", "su2_%s_gpu" % tag)
    if not os.path.exists(exe):
        pytest.skip("reference build with the binding absent")
    with open(os.path.join(d, "config"), "w") as f:
        f.write(cfg)
    r = subprocess.run([exe, "config"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


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
    cfg = CONFIG.format(model=model, it=8, do_flow=0).replace(
        "seed 4242", "seed %d" % BINDING_SEED).replace("checkpoint 1000", "checkpoint 4") + extra
    dr, dc = tmp_path / "ref", tmp_path / "su2hip"
    dr.mkdir()
    dc.mkdir()
    out_ref = run_ref_gpu(tag, str(dr), cfg)
    out_cli = run_cli(str(dc), cfg)
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
    cfg = CONFIG.format(model="higgs+u1", it=8, do_flow=0).replace(
        "seed 4242", "seed %d" % BINDING_SEED).replace("multicanonical 0", "multicanonical 1")
    cfg = cfg.replace("checkpoint 1000", "checkpoint 4")
    cfg += ("weightfile weight\nbins 40\nmin 0.0\nmax 2.0\nchecks_per_sweep 2\n"
            "orderparam phisq\nmuca_mode %d\nmuca_delta 0.3\n" % mode)
    dr, dc = tmp_path / "ref", tmp_path / "su2hip"
    for d in (dr, dc):
        d.mkdir()
        write_weight(str(d / "weight"), 40, 0.0, 2.0)
    out_ref = run_ref_gpu("hu", str(dr), cfg)
    out_cli = run_cli(str(dc), cfg)
    ref = open(dr / "measure").read().splitlines(keepends=True)
    ours = open(dc / "measure").read().splitlines(keepends=True)
    assert len(ref) == 8
    same_lines(ours, ref)
    acc = acceptance_lines(out_ref)
    assert len(acc) == 2 and acc == acceptance_lines(out_cli), (acc, acceptance_lines(out_cli))
