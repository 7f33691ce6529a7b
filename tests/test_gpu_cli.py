"""The `su2hip <config>` driver end to end: the files it writes against the
reference build following the same Philox trajectory (same seed, pass
counter from 0), in FP64."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import orc
from su2higgs_amd import abi
from su2higgs_amd.abi import make_params, ptr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "su2higgs_amd", "bin", "su2hip")

CONFIG = """dim 3
L1 4
L2 6
L3 8
model {model}
seed 4242
reset 1
iterations {it}
n_thermalize 0
interval 1
checkpoint 1000
run_checks 0
resultsfile measure
latticefile lattice
betasu2 8.0
betau1 9.0
r_u1 1
msq -0.25
lambda 0.02
phi0 1.0
algorithm_su2link heatbath
algorithm_u1link metropolis
algorithm_su2doublet overrelax
scalar_sweeps 3
update_doublet 1
update_links 1
multicanonical 0
measure_local 0
do_flow {do_flow}
flow_interval 5
flow_meas_interval 3
flow_dt 0.05
flow_t_max 0.5
"""


def same_lines(ours, ref):
    """measure lines byte for byte, except columns that are 0 up to rounding
    noise whose digits depend on the summation order (the total magnetic
    charge, a cold start's covariant singlet term)."""
    assert len(ours) == len(ref)
    for a, b in zip(ours, ref):
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb)
        for x, y in zip(ta, tb):
            assert x == y or (abs(float(x)) < 1e-10 and abs(float(y)) < 1e-10), (a, b)


def run_cli(d, cfg):
    with open(os.path.join(d, "config"), "w") as f:
        f.write(cfg)
    r = subprocess.run([EXE, "config"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def ref_trajectory(tag, p, n, flow_every=0, d=None, corr_every=0):
    """measure lines of the reference (and measure_flow / correl0 written
    into d), in the order of its main loop (src/main.c:302-353)."""
    r = orc.Ref(p, gradflow=bool(flow_every), correlators=bool(corr_every))
    o = orc.Oracle(p)
    o.cold_start(phi0=1.0)
    r.set_fields(o.get_fields())
    r.rng_set(p.seed, 0)
    lines = []
    cwd = os.getcwd()
    if d:
        os.chdir(d)
    try:
        for it in range(1, n + 1):
            lines.append(r.measure_text())
            if flow_every and it % flow_every == 0:
                r.lib.ref_grad_flow(0.5, 0.05, 3, it // flow_every)
            if corr_every and it % corr_every == 0:
                r.lib.ref_measure_correlators(b"correl0", r.lib.ref_longest_dir(), it // corr_every)
            r.update(1)
    finally:
        os.chdir(cwd)
    return r, lines


def test_cli_measure_and_lattice_match_reference(tmp_path):
    if not orc.ref_available("hu"):
        pytest.skip("reference build absent")
    d = str(tmp_path)
    run_cli(d, CONFIG.format(model="higgs+u1", it=6, do_flow=0))
    p = make_params(L=(4, 6, 8), nhiggs=1, u1=1, betasu2=8.0, betau1=9.0, msq_phi=-0.25,
                    lambda_phi=0.02, scalar_sweeps=3, seed=4242)
    r, lines = ref_trajectory("hu", p, 6)
    ours = open(os.path.join(d, "measure")).read().splitlines(keepends=True)
    assert ours == lines
    # labels of the same build
    os.chdir(d)
    r.lib.ref_print_labels()
    # (ref_print_labels overwrote ./labels with the reference's; the driver wrote the same)
    lib = abi.load()
    assert lib.su2hip_labels_write(b"labels_ours", C.byref(p)) == 0
    assert open("labels_ours").read() == open("labels").read()
    # final lattice: header iter = iterations + 1, fields within rounding
    g = abi.field_arrays(p)
    it = C.c_int64()
    assert lib.su2hip_lattice_read(b"lattice", C.byref(p), C.byref(it), None, None,
                                   *[ptr(g[k]) for k in orc.FIELDS]) == 0
    assert it.value == 7
    f = r.get_fields()
    for k in orc.FIELDS:
        if g[k] is not None:
            assert np.abs(g[k] - f[k]).max() < 1e-10


TWO_HIGGS = dict(msq_phi2=-0.15, m12sq_re=0.02, m12sq_im=0.01, lam2=0.03, lam3=0.015, lam4=-0.01,
                 lam5_re=0.008, lam5_im=0.004, lam6_re=0.005, lam6_im=-0.003, lam7_re=0.004,
                 lam7_im=0.002)


def test_cli_two_doublets_match_reference(tmp_path):
    """The two-doublet model (the reference's -DNHIGGS=2 -DU1 build) from a
    cold start: measure lines, labels and the final lattice."""
    if not orc.ref_available("du"):
        pytest.skip("reference build absent")
    d = str(tmp_path)
    extra = "".join("%s %r\n" % kv for kv in TWO_HIGGS.items())
    run_cli(d, CONFIG.format(model="2hdm+u1", it=6, do_flow=0) + extra)
    p = make_params(L=(4, 6, 8), nhiggs=2, u1=1, betasu2=8.0, betau1=9.0, msq_phi=-0.25,
                    lambda_phi=0.02, scalar_sweeps=3, seed=4242, **TWO_HIGGS)
    r, lines = ref_trajectory("du", p, 6)
    ours = open(os.path.join(d, "measure")).read().splitlines(keepends=True)
    same_lines(ours, lines)
    lib = abi.load()
    os.chdir(d)
    r.lib.ref_print_labels()
    assert lib.su2hip_labels_write(b"labels_ours", C.byref(p)) == 0
    assert open("labels_ours").read() == open("labels").read()
    g = abi.field_arrays(p)
    it = C.c_int64()
    assert lib.su2hip_lattice_read(b"lattice", C.byref(p), C.byref(it), None, None,
                                   *[ptr(g[k]) for k in orc.FIELDS]) == 0
    f = r.get_fields()
    for k in ("su2link", "u1link", "doublet"):
        assert np.abs(g[k] - f[k]).max() < 1e-10, k


def test_cli_heatbath_trajectories_match_reference(tmp_path):
    """The HB_TRAJECTORY build's main loop (src/main.c:321-328): a
    multicanonical run with a set of trajectories at iteration 3; the
    measure lines and the trajectory file against the reference's own
    make_realtime_trajectories on the same Philox stream."""
    if not orc.ref_available("hr"):
        pytest.skip("reference build absent")
    from test_gpu_traj import write_weight
    d = str(tmp_path)
    os.chdir(d)
    write_weight("weight", 40, 0.0, 2.0)
    cfg = CONFIG.format(model="higgs+trajectory", it=4, do_flow=0).replace(
        "multicanonical 0", "multicanonical 1")
    cfg += ("do_trajectory 1\nweightfile weight\nbins 40\nmin 0.0\nmax 2.0\nchecks_per_sweep 2\n"
            "orderparam phisq\nmuca_mode 0\nmuca_delta 0.3\n")
    with open("realtime_config", "w") as f:
        f.write("n_traj 3\ninterval 1\nmode_interval 3\nmin 0.7\nmax 0.95\n")
    run_cli(d, cfg)
    ours_m = open("measure").read().splitlines(keepends=True)
    ours_t = open("trajectory").read()
    os.rename("trajectory", "trajectory_ours")
    p = make_params(L=(4, 6, 8), nhiggs=1, betasu2=8.0, msq_phi=-0.25, lambda_phi=0.02,
                    scalar_sweeps=3, seed=4242)
    r = orc.Ref(p, trajectory=True)
    o = orc.Oracle(p)
    o.cold_start(phi0=1.0)
    r.set_fields(o.get_fields())
    r.rng_set(p.seed, 0)
    m = abi.Muca()
    m.enabled, m.orderparam, m.mode, m.checks_per_sweep = 1, abi.PHISQ, abi.READONLY, 2
    m.bins, m.min, m.max, m.delta = 40, 0.0, 2.0, 0.3
    assert r.lib.ref_muca_load(b"weight", C.byref(m)) == 0
    lines = []
    for it in range(1, 5):
        lines.append(r.measure_text())
        if it % 3 == 0:
            r.lib.ref_trajectories(3, 1, 0.7, 0.95, it // 3)
        r.update(1)
    same_lines(ours_m, lines)
    ref_t = "===== Realtime trajectories: min=0.7, max=0.95, n_traj=3 ===== \n" + \
        open("trajectory").read()
    a, c = ours_t.splitlines(), ref_t.splitlines()
    assert len(a) == len(c) > 10
    for x, y in zip(a, c):
        if x != y:
            u, v = x.split(), y.split()
            assert len(u) == len(v) == 2 and u[0] == v[0], (x, y)
            assert abs(float(u[1]) - float(v[1])) <= 1.01e-8, (x, y)


def test_cli_resume_is_exact(tmp_path):
    """6 sweeps in one run = 3 + 3 with a restart from the final checkpoint
    (the Philox position travels in lattice.rng).  As in the reference, the
    final save stores iter = iterations + 1 and a resumed run starts at
    iter + 1 (src/main.c:284-285, :389), so the second run continues to
    iteration 7 to do three more sweeps."""
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    os.makedirs(a)
    os.makedirs(b)
    run_cli(a, CONFIG.format(model="higgs+u1", it=6, do_flow=0))
    run_cli(b, CONFIG.format(model="higgs+u1", it=3, do_flow=0))
    cfg = CONFIG.format(model="higgs+u1", it=7, do_flow=0).replace("reset 1", "reset 0")
    run_cli(b, cfg)
    ma = open(os.path.join(a, "measure")).read()
    mb = open(os.path.join(b, "measure")).read()
    assert ma == mb and len(ma.splitlines()) == 6
    la = open(os.path.join(a, "lattice"), "rb").read()
    lb = open(os.path.join(b, "lattice"), "rb").read()
    # same fields; the headers differ in the stored iteration and wall time
    assert la[:20] == lb[:20] and la[44:] == lb[44:]


def test_cli_gradient_flow_matches_reference(tmp_path):
    if not orc.ref_available("hg"):
        pytest.skip("reference build absent")
    d_ours, d_ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    os.makedirs(d_ours)
    os.makedirs(d_ref)
    run_cli(d_ours, CONFIG.format(model="higgs+gradflow", it=10, do_flow=1))
    p = make_params(L=(4, 6, 8), nhiggs=1, betasu2=8.0, msq_phi=-0.25, lambda_phi=0.02,
                    scalar_sweeps=3, seed=4242)
    r, lines = ref_trajectory("hg", p, 10, flow_every=5, d=d_ref)
    assert open(os.path.join(d_ours, "measure")).read().splitlines(keepends=True) == lines
    assert open(os.path.join(d_ours, "measure_flow")).read() == \
        open(os.path.join(d_ref, "measure_flow")).read()


@pytest.mark.parametrize("model,tag,kw", [
    ("higgs+correlators", "hc", dict(nhiggs=1)),
    ("triplet+correlators", "tc", dict(nhiggs=0, triplet=1)),
])
def test_cli_correlators_match_reference(tmp_path, model, tag, kw):
    """`correl0` and `labels_correl` of a run with correlators every 2nd
    iteration, byte for byte against the reference's CORRELATORS build on
    the same trajectory."""
    if not orc.ref_available(tag):
        pytest.skip("reference build absent")
    d_ours, d_ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    os.makedirs(d_ours)
    os.makedirs(d_ref)
    cfg = CONFIG.format(model=model, it=6, do_flow=0)
    cfg += "do_correlators 1\ncorrelator_interval 2\nmsq_triplet -0.3\nb4 0.1\nsigma0 0.4\n"
    cfg += "algorithm_su2triplet overrelax\nupdate_triplet 1\n"
    run_cli(d_ours, cfg)
    p = make_params(L=(4, 6, 8), **kw, betasu2=8.0, msq_phi=-0.25, lambda_phi=0.02, scalar_sweeps=3,
                    msq_triplet=-0.3, b4=0.1, seed=4242)
    r = orc.Ref(p, correlators=True)
    o = orc.Oracle(p)
    o.cold_start(phi0=1.0, sigma0=0.4)
    r.set_fields(o.get_fields())
    r.rng_set(p.seed, 0)
    cwd = os.getcwd()
    os.chdir(d_ref)
    try:
        lines = []
        for it in range(1, 7):
            lines.append(r.measure_text())
            if it % 2 == 0:
                r.lib.ref_measure_correlators(b"correl0", r.lib.ref_longest_dir(), it // 2)
            r.update(1)
        r.lib.ref_print_labels_correlators()
    finally:
        os.chdir(cwd)
    same_lines(open(os.path.join(d_ours, "measure")).read().splitlines(keepends=True), lines)
    assert open(os.path.join(d_ours, "labels_correl")).read() == \
        open(os.path.join(d_ref, "labels_correl")).read()
    ours = open(os.path.join(d_ours, "correl0")).read()
    assert ours == open(os.path.join(d_ref, "correl0")).read()
    assert ours.count("Measurement id") == 3


def test_cli_zprofile_matches_reference(tmp_path):
    """`measure_z` and `labels_z` of a run with z profiles every 2nd
    iteration against the reference's MEASURE_Z build on the same
    trajectory (cold start; setup_wall 0)."""
    if not orc.ref_available("hsuz"):
        pytest.skip("reference build absent")
    d_ours, d_ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    os.makedirs(d_ours)
    os.makedirs(d_ref)
    cfg = CONFIG.format(model="higgs+u1+singlet+measure_z", it=6, do_flow=0)
    cfg += ("interval_z 2\ndo_z_meas 1\nsetup_wall 0\nupdate_singlet 1\nsinglet0 0.3\n"
            "algorithm_singlet overrelax\nmsq_s -0.1\nb1_s 0.2\nb3_s 0.1\nb4_s 0.3\n"
            "a1_s 0.3\na2_s 0.25\n")
    run_cli(d_ours, cfg)
    p = make_params(L=(4, 6, 8), nhiggs=1, u1=1, singlet=1, betasu2=8.0, betau1=9.0,
                    msq_phi=-0.25, lambda_phi=0.02, scalar_sweeps=3, msq_s=-0.1, b1_s=0.2,
                    b3_s=0.1, b4_s=0.3, a1_s=0.3, a2_s=0.25, seed=4242)
    r = orc.Ref(p, measure_z=True)
    o = orc.Oracle(p)
    o.cold_start(phi0=1.0, singlet0=0.3)
    r.set_fields(o.get_fields())
    r.rng_set(p.seed, 0)
    cwd = os.getcwd()
    os.chdir(d_ref)
    try:
        r.lib.ref_print_z_labels()
        lines = []
        for it in range(1, 7):
            lines.append(r.measure_text())
            if it % 2 == 0:
                r.lib.ref_measure_along_z(it // 2)
            r.update(1)
    finally:
        os.chdir(cwd)
    same_lines(open(os.path.join(d_ours, "measure")).read().splitlines(keepends=True), lines)
    assert open(os.path.join(d_ours, "labels_z")).read() == open(os.path.join(d_ref, "labels_z")).read()
    ours = open(os.path.join(d_ours, "measure_z")).read().split()
    ref = open(os.path.join(d_ref, "measure_z")).read().split()
    assert len(ours) == len(ref) == 3 * 8 * 4  # z, phi^2, S, S^2 per plane
    for x, y in zip(ours, ref):
        assert x == y or abs(float(x) - float(y)) <= 1e-5 * abs(float(y))


def test_cli_blocked_correlators_match_reference(tmp_path):
    """`correl0`, `correl1`, `correl2` of a triplet run with blocking (blocks
    3, capped at 2 levels for 8 x 8 x 12) against the reference's BLOCKING
    build on the same trajectory."""
    if not orc.ref_available("tcb"):
        pytest.skip("reference build absent")
    d_ours, d_ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    os.makedirs(d_ours)
    os.makedirs(d_ref)
    cfg = CONFIG.format(model="triplet+blocking", it=4, do_flow=0).replace("L1 4\n", "L1 8\n")
    cfg = cfg.replace("L2 6\n", "L2 8\n").replace("L3 8\n", "L3 12\n")
    cfg += ("do_correlators 1\ncorrelator_interval 2\nblocks 3\nmsq_triplet -0.3\nb4 0.1\n"
            "sigma0 0.4\nalgorithm_su2triplet overrelax\nupdate_triplet 1\n")
    run_cli(d_ours, cfg)
    p = make_params(L=(8, 8, 12), nhiggs=0, triplet=1, betasu2=8.0, scalar_sweeps=3, msq_triplet=-0.3,
                    b4=0.1, seed=4242)
    r = orc.Ref(p, blocking=True)
    o = orc.Oracle(p)
    o.cold_start(sigma0=0.4)
    r.set_fields(o.get_fields())
    r.rng_set(p.seed, 0)
    cwd = os.getcwd()
    os.chdir(d_ref)
    try:
        assert r.lib.ref_blocked_setup(3) == 2
        for it in range(1, 5):
            if it % 2 == 0:
                r.lib.ref_measure_correlators(b"correl0", r.lib.ref_longest_dir(), it // 2)
                r.lib.ref_blocked_correlators(it // 2)
            r.update(1)
    finally:
        os.chdir(cwd)
    for fn in ("correl0", "correl1", "correl2"):
        ours = open(os.path.join(d_ours, fn)).read().split()
        ref = open(os.path.join(d_ref, fn)).read().split()
        assert len(ours) == len(ref) > 0, fn
        for x, y in zip(ours, ref):
            if x == y:
                continue
            fx, fy = float(x), float(y)
            assert abs(fx - fy) <= 1e-5 * abs(fy) or (abs(fx) < 1e-12 and abs(fy) < 1e-12), (fn, x, y)
    assert not os.path.exists(os.path.join(d_ours, "correl3"))
