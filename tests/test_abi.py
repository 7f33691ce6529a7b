"""The C ABI without a GPU: libsu2hip.so loads and exports every function the
headers in include/ declare, the ctypes mirrors have the C layouts, and the
host-only parts of the request loop (synthetic stream, update counts)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from su2higgs_amd import abi, serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
HEADERS = sorted(f for f in os.listdir(INCLUDE) if f.endswith(".h"))


def declared():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(su2hip_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_function_is_exported():
    names = declared()
    assert len(names) > 40
    lib = C.CDLL(abi.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signatures cover them all
    assert set(names) <= set(_signed()), sorted(set(names) - set(_signed()))


def _signed():
    lib = abi.load()
    return [n for n in declared() if getattr(lib, n).argtypes is not None or
            getattr(lib, n).restype is not C.c_int]


STRUCTS = {
    "su2hip_params": abi.Params,
    "su2hip_counters": abi.Counters,
    "su2hip_muca": abi.Muca,
    "su2hip_run_config": abi.RunConfig,
    "su2hip_request": abi.Request,
    "su2hip_result": abi.Result,
    "su2hip_serve_config": abi.ServeConfig,
    "su2hip_serve_stats": abi.ServeStats,
    "su2hip_trajectory_config": abi.TrajectoryConfig,
}
RENAME = {"pass_": "pass"}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_struct_layouts_match_the_headers(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "su2hip.h"',
             '#include "su2hip_io.h"', '#include "su2hip_serve.h"', '#include "su2hip_traj.h"',
             "int main(void) {"]
    expect = []
    for cname, py in STRUCTS.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        expect.append(C.sizeof(py))
        for f, _ in py._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, RENAME.get(f, f)))
            expect.append(getattr(py, f).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I" + INCLUDE, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                                          check=True).stdout.split()]
    assert got == expect


def test_updates_per_sweep():
    # 16^3 mix: heatbath + 4 link OR rounds (15 link updates), 4 doublet OR +
    # the final Metropolis round (5)
    p = abi.make_params(L=(16, 16, 16), link_overrelax=4, scalar_sweeps=4)
    assert serve.updates_per_sweep(p) == 20
    p.u1 = 1
    assert serve.updates_per_sweep(p) == 23
    p = abi.make_params(L=(8, 8, 8), nhiggs=0, triplet=1, scalar_sweeps=2,
                        algorithm_su2triplet=abi.METROPOLIS)
    assert serve.updates_per_sweep(p) == 3 + 2


def test_stream_is_the_stated_distribution():
    n = 4000
    r = serve.stream(7, 20.0, n, flow_every=10)
    t = np.array([x.arrival for x in r])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert (gaps > 0).all()
    assert abs(gaps.mean() - 1 / 20) < 5 * (1 / 20) / np.sqrt(n)
    assert abs(gaps.std() - 1 / 20) < 0.1 / 20  # exponential: sd = mean
    L = np.array([x.L for x in r])
    flow = np.array([x.flow for x in r])
    assert (flow[9::10] == 1).all() and flow.sum() == n // 10
    assert (L[flow == 1] == 64).all()
    other = L[flow == 0]
    vals, cnt = np.unique(other, return_counts=True)
    assert list(vals) == [24, 32, 40, 48, 56, 64]
    assert (np.abs(cnt / len(other) - 1 / 6) < 0.03).all()
    for k, lo, hi in (("betasu2", 6, 12), ("msq_phi", -0.3, -0.1), ("lambda_phi", 0.005, 0.03)):
        v = np.array([getattr(x, k) for x in r])
        assert v.min() >= lo and v.max() <= hi
        assert abs(v.mean() - (lo + hi) / 2) < 0.02 * (hi - lo)
    s = np.array([x.sweeps for x in r])
    assert s.min() >= 100 and s.max() <= 1000 and abs(s.mean() - 550) < 15
    assert len({x.seed for x in r}) == n
    assert [x.id for x in r] == list(range(n))
    # deterministic in the seed
    r2 = serve.stream(7, 20.0, 50, flow_every=10)
    assert [(a.arrival, a.seed) for a in r2] == [(a.arrival, a.seed) for a in list(r)[:50]]


@pytest.mark.parametrize("kw,code,msg", [
    (dict(L=(7, 8, 8)), abi.EINVAL, b"even"),
    (dict(L=(0, 8, 8)), abi.EINVAL, b"even"),
    (dict(L=(1024, 1024, 512)), abi.ERANGE, b"too large"),       # 2^29 sites
    (dict(nhiggs=3), abi.EUNSUPPORTED, b"nhiggs"),
    (dict(nhiggs=2, triplet=1), abi.EUNSUPPORTED, b"not supported"),
    (dict(nhiggs=0, u1=1), abi.EUNSUPPORTED, b"not supported"),
    (dict(precision=16), abi.EINVAL, b"precision"),
    (dict(algorithm_su2link=abi.OVERRELAX), abi.EINVAL, b"algorithm_su2link"),
    (dict(scalar_sweeps=-1), abi.EINVAL, b"negative"),
    (dict(betasu2=0.0), abi.EINVAL, b"betasu2"),
])
def test_invalid_parameters_rejected_before_any_device_call(kw, code, msg):
    """su2hip_create and su2hip_footprint check the parameters before they
    touch a device: the error code and message, and no handle."""
    lib = abi.load()
    p = abi.make_params(**kw)
    h = C.c_void_p()
    assert lib.su2hip_create(1, C.byref(p), 1, 0, C.byref(h)) == code
    assert msg in lib.su2hip_last_error()
    assert not h.value
    nb = C.c_uint64()
    assert lib.su2hip_footprint(C.byref(p), 0, C.byref(nb)) == code


def test_batch_rules_and_largest_volume_without_a_device():
    lib = abi.load()
    h = C.c_void_p()
    # lattices of one batch share field content, precision and schedule
    ps = (abi.Params * 2)(abi.make_params(), abi.make_params(precision=abi.FP32))
    assert lib.su2hip_create(2, ps, 2, 0, C.byref(h)) == abi.EINVAL
    assert b"differs" in lib.su2hip_last_error()
    # an empty handle needs room for a lattice
    assert lib.su2hip_create(0, ps, 0, 0, C.byref(h)) == abi.EINVAL
    # the largest accepted volume (640^3 <= 2^28 sites): its FP64 fields
    # (12 link + 4 doublet + 3 U(1) words per site) fit one GPU's 288 GB
    p = abi.make_params(L=(640, 640, 640), u1=1)
    nb = C.c_uint64()
    assert lib.su2hip_footprint(C.byref(p), 1, C.byref(nb)) == abi.OK
    V = 640 ** 3
    assert 19 * 8 * V <= nb.value < 288e9
