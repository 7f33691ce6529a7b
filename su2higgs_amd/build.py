"""Build libsu2hip.so (HIP kernels for gfx950 + C ABI + host I/O layer) and
the su2hip command-line driver, in-tree.

    python -m su2higgs_amd.build
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
OBJDIR = os.path.join(ROOT, "_build")

ARCH = os.environ.get("SU2HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# no packed FP32 instructions in any HIP code object
NOPK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIPFLAGS = ["-O3", "--offload-arch=" + ARCH, "-fPIC", "-std=c++17", "-Wall",
            "-Wno-unused-function", "-Wno-unused-variable"] + NOPK
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                  if f.endswith(".c") and f != "su2hip_main.c")


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    objs = []
    for src in ("su2hip_kernels.hip", "su2hip_api.cpp"):
        o = os.path.join(OBJDIR, src + ".o")
        cmd = [HIPCC] + HIPFLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", o]
        jobs.append(cmd)
        objs.append(o)
    for src in host_sources():
        o = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        jobs.append(["gcc"] + CFLAGS + ["-I" + os.path.join(ROOT, "include"), "-c", src, "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=8) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(LIBDIR, "libsu2hip.so")
    tmp = lib + ".tmp"
    # roctx ranges on the host entry points (rocprofiler-sdk-roctx)
    _run([HIPCC, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lrocprofiler-sdk-roctx", "-lm"])
    os.replace(tmp, lib)
    main_c = os.path.join(HOST, "su2hip_main.c")
    if os.path.exists(main_c):
        os.makedirs(BINDIR, exist_ok=True)
        _run(["gcc"] + CFLAGS + ["-I" + os.path.join(ROOT, "include"), main_c, "-o",
                                 os.path.join(BINDIR, "su2hip"), "-L" + LIBDIR, "-lsu2hip",
                                 "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    return lib


def build_oracle(reference=True):
    """The oracle library, and the reference build when its sources exist."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s", "liboracle.so"])
    if reference and os.path.isdir("/root/reference/src"):
        _run(["make", "-C", odir, "-s", "-j8", "ref"])


if __name__ == "__main__":
    print(build(verbose=True))
    build_oracle()
    if shutil.which("hipcc") is None and not os.path.exists(HIPCC):
        sys.exit("hipcc not found")
