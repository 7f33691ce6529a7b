"""The reference-side binding su2higgs_amd/bind/su2hip_bind.c, which
INTEGRATION.md shows in full, compiles against the reference's own headers and
include/su2hip.h for the builds it covers: one doublet with U(1), a singlet or
a triplet, two doublets, and the adjoint-only build.  Compile only (gcc -c
-Wall -Werror); skipped where the reference sources are absent.  Running it
inside the reference's main loop is tests/test_gpu_binding.py."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
BUILDS = ["-DNHIGGS=1 -DU1", "-DNHIGGS=2", "-DNHIGGS=2 -DU1", "-DNHIGGS=1 -DTRIPLET",
          "-DNHIGGS=1 -DSINGLET -DU1", "-DNHIGGS=0 -DTRIPLET"]


BINDING = os.path.join(ROOT, "su2higgs_amd", "bind", "su2hip_bind.c")


def test_integration_shows_the_binding_verbatim():
    """The one ```c block that opens with the file's own header line is the
    file, byte for byte; the other blocks are call sequences."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = [b for b in re.findall(r"```c\n(.*?)```", text, re.S)
              if b.startswith("/* su2hip_bind.c")]
    assert blocks == [open(BINDING).read()]
    # the call sequence names only functions the binding defines
    defined = set(re.findall(r"^int (su2hip_bind_\w+)\(", blocks[0], re.M))
    called = set(re.findall(r"(su2hip_bind_\w+)\(", text))
    assert called <= defined, called - defined


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "su2.h")) or not shutil.which("gcc"),
                    reason="reference sources or gcc absent")
@pytest.mark.parametrize("flags", BUILDS)
def test_binding_stub_compiles_against_reference_headers(flags, tmp_path):
    cmd = (["gcc", "-std=gnu11", "-Wall", "-Werror", "-c"] + flags.split() +
           ["-I" + REF_SRC, "-I" + os.path.join(ROOT, "include"), BINDING, "-o",
            str(tmp_path / "bind.o")])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
