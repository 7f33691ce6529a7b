"""su2higgs_amd: MI355X-native lattice update sweep with su2higgs-compatible I/O.

The compute path is libsu2hip.so (HIP kernels for gfx950 behind the C ABI in
include/su2hip.h); `abi` is its ctypes mirror and `lattice` a thin Python
handle over it.  Nothing here computes on the CPU.
"""
from . import abi  # noqa: F401
