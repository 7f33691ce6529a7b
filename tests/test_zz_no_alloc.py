"""su2hip_measure_noalloc through the C ABI: the same sums as su2hip_measure,
bit for bit, on two calls in a row whose second lattice is the smaller one
(the scratch keeps the first call's partials, and every block overwrites
its own), in FP64 and FP32 and for two field contents; a lattice beyond
the scratch is refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from su2higgs_amd import abi
from su2higgs_amd.abi import FP32, FP64, make_params
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu

HIP_D2H = 2  # hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def device_sums():
    """Caller-owned device memory for the sums (the entry point allocates
    nothing; the caller passes its output)."""
    hip = C.CDLL("libamdhip64.so")
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), C.c_size_t(8 * abi.NSUMS)) == 0
    yield hip, ptr
    hip.hipFree(ptr)


def noalloc_sums(lib, hip, ptr, b, slot):
    rc = lib.su2hip_measure_noalloc(b.h, slot, C.cast(ptr, C.POINTER(C.c_double)))
    abi.check(rc)
    b.sync()
    out = np.zeros(abi.NSUMS)
    assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr, C.c_size_t(8 * abi.NSUMS), HIP_D2H) == 0
    return out


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("model", [dict(nhiggs=1, u1=1), dict(nhiggs=0, triplet=1)],
                         ids=["higgs_u1", "adjoint"])
def test_measure_noalloc_equals_measure(device_sums, precision, model):
    hip, ptr = device_sums
    lib = abi.load()
    big = make_params(L=(32, 24, 40), precision=precision, seed=3, **model)
    small = make_params(L=(8, 6, 4), precision=precision, seed=4, **model)
    b = Batch([big, small])
    for slot in (0, 1):
        b.hot_start(slot)
    b.update(2)
    for slot in (0, 1):  # the big lattice first, then the small one
        ref = b.measure(slot)
        got = noalloc_sums(lib, hip, ptr, b, slot)
        assert np.array_equal(got, ref), (slot, np.abs(got - ref).max())
    # and once more after further sweeps, small then big
    b.update(1)
    for slot in (1, 0):
        assert np.array_equal(noalloc_sums(lib, hip, ptr, b, slot), b.measure(slot))
    b.close()


def test_measure_noalloc_refuses_a_lattice_beyond_its_scratch(device_sums):
    hip, ptr = device_sums
    lib = abi.load()
    limit = 2 * 256 * 8192  # sites: SU2HIP_NOALLOC_MAX_BLOCKS blocks of 256 per parity
    p = make_params(L=(164, 160, 160), nhiggs=0, precision=FP32)
    assert abi.volume(p) > limit
    b = Batch(p)
    rc = lib.su2hip_measure_noalloc(b.h, 0, C.cast(ptr, C.POINTER(C.c_double)))
    assert rc == abi.ERANGE
    assert b"exceeds the scratch" in lib.su2hip_last_error()
    ok = make_params(L=(160, 160, 160), nhiggs=0, precision=FP32)  # 4096000 sites: fits
    assert abi.volume(ok) <= limit
    c = Batch(ok)
    c.hot_start(0)
    assert np.array_equal(noalloc_sums(lib, hip, ptr, c, 0), c.measure(0))
    b.close()
    c.close()
