"""The 48^3 SU(2)+adjoint multicanonical run (Sigmasq, READONLY weight of
200 bins on [0,2], 500 + 2000 sweeps) through the su2hip driver against the
reference's own four chains recorded in tests/golden/muca48_reference.npz:
the order parameter's blocked mean within 4 combined standard errors and
the same spread; the weight file is left unchanged (READONLY)."""
import os

import pytest

import muca48

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [True, False], ids=["fp32", "fp64"])
def test_multicanonical_48_matches_reference_statistics(fp32):
    if not os.path.exists(muca48.FIXTURE):
        pytest.skip("fixture absent")
    ours = muca48.run(fp32=fp32, seed=48 if fp32 else 49)
    assert len(ours["orderparam"]) == 2000
    assert ours["weight_unchanged"]
    c = muca48.compare(ours, muca48.reference())
    print(c)
    assert abs(c["z"]) < 4, c
    assert 0.5 < c["spread_ratio"] < 2.0, c
