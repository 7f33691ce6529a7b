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


@pytest.mark.parametrize("fp32", [True, False], ids=["fp32", "fp64"])
def test_multicanonical_48_many_chains_match_reference_histogram(fp32):
    """16 chains of the same run in one handle (every launch covers all of
    them), pooled, against the reference's four chains pooled: the order
    parameter's mean within 3.5 combined standard errors (ours from the
    chain-to-chain spread, the reference's blocked), the same per-chain
    spread, and a histogram distance no larger than the largest distance
    between one reference chain and the other three (the reference's own
    chain-to-chain noise, which a biased distribution would exceed)."""
    if not os.path.exists(muca48.FIXTURE):
        pytest.skip("fixture absent")
    ours, _ = muca48.run_batch(fp32=fp32, chains=16, seed0=480000 if fp32 else 490000)
    c = muca48.compare_chains(ours, muca48.reference())
    print(c)
    assert abs(c["z"]) < 3.5, c
    assert 0.8 < c["spread_ratio"] < 1.25, c
    assert c["hist_tv_distance"] < max(c["ref_leave_one_out_tv"]), c
