"""Launch shapes: results do not depend on the workgroup chunk or the
occupancy target of the pass kernels, and the autotuner leaves the lattices
exactly as it found them."""
import numpy as np
import pytest

from su2higgs_amd import abi
from su2higgs_amd.abi import FP32, FP64, make_params
from su2higgs_amd.lattice import Batch

pytestmark = pytest.mark.gpu

PASS_KINDS = range(abi.K_MEASURE)


def batch(precision):
    ps = [make_params(L=L, nhiggs=1, u1=1, precision=precision, link_overrelax=2,
                      reunitarize=1 if precision == FP32 else 0, seed=40 + i,
                      msq_phi=-0.2 - 0.01 * i)
          for i, L in enumerate([(8, 8, 8), (12, 10, 8), (16, 16, 16), (6, 8, 10)])]
    b = Batch(ps)
    for i in range(len(ps)):
        b.hot_start(i)
    return b, len(ps)


def state(b, n):
    return [(b.get_fields(i), b.counters(i), b.rng_get(i)) for i in range(n)]


def same(a, b):
    for (fa, ca, ra), (fb, cb, rb) in zip(a, b):
        if ra != rb or ca != cb:
            return False
        for k in fa:
            if fa[k] is not None and not np.array_equal(fa[k], fb[k]):
                return False
    return True


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("graph", [False, True])
def test_results_do_not_depend_on_the_chunk(precision, graph):
    out = {}
    for chunk in (64, 128, 256, 0):
        b, n = batch(precision)
        b.use_graph(graph)
        for k in PASS_KINDS:
            b.set_chunk(k, chunk)
        if chunk:
            assert b.chunk(abi.K_LINK_OVERRELAX) == chunk
        b.update(3)
        out[chunk] = state(b, n)
        b.close()
    assert same(out[64], out[256]) and same(out[128], out[256]) and same(out[0], out[256])


def test_autotune_restores_the_state_and_picks_a_chunk():
    b, n = batch(FP32)
    b.update(2)
    before = state(b, n)
    budgets = b.budgets()
    table = b.autotune(repeats=20)
    assert set(table) >= {"link_heatbath", "link_overrelax", "u1_metropolis", "doublet_overrelax"}
    for name, ms in table.items():
        assert all(t > 0 for t in ms), (name, ms)
        k = abi.KIND_NAMES.index(name)
        assert b.chunk(k) in (64, 128, 256)
        assert b.waves(k) in ((0, 5) if k == abi.K_LINK_OVERRELAX else (0,))
    assert same(state(b, n), before)
    assert list(b.budgets()) == list(budgets)
    # and the tuned handle continues the chains exactly as an untuned one
    b.update(2)
    c, _ = batch(FP32)
    c.update(4)
    assert same(state(b, n), state(c, n))


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("graph", [False, True])
def test_results_do_not_depend_on_the_occupancy_target(precision, graph):
    """The link overrelaxation built for 5 waves per SIMD and built with the
    compiler's own registers make the same moves, bit for bit, launched one
    by one or from a captured graph."""
    out = {}
    for waves in (5, 0):
        b, n = batch(precision)
        b.use_graph(graph)
        b.set_waves(abi.K_LINK_OVERRELAX, waves)
        assert b.waves(abi.K_LINK_OVERRELAX) == waves
        b.update(3)
        out[waves] = state(b, n)
        b.close()
    assert same(out[5], out[0])


def test_occupancy_targets_each_kind_is_built_with():
    b, _ = batch(FP32)
    assert b.waves(abi.K_LINK_OVERRELAX) == 5  # the default
    assert b.waves(abi.K_LINK_HEATBATH) == 0
    for kind, waves in ((abi.K_LINK_HEATBATH, 5), (abi.K_LINK_OVERRELAX, 3),
                        (abi.K_DOUBLET_OVERRELAX, 5), (abi.K_MEASURE + 5, 0)):
        with pytest.raises(abi.Su2hipError):
            b.set_waves(kind, waves)
    b.set_waves(abi.K_LINK_HEATBATH, 0)
    b.set_waves(abi.K_LINK_OVERRELAX, 0)
    assert b.waves(abi.K_LINK_OVERRELAX) == 0
    b.close()
