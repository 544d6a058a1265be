"""GPU parity at the large stream shape: a 10,000-taxon tree (19,997 edges),
1,500 columns, GTR+Gamma4, the GLVs of every node resident on the device.

The oracle's precompute at this size takes ~15 s, so a handful of queries are
checked: stage-1 scores at all 19,997 edges (every supernode's masked logdot)
within 1e-12 relative with identical rankings, and full placements (FP64: the
same edges, likelihoods within 1e-9 relative; FP32 form: 1e-4 absolute) as
at 2,000 taxa.
"""
import numpy as np
import pytest

from tests.helpers import oracle_setup
from tests.test_gpu_parity import _compare_placements

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


@pytest.fixture(scope="module")
def gtr10000(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload gtr10000 seed 3, 5 queries")
    wl = synth.make_workload("gtr10000", seed=3, n_queries=5)
    model, pre, core, queries = oracle_setup(oracle, wl)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    assert q2 == queries
    yield dict(wl=wl, core=core, pl=pl, queries=queries)
    pl.close()


def test_scale_stage1_every_edge(gtr10000):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("gtr10000 stage1")
    pl, core, queries = gtr10000["pl"], gtr10000["core"], gtr10000["queries"]
    s64 = pl.stage1_scores(queries[:2], 0.1, PP_FP64)
    s32 = pl.stage1_scores(queries[:2], 0.1, PP_FP32)
    for q in range(2):
        want = core.stage1_scores(queries[q])
        assert want.shape == (19997,)
        np.testing.assert_allclose(s64[q], want, rtol=1e-12, atol=0)
        np.testing.assert_array_equal(core.rank(s64[q]), core.rank(want))
        assert np.max(np.abs(s32[q] - want)) < 5e-3


def test_scale_placements(gtr10000):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("gtr10000 place fp64/fp32")
    pl, core, queries = gtr10000["pl"], gtr10000["core"], gtr10000["queries"]
    got, status, stats = pl.place(queries, precision=PP_FP64)
    assert _compare_placements(core, queries, got, status, 1e-9, True) == len(queries)
    got32, status32, _ = pl.place(queries, precision=PP_FP32)
    ties = [0]
    assert _compare_placements(core, queries, got32, status32, 1e-4, False, ties) + ties[0] == len(queries)
