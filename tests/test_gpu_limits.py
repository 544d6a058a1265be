"""GPU tests at the limits of a placement pass.

* An empty batch places nothing and reports no work.
* max_pitches (core.ml:300: stop once the results held max_pitches entries
  before the newest) bounds the baseball: placements equal the oracle's under the same
  preference, at 1 and 3 pitches.
* A query whose masked length exceeds what a stage-2 workgroup holds in LDS
  (896 columns at 2,000 taxa x 1,600 columns) gets status 7 ("query too long
  for the GPU path"), and the other queries of its batch place exactly as
  without it.
"""
import numpy as np
import pytest

from tests.helpers import oracle_setup
from tests.test_gpu_parity import _compare_placements

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _fields(rows):
    return [[(p.edge_num, p.likelihood, p.like_weight_ratio, p.distal_length, p.pendant_length)
             for p in r] for r in rows]


@pytest.fixture(scope="module")
def gtr50_limits(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload gtr50 seed 11, 24 queries")
    wl = synth.make_workload("gtr50", seed=11, n_queries=24)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    yield dict(wl=wl, pl=pl, queries=queries)
    pl.close()


def test_empty_batch(gtr50_limits):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    pl = gtr50_limits["pl"]
    for prec in (PP_FP64, PP_FP32):
        _print_case("empty batch, precision %d" % prec)
        got, status, stats = pl.place([], precision=prec)
        assert got == []
        assert stats.n_logdots == 0 and stats.n_like_calls == 0 and stats.n_pitched == 0
    # the context still places afterwards
    got, status, _ = pl.place(gtr50_limits["queries"][:2], precision=PP_FP64)
    assert list(status[:2]) == [0, 0] and all(len(g) > 0 for g in got)


@pytest.mark.parametrize("max_pitches", [1, 3])
def test_max_pitches_matches_oracle(gtr50_limits, oracle, max_pitches):
    from pplacer_amd.placer import PP_FP64, default_prefs
    pl, queries = gtr50_limits["pl"], gtr50_limits["queries"]
    _print_case("max_pitches %d" % max_pitches)
    model, pre, core, oq = oracle_setup(oracle, gtr50_limits["wl"], max_pitches=max_pitches)
    assert oq == queries
    got, status, stats = pl.place(queries, prefs=default_prefs(max_pitches=max_pitches),
                                  precision=PP_FP64)
    assert _compare_placements(core, queries, got, status, 1e-9, True) == len(queries)
    # results hold at most max_pitches + 1 pitched edges per query
    assert stats.n_pitched <= len(queries) * (max_pitches + 1)


def test_too_long_query_gets_status_7():
    from pplacer_amd import placer, synth
    from pplacer_amd.placer import PP_FP64
    _print_case("workload gtr2000 seed 4, 6 queries + a full-length one")
    wl = synth.make_workload("gtr2000", seed=4, n_queries=6)
    full = next(iter(wl.leaf_seqs().values()))  # a reference sequence: no gaps
    wl.queries.insert(3, ("full_length", full))
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    try:
        N = len(queries[0])
        assert N > 896 and "-" not in queries[3]
        _print_case("gtr2000 batch with a %d-column query" % N)
        got, status, _ = pl.place(queries, precision=PP_FP64)
        assert list(status[:len(queries)]) == [0, 0, 0, 7, 0, 0, 0]
        assert got[3] == []
        _print_case("gtr2000 batch without it")
        rest = queries[:3] + queries[4:]
        want, wstatus, _ = pl.place(rest, precision=PP_FP64)
        assert np.all(wstatus[:len(rest)] == 0)
        assert _fields(got[:3] + got[4:]) == _fields(want)
    finally:
        pl.close()
