"""GPU tests at the limits of a placement pass.

* An empty batch places nothing and reports no work.
* max_pitches (core.ml:300: stop once the results held max_pitches entries
  before the newest) bounds the baseball: placements equal the oracle's under the same
  preference, at 1 and 3 pitches.
* Queries longer than the register-resident stage-2 classes hold (gap-free,
  full-length reads) take the streaming form and are placed as the oracle
  places them (Glv.mask_into places any masked length, gmix_model.ml:121-141):
  a 1,600-column query at 2,000 taxa, 1,500-column queries at 10,000 taxa, in
  both forms; the other queries of the batch place exactly as without them.
* keep_at_most 0 keeps no placement and leaves the query placed with an empty
  placement list (core.ml:346-351, partitioni with i < 0 never true).
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


def _mutate(seq, rate, seed):
    """seq with a fraction `rate` of its bases substituted (seeded)."""
    rng = np.random.default_rng(seed)
    out = list(seq)
    for i in np.flatnonzero(rng.random(len(seq)) < rate):
        if out[i] in "ACGT":
            out[i] = "ACGT"[("ACGT".index(out[i]) + 1 + int(rng.integers(3))) % 4]
    return "".join(out)


def _long_query_case(oracle, name, seed, n_short, n_long):
    from pplacer_amd import placer, synth
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("workload %s seed %d, %d queries + %d gap-free ones" % (name, seed, n_short, n_long))
    wl = synth.make_workload(name, seed=seed, n_queries=n_short)
    leaves = list(wl.leaf_seqs().values())
    for k in range(n_long):
        wl.queries.insert(1 + 2 * k, ("full_%d" % k, _mutate(leaves[7 * k + 3], 0.02, seed + k)))
    model, pre, core, queries = oracle_setup(oracle, wl)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    assert q2 == queries
    try:
        long_idx = [1 + 2 * k for k in range(n_long)]
        N = len(queries[0])
        for i in long_idx:
            assert "-" not in queries[i] and len(queries[i]) == N
        _print_case("%s: fp64 batch with %d-column queries" % (name, N))
        got, status, _ = pl.place(queries, precision=PP_FP64)
        assert list(status[:len(queries)]) == [0] * len(queries)
        assert _compare_placements(core, queries, got, status, 1e-9, True) == len(queries)
        _print_case("%s: fp32 batch with %d-column queries" % (name, N))
        got32, status32, _ = pl.place(queries, precision=PP_FP32)
        assert _compare_placements(core, queries, got32, status32, 1e-4, False) == len(queries)
        _print_case("%s: fp64 batch without them" % name)
        rest = [q for i, q in enumerate(queries) if i not in long_idx]
        want, wstatus, _ = pl.place(rest, precision=PP_FP64)
        assert np.all(wstatus[:len(rest)] == 0)
        assert _fields([g for i, g in enumerate(got) if i not in long_idx]) == _fields(want)
    finally:
        pl.close()


def test_full_length_query_2000_taxa(oracle):
    _long_query_case(oracle, "gtr2000", 4, 6, 1)


def test_full_length_queries_10000_taxa(oracle):
    _long_query_case(oracle, "gtr10000", 5, 3, 2)


def test_keep_at_most_zero(gtr50_limits, oracle):
    from pplacer_amd.placer import PP_FP64, default_prefs
    pl, queries = gtr50_limits["pl"], gtr50_limits["queries"][:4]
    _print_case("keep_at_most 0")
    model, pre, core, oq = oracle_setup(oracle, gtr50_limits["wl"], keep_at_most=0)
    got, status, _ = pl.place(queries, prefs=default_prefs(keep_at_most=0), precision=PP_FP64)
    for q, query in enumerate(queries):
        st, want, _ = core.place(query)
        assert st == 0 and want == []
        assert status[q] == 0 and got[q] == []
