"""GPU parity of a placement pass longer than one device chunk (set to 32,768
queries here).  Such a pass runs as a pipeline: chunks of even size, stage 1 of the
next chunk on its own stream into the second score buffer while stage 2 of the
current one runs, and (in the 4-state form) concurrent buckets.  Per-query
results must not depend on how a pass is cut into chunks.

* 80,000 queries (50 taxa, 1,000 columns) placed in one pass (three chunks of
  26,666-26,667, so a score buffer is reused) equal, field for field and bit
  for bit, the same queries placed in passes of 20,000, 30,000 and 30,000, each
  of which fits in one chunk.  This is checked in the FP64 and FP32 forms.
* Queries of every chunk of the long pass (first, middle, last) match the oracle
  to the FP64 bar: the same best edge, likelihood within 1e-9 relative.
"""
import pytest

from tests.helpers import oracle_setup
from tests.test_gpu_parity import _compare_placements

pytestmark = pytest.mark.gpu

NQ = 80000
PARTS = [(0, 20000), (20000, 50000), (50000, NQ)]


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


@pytest.fixture(scope="module")
def long_pass():
    from pplacer_amd import placer, synth
    _print_case("workload gtr50 seed 5, %d queries" % NQ)
    wl = synth.make_workload("gtr50", seed=5, n_queries=NQ)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    pl.set_launch_config(chunk=32768)
    yield dict(wl=wl, pl=pl, queries=queries)
    pl.close()


def _fields(rows):
    return [[(p.edge_num, p.likelihood, p.like_weight_ratio, p.distal_length, p.pendant_length)
             for p in r] for r in rows]


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_long_pass_equals_single_chunk_passes(long_pass, form):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    prec = PP_FP64 if form == "fp64" else PP_FP32
    pl, queries = long_pass["pl"], long_pass["queries"]
    _print_case("long pass %s, %d queries in one pass" % (form, NQ))
    got, status, _ = pl.place(queries, precision=prec)
    want, want_status = [], []
    for lo, hi in PARTS:
        _print_case("long pass %s, queries %d..%d in one pass" % (form, lo, hi))
        r, st, _ = pl.place(queries[lo:hi], precision=prec)
        want += _fields(r)
        want_status += list(st[:hi - lo])
    assert list(status[:NQ]) == want_status
    assert _fields(got) == want


def test_long_pass_chunks_match_oracle(long_pass, oracle):
    from pplacer_amd.placer import PP_FP64
    pl, queries = long_pass["pl"], long_pass["queries"]
    _print_case("long pass fp64 vs oracle")
    got, status, _ = pl.place(queries, precision=PP_FP64)
    # chunk boundaries 0, 26666, 53333, 80000
    pick = [0, 1, 26665, 26666, 26667, 40000, 53332, 53333, 79998, 79999]
    model, pre, core, oq = oracle_setup(oracle, long_pass["wl"])
    sub_q = [oq[i] for i in pick]
    assert sub_q == [queries[i] for i in pick]
    n = _compare_placements(core, sub_q, [got[i] for i in pick], [status[i] for i in pick], 1e-9, True)
    assert n == len(pick)
