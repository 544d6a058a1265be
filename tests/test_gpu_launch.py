"""GPU tests of the launch shape of a placement pass (pp_launch_config_get /
pp_launch_config_set) and of the autotuner that sets it.

* The defaults read back; out-of-range values are rejected with PP_EINVAL
  and change nothing.
* Placements do not depend on the launch shape: 3,000 queries (50 taxa)
  placed with every phase on the main stream and 1,000-query chunks, or with
  2 and 8 bucket streams, equal the default shape's results field for field
  and bit for bit, in both forms.
* The tuner leaves a valid shape set, and placements under it equal the
  default shape's.
* Graph mode (the pass's launches captured into a HIP graph, replayed by the
  next passes with the same plan): every replay equals the eager results bit
  for bit, with eager passes, other forms and other chunkings in between, so
  a captured graph that went stale or a replay that read stale counters
  would show.
"""
import pytest

pytestmark = pytest.mark.gpu

NQ = 3000
DEFAULT = {"pitch_streams": [0, 8], "refine_streams": [4, 4], "chunk": 65536, "graph": 0}


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _fields(rows):
    return [[(p.edge_num, p.likelihood, p.like_weight_ratio, p.distal_length, p.pendant_length)
             for p in r] for r in rows]


@pytest.fixture(scope="module")
def gtr50_3k():
    from pplacer_amd import placer, synth
    _print_case("workload gtr50 seed 9, %d queries" % NQ)
    wl = synth.make_workload("gtr50", seed=9, n_queries=NQ)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    assert pl.launch_config() == DEFAULT
    want = {}
    for form in ("fp64", "fp32"):
        prec = placer.PP_FP64 if form == "fp64" else placer.PP_FP32
        got, status, _ = pl.place(queries, precision=prec)
        want[form] = (_fields(got), list(status[:NQ]))
    yield dict(pl=pl, queries=queries, want=want)
    pl.close()


def test_launch_config_validation(gtr50_3k):
    from pplacer_amd._lib import PPError
    pl = gtr50_3k["pl"]
    _print_case("launch config validation")
    for bad in ({"pitch_streams": [9, 0]}, {"refine_streams": [0, -1]}, {"chunk": 0},
                {"chunk": 262145}, {"graph": 2}):
        with pytest.raises(PPError):
            pl.set_launch_config(**bad)
        assert pl.launch_config() == DEFAULT
    assert pl.set_launch_config(chunk=262144)["chunk"] == 262144
    assert pl.set_launch_config(**DEFAULT) == DEFAULT


@pytest.mark.parametrize("shape", [
    {"pitch_streams": [0, 0], "refine_streams": [0, 0], "chunk": 1000, "graph": 0},
    {"pitch_streams": [2, 2], "refine_streams": [8, 2], "chunk": 1200, "graph": 0},
    {"pitch_streams": [8, 4], "refine_streams": [2, 0], "chunk": 65536, "graph": 0},
    {"pitch_streams": [4, 8], "refine_streams": [4, 8], "chunk": 1000, "graph": 1},
])
def test_results_do_not_depend_on_launch_shape(gtr50_3k, shape):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    pl, queries = gtr50_3k["pl"], gtr50_3k["queries"]
    try:
        assert pl.set_launch_config(**shape) == shape
        for form, prec in (("fp64", PP_FP64), ("fp32", PP_FP32)):
            _print_case("launch shape %s, %s" % (shape, form))
            got, status, _ = pl.place(queries, precision=prec)
            assert (_fields(got), list(status[:NQ])) == gtr50_3k["want"][form]
    finally:
        pl.set_launch_config(**DEFAULT)


def test_tuner_sets_a_valid_shape(gtr50_3k):
    from pplacer_amd import placer, tune
    pl, queries = gtr50_3k["pl"], gtr50_3k["queries"]
    prefs = placer.default_prefs()
    try:
        for form, prec in (("fp64", placer.PP_FP64), ("fp32", placer.PP_FP32)):
            _print_case("tune %s" % form)
            pl.upload_queries(queries)
            best, trials = tune.tune(pl, prefs, prec, reps=1, chunks=(1000, 32768))
            assert pl.launch_config() == best
            assert len(trials) >= 2 and all(ms > 0 for _, ms in trials)
            got, status, _ = pl.place(queries, precision=prec)
            assert (_fields(got), list(status[:NQ])) == gtr50_3k["want"][form]
    finally:
        pl.set_launch_config(**DEFAULT)


def test_graph_replays_equal_eager_passes(gtr50_3k):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    pl, queries, want = gtr50_3k["pl"], gtr50_3k["queries"], gtr50_3k["want"]
    try:
        # FP64 form: capture, then replays with an eager pass, the other form
        # and another chunking (another plan, captured beside) in between
        for step, (form, shape) in enumerate([
                ("fp64", dict(DEFAULT, graph=1)), ("fp64", dict(DEFAULT, graph=1)),
                ("fp64", dict(DEFAULT, graph=0)), ("fp32", dict(DEFAULT, graph=1)),
                ("fp64", dict(DEFAULT, graph=1)), ("fp64", dict(DEFAULT, graph=1, chunk=1000)),
                ("fp32", dict(DEFAULT, graph=1)), ("fp64", dict(DEFAULT, graph=1))]):
            _print_case("graph replay step %d: %s, %s" % (step, form, shape))
            pl.set_launch_config(**shape)
            got, status, stats = pl.place(queries, precision=PP_FP64 if form == "fp64" else PP_FP32)
            assert (_fields(got), list(status[:NQ])) == want[form], step
            if shape["graph"]:
                assert stats.t_ml_ms > 0 and stats.t_pitch_ms == 0
    finally:
        pl.set_launch_config(**DEFAULT)
