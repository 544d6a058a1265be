"""Host logic of the launch-shape autotuner (pplacer_amd/tune.py), on a stand-in
for the placer whose pass time is a known function of the launch shape: the
search must end on the fastest shape it can reach coordinate-wise, leave that
shape set, and ignore gains within its noise margin."""
import pytest

from pplacer_amd import tune
from pplacer_amd._lib import PP_FP32, PP_FP64


class TimedShapes:
    """place_resident advances a virtual clock by cost(shape, form)."""

    def __init__(self, cost, monkeypatch):
        self.cfg = {"pitch_streams": [0, 8], "refine_streams": [4, 8], "chunk": 32768, "graph": 0}
        self.cost = cost
        self.now = 0.0
        self.passes = 0
        monkeypatch.setattr(tune.time, "perf_counter", lambda: self.now)

    def launch_config(self):
        return {k: (list(v) if isinstance(v, list) else v) for k, v in self.cfg.items()}

    def set_launch_config(self, pitch_streams=None, refine_streams=None, chunk=None, graph=None):
        if pitch_streams is not None:
            self.cfg["pitch_streams"] = list(pitch_streams)
        if refine_streams is not None:
            self.cfg["refine_streams"] = list(refine_streams)
        if chunk is not None:
            self.cfg["chunk"] = chunk
        if graph is not None:
            self.cfg["graph"] = graph
        return self.launch_config()

    def place_resident(self, prefs, precision):
        self.passes += 1
        self.now += self.cost(self.cfg, 0 if precision == PP_FP64 else 1) / 1e3

    def synchronize(self):
        pass


def test_tune_finds_the_fastest_shape(monkeypatch):
    # FP32 passes: best at 4 baseball streams, 2 refinement streams, 65,536 queries per chunk,
    # graph replay
    def cost(cfg, f):
        return (100 + 10 * abs(cfg["pitch_streams"][f] - 4) + 7 * abs(cfg["refine_streams"][f] - 2)
                + (0 if cfg["chunk"] == 65536 else 5) - 3 * cfg["graph"])
    pl = TimedShapes(cost, monkeypatch)
    best, trials = tune.tune(pl, None, PP_FP32, reps=1)
    assert best == {"pitch_streams": [0, 4], "refine_streams": [4, 2], "chunk": 65536, "graph": 1}
    assert pl.launch_config() == best
    # the other form's entries are untouched
    assert best["pitch_streams"][0] == 0 and best["refine_streams"][0] == 4
    # the first trial is the starting shape; every trial timed one pass
    assert trials[0][0] == {"pitch_streams": [0, 8], "refine_streams": [4, 8], "chunk": 32768, "graph": 0}
    assert pl.passes == 1 + len(trials)


def test_tune_ignores_gains_within_the_margin(monkeypatch):
    # 0.5% faster with 2 FP64 baseball streams: below the 1% margin
    def cost(cfg, f):
        return 200 - (1.0 if cfg["pitch_streams"][f] == 2 else 0.0)
    pl = TimedShapes(cost, monkeypatch)
    best, _ = tune.tune(pl, None, PP_FP64, reps=1)
    assert best == {"pitch_streams": [0, 8], "refine_streams": [4, 8], "chunk": 32768, "graph": 0}


@pytest.mark.parametrize("reps", [1, 3])
def test_pass_ms_takes_the_best_of_reps(monkeypatch, reps):
    times = iter([50.0, 40.0, 45.0])
    pl = TimedShapes(lambda cfg, f: next(times), monkeypatch)
    assert tune.pass_ms(pl, None, PP_FP64, reps) == pytest.approx(50.0 if reps == 1 else 40.0)
