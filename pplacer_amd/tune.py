"""Launch-shape autotuner for placement passes.

Times whole passes over the resident queries (`Placer.place_resident`, no host
transfer inside) under candidate launch shapes and keeps the fastest.  A
launch shape is:

* the number of streams over which the 4-state baseball phase and the
  refinement phase spread their masked-length buckets, per precision form;
* the number of queries per device chunk;
* whether the pass's launches are captured into a HIP graph and replayed
  (the first pass of a new plan captures, the next ones replay; `pass_ms`
  keeps the best of its passes, so it times a replay).

Per-query results do not depend on the launch shape; only the time does, and
it depends on the device, the tree and the query mix.  The search is
coordinate-wise for one precision form: baseball streams, then refinement
streams, then chunk size, then graph mode, each against the best shape so far.  A candidate
replaces the best only if it is faster by more than `min_gain`, so that timing
noise does not move the choice.
"""
from __future__ import annotations

import time

from ._lib import PP_FP64

PITCH_STREAMS = (0, 2, 4, 8)
REFINE_STREAMS = (0, 2, 4, 8)
CHUNKS = (16384, 32768, 65536)
GRAPHS = (0, 1)


def pass_ms(pl, prefs, precision, reps):
    """Best wall time of `reps` passes over the resident queries, in ms."""
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        pl.place_resident(prefs, precision)
        pl.synchronize()
        best = min(best, 1e3 * (time.perf_counter() - t0))
    return best


def tune(pl, prefs, precision=PP_FP64, reps=2, pitch_streams=PITCH_STREAMS,
         refine_streams=REFINE_STREAMS, chunks=CHUNKS, graphs=GRAPHS, min_gain=0.01, log=None):
    """Tunes the launch shape of `precision` passes over the queries already
    resident in `pl` and leaves the best shape set.  Returns (best shape,
    trials), trials being (shape, ms) pairs in the order timed."""
    f = 0 if precision == PP_FP64 else 1
    pl.place_resident(prefs, precision)  # first pass: buffers and code objects
    pl.synchronize()
    trials = []

    def trial(cfg):
        shape = pl.set_launch_config(**cfg)
        ms = pass_ms(pl, prefs, precision, reps)
        trials.append((shape, ms))
        if log:
            log("tune %s: %.1f ms" % (shape, ms))
        return shape, ms

    best, best_ms = trial(pl.launch_config())
    for field, values in (("pitch_streams", pitch_streams), ("refine_streams", refine_streams)):
        for v in values:
            if v == best[field][f]:
                continue
            cand = dict(best)
            cand[field] = list(best[field])
            cand[field][f] = v
            shape, ms = trial(cand)
            if ms < best_ms * (1 - min_gain):
                best, best_ms = shape, ms
    for ch in chunks:
        if ch == best["chunk"]:
            continue
        shape, ms = trial(dict(best, chunk=ch))
        if ms < best_ms * (1 - min_gain):
            best, best_ms = shape, ms
    for g in graphs:
        if g == best["graph"]:
            continue
        shape, ms = trial(dict(best, graph=g))
        if ms < best_ms * (1 - min_gain):
            best, best_ms = shape, ms
    pl.set_launch_config(**best)
    return best, trials
