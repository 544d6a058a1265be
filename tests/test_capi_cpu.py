"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/pplacer_hip.h declares, and its host-only entry points (model
building blocks, pre-masking) agree with the oracle.  No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests.helpers import GOLDEN  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pplacer_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pplacer_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python layer binds all of them
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)


def test_discrete_gamma_matches_reference_values(oracle):
    from pplacer_amd.placer import discrete_gamma
    want = [0.0333877533835995541, 0.251915917593438454, 0.820268481973647878,
            2.89442784704931411]
    np.testing.assert_allclose(discrete_gamma(4, 0.5), want, rtol=1e-9)
    for k, a in [(1, 1.0), (4, 0.771), (8, 0.2)]:
        np.testing.assert_allclose(discrete_gamma(k, a), oracle.discrete_gamma(k, a), rtol=1e-12)


def test_emper_freq_and_premask_match_oracle(oracle):
    from pplacer_amd import placer, synth
    wl = synth.make_workload("gtr12", seed=9)
    ref = [s for _, s in wl.ref_aln]
    qs = [s for _, s in wl.queries]
    np.testing.assert_array_equal(placer.emper_freq(4, ref), oracle.emper_freq(4, ref))
    mask, keep, n = placer.premask(4, ref, qs)
    st, omask, okeep, on = oracle.premask(4, ref, qs)
    assert st == 0 and n == on
    np.testing.assert_array_equal(mask, omask)
    np.testing.assert_array_equal(keep, okeep)
    with pytest.raises(Exception):
        placer.premask(4, ref, ["-" * len(ref[0])])


def test_prefs_defaults():
    from pplacer_amd.placer import default_prefs
    p = default_prefs()
    assert (p.start_pend, p.max_pend, p.initial_tolerance, p.strike_box, p.keep_factor) == \
        (0.1, 2.0, 0.01, 3.0, 0.01)
    assert (p.max_strikes, p.max_pitches, p.keep_at_most) == (6, 40, 7)
