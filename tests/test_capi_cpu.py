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


def test_price_cat_phylo_model_fixture():
    """The donovan refpkg's Price-CAT phylo_model (refpkg.ml:38-63,
    gcat_model.ml:453-478): 20 FastTree rates, 8062 1-based categories
    shifted to 0-based."""
    from pplacer_amd import refio
    from tests.helpers import ar53_price_cat
    pm = refio.read_phylo_model(os.path.join(GOLDEN, "price_cat", "ar53_phylo_model.json"))
    assert pm.ras_model == "Price-CAT" and pm.subs_model == "WAG"
    assert not pm.empirical_frequencies
    assert pm.n_cats == 20 and len(pm.cat_rates) == 20
    assert pm.cat_rates[0] == 0.046161 and pm.cat_rates[-1] == 18.464318
    rates, cats = ar53_price_cat()
    assert cats.shape == (8062,) and cats.min() == 0 and cats.max() == 16
    assert cats[:4].tolist() == [10, 8, 8, 7]  # the file's 11, 9, 9, 8
    with pytest.raises(ValueError, match="invalid ras_model"):
        refio.parse_phylo_model({"ras_model": "free", "subs_model": "WAG",
                                 "empirical_frequencies": False})


def test_cat_sites_follow_the_premask():
    """Gcat_model.mask_sites (gcat_model.ml:429-439) keeps the categories of
    the kept columns; a length mismatch is an error."""
    from pplacer_amd.run import cat_sites
    mask = np.array([1, 0, 1, 1, 0], dtype=np.uint8)
    assert cat_sites([4, 3, 2, 1, 0], mask).tolist() == [4, 2, 1]
    with pytest.raises(ValueError, match="site categories"):
        cat_sites([1, 2, 3], mask)


def test_cat_model_entry_points_reject_bad_input():
    """pp_model_build_cat / pp_site_categories argument checks run on the
    host before any device work."""
    from pplacer_amd import _lib
    L = _lib.lib()
    S = 4
    exch = np.ones((S, S))
    statd = np.full(S, 0.25)
    rates = np.array([0.5, -1.0])
    dp = C.POINTER(C.c_double)
    rc = L.pp_model_build_cat(None, S, 2, exch.ctypes.data_as(dp), statd.ctypes.data_as(dp),
                              rates.ctypes.data_as(dp))
    assert rc != 0
    cats = np.zeros(3, dtype=np.int32)
    assert L.pp_site_categories(None, 3, cats.ctypes.data_as(C.POINTER(C.c_int))) != 0
