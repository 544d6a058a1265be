"""Pins the oracle's Price-CAT model (gcat_model.ml): per-site transition
matrices by rate category, rate-free GLVs.

The reference ships no CAT likelihood fixture, so the pins are (1) the
reduction to the gamma model -- every site in one category of rate r is the
Gamma model with rates = [r], bit for bit -- and (2) brute-force Felsenstein
pruning with each site at its own category's rate (tests/bruteforce.py,
scipy expm), at every edge and at three-taxon placements.
"""
import numpy as np
import pytest

from tests import bruteforce as bf

CAT_RATES = [0.05, 0.3, 0.8, 1.4, 2.9]


def _setup(oracle, name, seed, site_cat_fn):
    from pplacer_amd import synth
    wl = synth.make_workload(name, seed=seed)
    S = wl.n_states
    ref_seqs = [s for _, s in wl.ref_aln]
    statd = oracle.emper_freq(S, ref_seqs)
    leaf = wl.leaf_seqs()
    N = len(ref_seqs[0])
    site_cat = site_cat_fn(N)
    model = oracle.Model(wl.exch, statd, cat_rates=CAT_RATES, site_cat=site_cat)
    pre = oracle.Precompute(model, wl.tree, leaf, N)
    pre.pull_and_supernodes()
    return wl, statd, leaf, N, site_cat, model, pre


def test_cat_single_category_equals_gamma(oracle):
    wl, statd, leaf, N, site_cat, cat, pre_c = _setup(oracle, "gtr12", 5, lambda n: np.full(n, 2))
    gam = oracle.Model(wl.exch, statd, [CAT_RATES[2]])
    pre_g = oracle.Precompute(gam, wl.tree, leaf, N)
    pre_g.pull_and_supernodes()
    assert np.array_equal(pre_c.darr, pre_g.darr) and np.array_equal(pre_c.parr, pre_g.parr)
    assert np.array_equal(pre_c.snodes, pre_g.snodes)
    assert np.array_equal(pre_c.darr_e, pre_g.darr_e)
    core_c, core_g = oracle.Core(pre_c), oracle.Core(pre_g)
    q = [s for _, s in wl.queries][0]
    assert np.array_equal(core_c.stage1_scores(q), core_g.stage1_scores(q))
    assert core_c.place(q)[1] == core_g.place(q)[1]


@pytest.mark.parametrize("name,seed", [("gtr12", 7), ("prot10", 3)])
def test_cat_tree_likelihood_bruteforce(oracle, name, seed):
    rng = np.random.default_rng(seed)
    wl, statd, leaf, N, site_cat, model, pre = _setup(
        oracle, name, seed, lambda n: rng.integers(0, len(CAT_RATES), size=n))
    want = bf.tree_loglike_cat(wl.tree, leaf, wl.exch, statd, CAT_RATES, site_cat)
    one = (np.ones((1, N, wl.n_states)), np.zeros(N, dtype=np.int64))
    for i in range(wl.tree.n_nodes - 1):
        got = model.logdot((pre.snodes[i], pre.snodes_e[i]), one)
        assert abs(got - want) < 1e-9 * abs(want), (i, got, want)


def test_cat_three_taxon_bruteforce(oracle):
    rng = np.random.default_rng(11)
    wl, statd, leaf, N, site_cat, model, pre = _setup(
        oracle, "gtr12", 11, lambda n: rng.integers(0, len(CAT_RATES), size=n))
    core = oracle.Core(pre)
    queries = [s for _, s in wl.queries]
    E = wl.tree.n_nodes - 1
    for q in queries[:3]:
        for loc in (0, E // 2, E - 1):
            st, ll, pend, dist, calls = core.optimize(q, loc, 0.1, wl.tree.bl[loc] / 2, 0.01)
            assert st == 0
            want = bf.placement_loglike_cat(wl.tree, leaf, wl.exch, statd, CAT_RATES, site_cat,
                                            loc, dist, pend, q)
            assert abs(ll - want) < 1e-9 * abs(want), (loc, ll, want)
