"""GPU parity of the Price-CAT model (gcat_model.ml) against the oracle,
through the C ABI: rate-free GLVs where each site evolves by the transition
matrix of its category's rate.

Categories are the donovan refpkg's real ones (20 FastTree rates, the
per-column categories of its first columns) or seeded random draws.  The bar
is the gamma path's: GLVs and exponents as the oracle's (exponents equal,
values within 1e-12 relative), stage-1 scores within 1e-12 relative in FP64,
placements with the same best edge (or a near-tie), likelihood within 1e-9
relative in FP64 and 1e-4 absolute in FP32.
"""
import numpy as np
import pytest

from tests.helpers import ar53_price_cat, oracle_setup_cat
from tests.test_gpu_parity import _compare_placements

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _gpu_setup(wl, o, cat_rates):
    from pplacer_amd import placer
    pl = placer.Placer(0)
    pl.build_model_cat(wl.exch, o["statd"], cat_rates, o["site_cat"])
    pl.set_tree(wl.tree, o["leaf"], o["N"])
    return pl


def _case(oracle, name, seed, n_queries, how):
    from pplacer_amd import synth
    wl = synth.make_workload(name, seed=seed, n_queries=n_queries)
    N0 = len(wl.ref_aln[0][1])
    if how == "ar53":
        cat_rates, cats = ar53_price_cat()
        site_cat = cats[:N0]
    else:
        cat_rates = [0.05, 0.3, 0.8, 1.4, 2.9]
        site_cat = np.random.default_rng(seed).integers(0, len(cat_rates), size=N0)
    o = oracle_setup_cat(oracle, wl, cat_rates, site_cat)
    pl = _gpu_setup(wl, o, cat_rates)
    return dict(wl=wl, o=o, pl=pl)


@pytest.fixture(scope="module")
def gtr50cat(oracle):
    _print_case("workload gtr50 seed 6, 24 queries, random categories")
    c = _case(oracle, "gtr50", 6, 24, "random")
    yield c
    c["pl"].close()


@pytest.fixture(scope="module")
def prot10cat(oracle):
    _print_case("workload prot10 seed 8, 12 queries, ar53 categories")
    c = _case(oracle, "prot10", 8, 12, "ar53")
    yield c
    c["pl"].close()


def _check_glvs(c):
    pl, pre = c["pl"], c["o"]["pre"]
    d, de, p, pe, s, se = pl.download_glvs()
    np.testing.assert_array_equal(de, pre.darr_e)
    np.testing.assert_array_equal(pe, pre.parr_e)
    np.testing.assert_allclose(d, pre.darr, rtol=1e-12, atol=0)
    np.testing.assert_allclose(p, pre.parr, rtol=1e-12, atol=0)
    E = c["wl"].tree.n_nodes - 1
    np.testing.assert_array_equal(se[:E], pre.snodes_e[:E])
    np.testing.assert_allclose(s[:E], pre.snodes[:E], rtol=1e-12, atol=0)


def _check_stage1(c):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    pl, core, queries = c["pl"], c["o"]["core"], c["o"]["queries"]
    s64 = pl.stage1_scores(queries, 0.1, PP_FP64)
    s32 = pl.stage1_scores(queries, 0.1, PP_FP32)
    for q in range(len(queries)):
        want = core.stage1_scores(queries[q])
        np.testing.assert_allclose(s64[q], want, rtol=1e-12, atol=0)
        assert np.max(np.abs(s32[q] - want)) < 2e-3
        np.testing.assert_array_equal(core.rank(s64[q]), core.rank(want))


def test_cat_nucleotide_glvs(gtr50cat):
    _print_case("gtr50 cat glvs")
    _check_glvs(gtr50cat)


def test_cat_nucleotide_stage1(gtr50cat):
    _print_case("gtr50 cat stage1")
    _check_stage1(gtr50cat)


def test_cat_nucleotide_place_fp64(gtr50cat):
    from pplacer_amd.placer import PP_FP64
    _print_case("gtr50 cat place fp64")
    core, queries = gtr50cat["o"]["core"], gtr50cat["o"]["queries"]
    got, status, stats = gtr50cat["pl"].place(queries, precision=PP_FP64)
    assert _compare_placements(core, queries, got, status, 1e-9, True) == len(queries)


def test_cat_nucleotide_place_fp32(gtr50cat):
    from pplacer_amd.placer import PP_FP32
    _print_case("gtr50 cat place fp32")
    core, queries = gtr50cat["o"]["core"], gtr50cat["o"]["queries"]
    got, status, stats = gtr50cat["pl"].place(queries, precision=PP_FP32)
    ties = [0]
    assert _compare_placements(core, queries, got, status, 1e-4, False, ties) + ties[0] == len(queries)


def test_cat_nucleotide_optimize(gtr50cat):
    from pplacer_amd.placer import PP_FP64
    _print_case("gtr50 cat three_tax optimize")
    c = gtr50cat
    pl, core, queries = c["pl"], c["o"]["core"], c["o"]["queries"]
    tree = c["wl"].tree
    E = tree.n_nodes - 1
    jobs = [(q, loc) for q in range(4) for loc in (0, E // 3, E - 1)]
    qi, locs = [j[0] for j in jobs], [j[1] for j in jobs]
    pend0 = [0.1] * len(jobs)
    dist0 = [float(tree.bl[loc]) / 2 for loc in locs]
    ll, pend, dist, calls, st = pl.optimize(queries, qi, locs, pend0, dist0, 0.01, 2.0, PP_FP64)
    for j, (q, loc) in enumerate(jobs):
        ost, oll, opend, odist, ocalls = core.optimize(queries[q], loc, pend0[j], dist0[j], 0.01)
        assert st[j] == ost == 0
        assert abs(ll[j] - oll) <= 1e-9 * abs(oll) + 1e-9, (j, ll[j], oll)
        assert abs(pend[j] - opend) <= 2 * 0.01 and abs(dist[j] - odist) <= 2 * 0.01


def test_cat_protein_glvs(prot10cat):
    _print_case("prot10 cat glvs")
    _check_glvs(prot10cat)


def test_cat_protein_stage1(prot10cat):
    _print_case("prot10 cat stage1")
    _check_stage1(prot10cat)


def test_cat_protein_place(prot10cat):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("prot10 cat place fp64/fp32")
    core, queries = prot10cat["o"]["core"], prot10cat["o"]["queries"]
    got, status, stats = prot10cat["pl"].place(queries, precision=PP_FP64)
    assert _compare_placements(core, queries, got, status, 1e-9, True) == len(queries)
    got32, status32, _ = prot10cat["pl"].place(queries, precision=PP_FP32)
    ties = [0]
    assert _compare_placements(core, queries, got32, status32, 1e-4, False, ties) + ties[0] == len(queries)


def test_cat_single_category_equals_gamma(oracle):
    """All sites in one category of rate r: the CAT path gives the gamma
    path's GLVs and placements with rates = [r]."""
    from pplacer_amd import placer, synth
    from pplacer_amd.placer import PP_FP64
    _print_case("gtr12 cat one category vs gamma")
    wl = synth.make_workload("gtr12", seed=9, n_queries=6)
    N0 = len(wl.ref_aln[0][1])
    o = oracle_setup_cat(oracle, wl, [0.1, 0.7, 1.9], np.full(N0, 1))
    cat = _gpu_setup(wl, o, [0.1, 0.7, 1.9])
    gam = placer.Placer(0)
    try:
        gam.build_model(wl.exch, o["statd"], [0.7])
        gam.set_tree(wl.tree, o["leaf"], o["N"])
        a, b = cat.download_glvs(), gam.download_glvs()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        r1, _, _ = cat.place(o["queries"], precision=PP_FP64)
        r2, _, _ = gam.place(o["queries"], precision=PP_FP64)
        for g1, g2 in zip(r1, r2):
            assert [p.edge_num for p in g1] == [p.edge_num for p in g2]
            for p1, p2 in zip(g1, g2):
                assert p1.likelihood == p2.likelihood
    finally:
        cat.close()
        gam.close()
