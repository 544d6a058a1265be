"""GPU parity of the HIP placement path against the oracle (FP64 restatement
of pplacer), through the C ABI.

Tolerances: FP64 form, log-likelihoods within 1e-9 relative (stage-1 scores
and twins much tighter, 1e-12); FP32 form, stage-1 scores within 2e-3
absolute and final placement likelihoods (FP64 refinement) within 1e-4
absolute.  Best edges must agree exactly; branch lengths within the Brent
tolerance they were optimised to.
"""
import math

import numpy as np
import pytest

from tests.helpers import oracle_setup

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


@pytest.fixture(scope="module")
def gtr50(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload gtr50 seed 3, 40 queries")
    wl = synth.make_workload("gtr50", seed=3, n_queries=40)
    model, pre, core, oq = oracle_setup(oracle, wl)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    assert queries == oq
    return dict(wl=wl, model=model, pre=pre, core=core, pl=pl, queries=queries)


# ---- linear_c.c twins --------------------------------------------------------

def test_twins_match_oracle(oracle):
    from pplacer_amd import linear
    _print_case("linear_c twins")
    rng = np.random.default_rng(1)
    for S in (4, 20):
        R, N = 4, 37
        x, y, z = (rng.random((R, N, S)) for _ in range(3))
        statd = rng.dirichlet(np.ones(S))
        mask = (rng.random(N) < 0.6).astype(np.uint16)
        util = np.zeros(N)
        got = linear.ten_masked_logdot(x, y, mask, util)
        assert abs(got - oracle.ten_masked_logdot(x, y, mask)) <= 1e-12 * abs(got)
        assert np.all(util[mask == 0] == 0)
        got = linear.ten_log_like3(statd, x, y, z)
        assert abs(got - oracle.ten_log_like3(statd, x, y, z)) <= 1e-12 * abs(got)
        got = linear.ten_bounded_logdot(x, y, 3, 20)
        assert abs(got - oracle.ten_bounded_logdot(x, y, 3, 20)) <= 1e-12 * abs(got)
        np.testing.assert_array_equal(linear.ten_pairwise_prod(None, x, y), oracle.ten_pairwise_prod(x, y))
        np.testing.assert_array_equal(linear.ten_statd_pairwise_prod(statd, None, x, y),
                                      oracle.ten_statd_pairwise_prod(statd, x, y))
        a, b = rng.random((S, S)), rng.random((N, S))
        np.testing.assert_array_equal(linear.gemmish(None, a, b), oracle.gemmish(a, b))
        u, lam, uit = rng.random((S, S)), rng.random(S), rng.random((S, S))
        np.testing.assert_array_equal(linear.dediagonalize(None, u, lam, uit),
                                      oracle.dediagonalize(u, lam, uit))
    # the mat_* stubs (Price-CAT's rate-free GLVs), linear_c.c:148-340
    for S in (4, 20):
        N = 41
        x, y, z = (rng.random((N, S)) for _ in range(3))
        statd = rng.dirichlet(np.ones(S))
        mask = (rng.random(N) < 0.6).astype(np.uint16)
        assert linear.mat_log_like3(statd, x, y, z) == oracle.mat_log_like3(statd, x, y, z)
        assert linear.mat_masked_logdot(x, y, mask) == oracle.mat_masked_logdot(x, y, mask)
        assert linear.mat_bounded_logdot(x, y, 5, 30) == oracle.mat_bounded_logdot(x, y, 5, 30)
        np.testing.assert_array_equal(linear.mat_pairwise_prod(None, x, y), x * y)
        np.testing.assert_array_equal(linear.mat_statd_pairwise_prod(statd, None, x, y),
                                      oracle.ten_statd_pairwise_prod(statd, x[None], y[None])[0])
    # tests/pplacer/test_linear.ml:4-12 closed form: log 3 + log 4
    m1 = np.array([[-1.0, 0.0, 1.0], [1000.0, 1.0, 3.0], [0.0, 0.0, 2.0]])
    m2 = np.array([[-1.0, 1.0, 2.0], [1.0, 3.0, 0.0], [0.0, 1.0, 2.0]])
    assert abs(linear.mat_masked_logdot(m1, m2, np.array([1, 0, 1], np.uint16))
               - (math.log(3) + math.log(4))) < 1e-12
    # tests/pplacer/test_linear.ml:14-28 closed form through the GPU twin
    t1 = np.array([[[-1.0, 0.0, 1.0], [1000.0, 1.0, 3.0], [0.0, 0.0, 2.0]]] * 2)
    t2 = np.array([[[-1.0, 1.0, 0.0], [1.0, 3.0, 0.0], [0.0, 1.0, 2.0]],
                   [[-1.0, 1.0, 2.0], [1.0, 3.0, 0.0], [0.0, 1.0, 3.0]]])
    want = (math.log(4) + math.log(10)) - 2 * math.log(2)
    assert abs(linear.ten_masked_logdot(t1, t2, np.array([1, 0, 1], np.uint16)) - want) < 1e-12


# ---- model and precompute -----------------------------------------------------

def test_model_and_tree_glvs_match_oracle(gtr50):
    _print_case("model + tree GLVs")
    pl, model, pre = gtr50["pl"], gtr50["model"], gtr50["pre"]
    x, l, xit = pl.model_diagd()
    np.testing.assert_allclose(x, model.x, rtol=0, atol=1e-14)
    np.testing.assert_allclose(l, model.l, rtol=1e-14)
    d, de, p, pe, s, se = pl.download_glvs()
    np.testing.assert_array_equal(de, pre.darr_e)
    np.testing.assert_array_equal(pe, pre.parr_e)
    np.testing.assert_allclose(d, pre.darr, rtol=1e-12, atol=0)
    np.testing.assert_allclose(p, pre.parr, rtol=1e-12, atol=0)
    E = gtr50["wl"].tree.n_nodes - 1
    np.testing.assert_array_equal(se[:E], pre.snodes_e[:E])
    np.testing.assert_allclose(s[:E], pre.snodes[:E], rtol=1e-12, atol=0)


# ---- stage 1 -------------------------------------------------------------------

def test_stage1_scores_fp64_and_fp32(gtr50):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("stage1 fp64/fp32")
    pl, core, queries = gtr50["pl"], gtr50["core"], gtr50["queries"]
    s64 = pl.stage1_scores(queries, 0.1, PP_FP64)
    s32 = pl.stage1_scores(queries, 0.1, PP_FP32)
    for q in range(len(queries)):
        want = core.stage1_scores(queries[q])
        np.testing.assert_allclose(s64[q], want, rtol=1e-12, atol=0)
        assert np.max(np.abs(s32[q] - want)) < 2e-3
        # ranking identical in FP64
        np.testing.assert_array_equal(core.rank(s64[q]), core.rank(want))


def test_rank_topk_matches_stable_sort(gtr50, oracle):
    _print_case("rank topk")
    pl = gtr50["pl"]
    rng = np.random.default_rng(5)
    s = rng.standard_normal((7, 300))
    s[0, 10:20] = 1.5          # ties keep ascending edge order
    s[1, 5] = np.nan           # nan ranks last
    s[2, :] = 0.0
    s[2, 3] = -0.0
    got = pl.rank_topk(s, 45)
    for q in range(7):
        np.testing.assert_array_equal(got[q], oracle.Core.rank(s[q])[:45])


# ---- stage 2: Three_tax.optimize -------------------------------------------------

def test_three_tax_optimize_matches_oracle(gtr50):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("three_tax optimize")
    pl, core, queries, wl = gtr50["pl"], gtr50["core"], gtr50["queries"], gtr50["wl"]
    jobs = []
    for q in range(8):
        order = core.rank(core.stage1_scores(queries[q]))
        for loc in order[:3]:
            jobs.append((q, int(loc)))
    qj = [j[0] for j in jobs]
    locs = [j[1] for j in jobs]
    bl = wl.tree.bl
    for tol in (0.01, 1e-5):
        p0 = [0.1] * len(jobs)
        d0 = [bl[l] / 2 for l in locs]
        ll, pe, di, calls, st = pl.optimize(queries, qj, locs, p0, d0, tol, 2.0, PP_FP64)
        assert np.all(st == 0)
        n_exact = 0
        for k, (q, loc) in enumerate(jobs):
            ost, oll, ope, odi, ocalls = core.optimize(queries[q], loc, 0.1, bl[loc] / 2, tol)
            assert ost == 0
            assert abs(ll[k] - oll) <= 1e-9 * abs(oll), (k, ll[k], oll)
            assert abs(pe[k] - ope) <= 2 * tol and abs(di[k] - odi) <= 2 * tol
            exact = ll[k] == oll and pe[k] == ope and di[k] == odi
            n_exact += exact
            # Three_tax.n_like_calls: where the result is the oracle's to the
            # last bit, Brent took the oracle's path, so the count of
            # log_like3 calls (the GSL set's three re-evaluations included)
            # is the oracle's
            if exact:
                assert calls[k] == ocalls, (k, calls[k], ocalls)
        print("tol %g: %d of %d jobs bit-equal to the oracle" % (tol, n_exact, len(jobs)))
        assert n_exact >= 0.9 * len(jobs)
        ll32, pe32, di32, _, st32 = pl.optimize(queries, qj, locs, p0, d0, tol, 2.0, PP_FP32)
        assert np.all(st32 == 0)
        # FP32 probes change Brent's path: most jobs agree to FP32 rounding,
        # a few settle in a neighbouring local optimum of the likelihood
        # surface (the FP32 form refines its kept placements in FP64)
        d = np.abs(ll32 - ll)
        assert np.median(d) < 1e-4 and np.max(d) < 0.5
        assert np.mean(d < 1e-3) >= 0.9


# ---- full placement --------------------------------------------------------------

def _compare_placements(core, queries, got, status, ll_atol, fp64, ties=None):
    """Number of queries whose best placement is the oracle's edge, with the
    likelihood within ll_atol (+1e-9 relative in FP64) and branch lengths
    within twice final_tolerance (core.ml:14) in FP64.  A different best edge
    is accepted only as a tie in the reference's own output: the oracle keeps
    that edge too, with a likelihood within 1e-6 of its best (which of two
    such edges comes first is decided by rounding); such queries are counted
    in ties[0] when a list is given, and fail the test otherwise."""
    n_same_best = 0
    bl_tol = 2e-5 if fp64 else 2e-4
    for q, query in enumerate(queries):
        st, want, _ = core.place(query)
        assert st == 0 and status[q] == 0
        g = got[q]
        assert len(g) > 0
        if g[0].edge_num != want[0][0]:
            wmap = {p[0]: p for p in want}
            tie = (ties is not None and g[0].edge_num in wmap
                   and abs(wmap[g[0].edge_num][2] - want[0][2]) < 1e-6
                   and abs(g[0].likelihood - want[0][2]) < 1e-6)
            assert tie, (q, g[0], want[:2])
            ties[0] += 1
            continue
        n_same_best += 1
        assert abs(g[0].likelihood - want[0][2]) <= ll_atol + (1e-9 * abs(want[0][2]) if fp64 else 0)
        assert abs(g[0].pendant_length - want[0][4]) < bl_tol, (q, g[0], want[0])
        assert abs(g[0].distal_length - want[0][3]) < bl_tol, (q, g[0], want[0])
        assert abs(sum(p.like_weight_ratio for p in g) - 1.0) < 1e-9
        wmap = {p[0]: p for p in want}
        for p in g:
            if p.edge_num in wmap and p.like_weight_ratio > 1e-3:
                assert abs(p.like_weight_ratio - wmap[p.edge_num][1]) < 1e-3
    return n_same_best


def test_place_fp64_matches_oracle(gtr50):
    from pplacer_amd.placer import PP_FP64
    _print_case("place fp64")
    pl, core, queries = gtr50["pl"], gtr50["core"], gtr50["queries"]
    got, status, stats = pl.place(queries, precision=PP_FP64)
    n = _compare_placements(core, queries, got, status, 1e-9, True)
    assert n == len(queries)
    assert stats.n_pitched >= len(queries)


def test_place_fp32_matches_oracle(gtr50):
    from pplacer_amd.placer import PP_FP32
    _print_case("place fp32")
    pl, core, queries = gtr50["pl"], gtr50["core"], gtr50["queries"]
    got, status, stats = pl.place(queries, precision=PP_FP32)
    ties = [0]
    n = _compare_placements(core, queries, got, status, 1e-4, False, ties)
    print("fp32: %d same best edge, %d ties of the reference's own output" % (n, ties[0]))
    assert n + ties[0] == len(queries)


def test_place_edge_cases(gtr50):
    from pplacer_amd.placer import PP_FP64, default_prefs
    _print_case("edge cases")
    pl, core, queries = gtr50["pl"], gtr50["core"], gtr50["queries"]
    N = len(queries[0])
    bad = ["-" * N, "A" * (N - 1) + "!", queries[0]]
    got, status, _ = pl.place(bad, precision=PP_FP64)
    assert list(status) == [1, 2, 0]
    assert got[0] == [] and got[1] == []
    # max_strikes = 0 evaluates every edge
    prefs = default_prefs(max_strikes=0)
    got, status, stats = pl.place(queries[:3], prefs=prefs, precision=PP_FP64)
    assert stats.n_pitched == 3 * (gtr50["wl"].tree.n_nodes - 1)
    assert np.all(status == 0)
    # keep_at_most = 1 returns the single best placement with ratio 1
    got, status, _ = pl.place(queries[:4], prefs=default_prefs(keep_at_most=1))
    assert all(len(g) == 1 and g[0].like_weight_ratio == 1.0 for g in got)


# ---- end to end on the reference's own test data --------------------------------

def test_run_placements_on_reference_fixture(tmp_path, oracle):
    """Queries cut from reference sequences of tests/data/like/gtr (HIV, 20
    taxa x 216 columns, PhyML GTR+G4) place on their own leaf edge, agree
    with the oracle, and are written as jplace."""
    import json as _json

    from pplacer_amd import jplace, refio, run
    from tests.helpers import GOLDEN, cut
    _print_case("run_placements on like/gtr fixture")
    d = GOLDEN + "/like/gtr/"
    tree = refio.read_newick(d + "hiv.phy_phyml_tree.txt")
    aln = refio.read_fasta(d + "hiv.fasta")
    pm = refio.read_phylo_model(d + "phylo_model.jplace")
    leaves = tree.leaves()
    queries = []
    for k, i in enumerate(leaves[:8]):
        seq = dict(aln)[tree.names[i]]
        a, b = 20 + 7 * k, 150 + 5 * k
        queries.append(("q_%s" % tree.names[i], "-" * a + seq[a:b] + "-" * (len(seq) - b)))
    # a duplicate query is placed once and reported under both names
    queries.append(("dup", queries[0][1]))
    pqueries, status = run.run_placements(tree, aln, queries, pm)
    assert np.all(status == 0)
    assert len(pqueries) == 8
    names = {nl[0][0]: (nl, recs) for nl, recs in pqueries}
    assert [n for n, _ in names["q_" + tree.names[leaves[0]]][0]] == ["q_" + tree.names[leaves[0]], "dup"]
    # oracle on the same inputs
    ref_seqs = [s for _, s in aln]
    model = oracle.Model(refio.gtr_exchangeabilities(pm.subs_rates), oracle.emper_freq(4, ref_seqs),
                         oracle.discrete_gamma(pm.n_cats, pm.alpha))
    st, mask, keep, n = oracle.premask(4, ref_seqs, [s for _, s in queries])
    leaf = {i: cut(dict(aln)[tree.names[i]], mask) for i in leaves}
    pre = oracle.Precompute(model, tree, leaf, n)
    pre.pull_and_supernodes()
    core = oracle.Core(pre)
    for k, i in enumerate(leaves[:8]):
        nl, recs = names["q_" + tree.names[i]]
        ost, want, _ = core.place(cut(queries[k][1], mask))
        assert recs[0].edge_num == want[0][0]
        assert abs(recs[0].likelihood - want[0][2]) <= 1e-9 * abs(want[0][2])
        # a fragment of a leaf's own sequence lands on that leaf's edge with a
        # short pendant
        assert recs[0].edge_num == i or recs[0].pendant_length < 0.02
    out = tmp_path / "out.jplace"
    jplace.write(str(out), tree, pqueries)
    obj = _json.loads(out.read_text())
    assert len(obj["placements"]) == 8 and obj["fields"] == jplace.FIELDS


# ---- the 16S-scale shape (2,000 taxa, 1,600 columns) -----------------------------

@pytest.fixture(scope="module")
def gtr2000(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload gtr2000 seed 1, 24 queries")
    wl = synth.make_workload("gtr2000", seed=1, n_queries=24)
    model, pre, core, oq = oracle_setup(oracle, wl)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    assert queries == oq
    return dict(wl=wl, model=model, pre=pre, core=core, pl=pl, queries=queries)


def test_full_size_glvs_and_stage1(gtr2000):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("gtr2000 glvs + stage1")
    pl, pre, core, queries = gtr2000["pl"], gtr2000["pre"], gtr2000["core"], gtr2000["queries"]
    d, de, p, pe, s, se = pl.download_glvs()
    np.testing.assert_array_equal(de, pre.darr_e)
    np.testing.assert_array_equal(pe, pre.parr_e)
    assert np.array_equal(d, pre.darr) and np.array_equal(p, pre.parr)
    s64 = pl.stage1_scores(queries, 0.1, PP_FP64)
    s32 = pl.stage1_scores(queries, 0.1, PP_FP32)
    assert np.all(np.isfinite(s64)) and np.all(np.isfinite(s32))
    for q in (0, 7):
        np.testing.assert_allclose(s64[q], core.stage1_scores(queries[q]), rtol=1e-12, atol=0)
    # FP32 error relative to the score's magnitude (M sites of logs)
    assert np.max(np.abs(s32 - s64) / np.abs(s64)) < 1e-6
    # identical top of the ranking for most queries
    top64 = np.argmax(s64, axis=1)
    top32 = np.argmax(s32, axis=1)
    assert np.mean(top64 == top32) >= 0.9


def test_full_size_placements(gtr2000):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("gtr2000 placements")
    pl, core, queries = gtr2000["pl"], gtr2000["core"], gtr2000["queries"]
    sub = queries[:8]
    got, status, stats = pl.place(sub, precision=PP_FP64)
    n = _compare_placements(core, sub, got, status, 1e-9, True)
    assert n == len(sub)
    got32, status32, stats32 = pl.place(sub, precision=PP_FP32)
    assert np.all(status32 == 0)
    ties = [0]
    n32 = _compare_placements(core, sub, got32, status32, 1e-4, False, ties)
    assert n32 + ties[0] == len(sub)
    # the FP32 pitches do not inflate the work (no spurious non-finite probes)
    assert stats32.n_pitched <= 1.5 * stats.n_pitched
