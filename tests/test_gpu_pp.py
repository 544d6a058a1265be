"""GPU parity of -p (posterior probabilities) against the oracle, through the
C ABI: Three_tax.calc_marg_prob per job (pp_marginal_batch) and full
placements with calc_pp (pp_place_batch).

The GPU runs the oracle's sequence of probes: the same find_upper_limit
Brent, the same QNG node order and sums, the same halving on tolerance
failures.  Probe likelihoods differ from the oracle's by summation order
(~1e-15 relative), so marginal log-likelihoods from the same branch lengths
are held to 1e-7 absolute (a different QNG level or Brent path would show as
~1e-4 or more); through full placements, see _check_place.
"""
import numpy as np
import pytest

from tests.helpers import oracle_setup, oracle_setup_cat

pytestmark = pytest.mark.gpu

TOL = 1e-7


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _jobs(core, queries, n_q, per_q):
    jobs = []
    for qi in range(n_q):
        st, pl, _ = core.place(queries[qi])
        assert st == 0
        for p in pl[:per_q]:
            jobs.append((qi, p[0], p[4], p[3]))  # query, loc, pend, dist
    return jobs


def _check_marginals(pl, core, queries, jobs, rel_err=0.01):
    qj = [j[0] for j in jobs]
    locs = [j[1] for j in jobs]
    pend = [j[2] for j in jobs]
    dist = [j[3] for j in jobs]
    m, calls, st = pl.marginal(queries, qj, locs, pend, dist, rel_err, 2.0)
    for k, (q, loc, pe, di) in enumerate(jobs):
        ost, om, ocalls = core.marg_prob(queries[q], loc, pe, di)
        assert st[k] == ost == 0, (k, st[k], ost)
        assert abs(m[k] - om) < TOL, (k, m[k], om)
        # the GPU reuses the three values GSL's fminimizer_set re-evaluates
        assert calls[k] in (ocalls, ocalls - 3), (k, calls[k], ocalls)


def _check_place(core, queries, got, status, best_rel=1e-9):
    """Full -p placements: the same kept edges in the same order; the best
    likelihood within 1e-9 relative as without -p; the other kept placements
    (-p keeps up to 20, down to 1e-3 of the best) end their final Brent
    within its tolerance of the oracle's, so their likelihoods, marginals and
    posteriors are held to 1e-6 absolute."""
    for q, query in enumerate(queries):
        st, want, _ = core.place(query)
        assert st == 0 and status[q] == 0
        g = got[q]
        assert [r.edge_num for r in g] == [w[0] for w in want]
        assert abs(g[0].likelihood - want[0][2]) <= best_rel * abs(want[0][2])
        for r, w in zip(g, want):
            assert abs(r.likelihood - w[2]) < 1e-6, (q, r, w)
            assert abs(r.marginal_like - w[6]) < 1e-6, (q, r, w)
            assert abs(r.post_prob - w[5]) < 1e-6, (q, r, w)
        assert abs(sum(r.post_prob for r in g) - 1.0) < 1e-12


@pytest.fixture(scope="module")
def gtr50pp(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload gtr50 seed 1, 10 queries, -p")
    wl = synth.make_workload("gtr50", seed=1, n_queries=10)
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    assert q2 == queries
    yield dict(wl=wl, core=core, pl=pl, queries=queries)
    pl.close()


def test_marginal_batch_matches_oracle(gtr50pp):
    _print_case("gtr50 marginal batch")
    c = gtr50pp
    _check_marginals(c["pl"], c["core"], c["queries"], _jobs(c["core"], c["queries"], 6, 3))


def test_place_calc_pp_matches_oracle(gtr50pp):
    from pplacer_amd.placer import PP_FP64, default_prefs
    _print_case("gtr50 place -p fp64")
    c = gtr50pp
    prefs = default_prefs(calc_pp=True)
    got, status, stats = c["pl"].place(c["queries"], prefs=prefs, precision=PP_FP64)
    _check_place(c["core"], c["queries"], got, status)
    assert stats.t_pp_ms > 0


def test_place_calc_pp_fp32_form(gtr50pp):
    """The FP32 form's marginals are FP64 from its (FP64-refined) kept
    placements: where the kept set matches the oracle's, so do they."""
    from pplacer_amd.placer import PP_FP32, default_prefs
    _print_case("gtr50 place -p fp32")
    c = gtr50pp
    got, status, _ = c["pl"].place(c["queries"], prefs=default_prefs(calc_pp=True),
                                   precision=PP_FP32)
    n_same = 0
    for q, query in enumerate(c["queries"]):
        st, want, _ = c["core"].place(query)
        assert status[q] == 0
        if [r.edge_num for r in got[q]] != [w[0] for w in want]:
            continue
        n_same += 1
        for r, w in zip(got[q], want):
            assert abs(r.marginal_like - w[6]) < 1e-4 and abs(r.post_prob - w[5]) < 1e-4
    assert n_same >= len(c["queries"]) - 1


@pytest.mark.parametrize("prior", ["uniform", "informative"])
def test_priors_match_oracle(oracle, gtr50pp, prior):
    _print_case("gtr50 marginal batch, %s prior" % prior)
    c = gtr50pp
    model, pre, core, queries = oracle_setup(oracle, c["wl"], calc_pp=True, prior=prior,
                                             prior_lower=0.02)
    c["pl"].set_prior(prior, 0.02)
    try:
        _check_marginals(c["pl"], core, queries, _jobs(core, queries, 3, 2))
    finally:
        c["pl"].set_prior("flat")


def test_protein_and_cat_marginals(oracle):
    from pplacer_amd import placer, synth
    from pplacer_amd.placer import PP_FP64, default_prefs
    _print_case("prot10 seed 2 marginal batch + place -p")
    wl = synth.make_workload("prot10", seed=2, n_queries=4)
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    try:
        _check_marginals(pl, core, queries, _jobs(core, queries, 3, 2))
        got, status, _ = pl.place(queries, prefs=default_prefs(calc_pp=True), precision=PP_FP64)
        # 20-state stage 2 evaluates in the eigenbasis: probes differ from the
        # oracle's P(t) form at ~1e-13, enough to resolve a near-tie in the final
        # Brent differently, which then stops elsewhere within its 1e-5 tolerance
        _check_place(core, queries, got, status, best_rel=1e-8)
    finally:
        pl.close()
    _print_case("gtr12 seed 4 Price-CAT marginal batch")
    wl = synth.make_workload("gtr12", seed=4, n_queries=4)
    N0 = len(wl.ref_aln[0][1])
    cats = [0.1, 0.6, 1.7]
    site_cat = np.random.default_rng(4).integers(0, 3, size=N0)
    o = oracle_setup_cat(oracle, wl, cats, site_cat, calc_pp=True)
    pl = placer.Placer(0)
    try:
        pl.build_model_cat(wl.exch, o["statd"], cats, o["site_cat"])
        pl.set_tree(wl.tree, o["leaf"], o["N"])
        _check_marginals(pl, o["core"], o["queries"], _jobs(o["core"], o["queries"], 3, 2))
    finally:
        pl.close()


def test_short_edge_and_tolerance_error(oracle):
    """cut_bl =~ 0 integrates the pendant length only; a tolerance below
    50 eps is GSL's uncaught EBADTOL on both sides."""
    from pplacer_amd import placer, synth
    _print_case("gtr12 seed 3 short edge / EBADTOL")
    wl = synth.make_workload("gtr12", seed=3, n_queries=2)
    wl.tree.bl[4] = 4e-6
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    try:
        _check_marginals(pl, core, queries, [(0, 4, 0.05, 2e-6), (1, 4, 0.2, 1e-6)])
        m, calls, st = pl.marginal(queries, [0], [3], [0.05], [wl.tree.bl[3] / 2], 1e-15, 2.0)
        _, _, core_bad, _ = oracle_setup(oracle, wl, calc_pp=True, pp_rel_err=1e-15)
        ost, om, oc = core_bad.marg_prob(queries[0], 3, 0.05, wl.tree.bl[3] / 2)
        assert st[0] != 0 and ost != 0
        assert calls[0] in (oc, oc - 3)  # both stop after find_upper_limit
    finally:
        pl.close()


def test_run_cli_writes_posterior_fields(tmp_path):
    """pplacer_amd.run -p on the reference's HIV fixture writes post_prob and
    marginal_like for every placement, summing to 1 per query."""
    import json
    import os
    from pplacer_amd import jplace, refio, run
    from tests.helpers import GOLDEN
    _print_case("run -p on the like/gtr fixture")
    d = os.path.join(GOLDEN, "like", "gtr")
    tree = refio.read_newick(os.path.join(d, "hiv.phy_phyml_tree.txt"))
    aln = refio.read_fasta(os.path.join(d, "hiv.fasta"))
    seqs = dict(aln)
    queries = []
    for k, i in enumerate(tree.leaves()[:3]):  # fragments of reference sequences
        s = seqs[tree.names[i]]
        a, b = 30 + 11 * k, 170 + 3 * k
        queries.append(("q%d" % k, "-" * a + s[a:b] + "-" * (len(s) - b)))
    fa = tmp_path / "all.fasta"
    refio.write_fasta(str(fa), list(aln) + queries)
    out = tmp_path / "out.jplace"
    rc = run.main(["--tree", os.path.join(d, "hiv.phy_phyml_tree.txt"), "--aln", str(fa),
                   "--model", os.path.join(d, "phylo_model.jplace"), "-o", str(out), "-p"])
    assert rc == 0
    obj = json.load(open(out))
    assert "post_prob" in obj["fields"] and "marginal_like" in obj["fields"]
    ip = obj["fields"].index("post_prob")
    for pq in obj["placements"]:
        assert abs(sum(p[ip] for p in pq["p"]) - 1.0) < 1e-9
    tree_s, fields, pqs = jplace.read(str(out))
    assert len(pqs) == len(queries)
