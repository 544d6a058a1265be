"""Pins the oracle's posterior-probability path (-p): QUADPACK QNG as GSL's
gsl_integration_qng runs it (integration.ml:28), the pendant priors
(core.ml:34-56, 86-92; pplacer_run.ml:380-389) and Three_tax.calc_marg_prob
(three_tax.ml:109-215).

GSL is not vendored, so QNG is pinned by (1) its tables' defining property
(each rule integrates the polynomials its degree covers exactly; the 10-point
rule is Gauss-Legendre), (2) the known answers of GSL's own QNG test cases
(results to 1e-15 relative and error estimates to 1e-7 relative, GSL's own
test tolerances), and (3) its error paths.  calc_marg_prob is pinned against
an adaptive scipy double integral of the same likelihood surface.
"""
import math
import os
import re

import numpy as np
import pytest
from scipy import integrate

from tests.helpers import oracle_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def qng_tables():
    text = open(os.path.join(ROOT, "pplacer_amd", "csrc", "pp_qng_tables.h")).read()
    out = {}
    for name, body in re.findall(r"pp_qng_(\w+)\[\d+\] = \{(.*?)\};", text, re.S):
        out[name] = np.array([float(v) for v in body.split(",") if v.strip()])
    return out


def rule(t, level):
    """(abscissae in [-1, 1], weights) of the 10/21/43/87-point rule."""
    if level == 10:
        xs, ws = t["x1"], t["w10"]
        return np.concatenate([xs, -xs]), np.concatenate([ws, ws])
    pos = {21: [t["x1"], t["x2"]], 43: [t["x1"], t["x2"], t["x3"]],
           87: [t["x1"], t["x2"], t["x3"], t["x4"]]}[level]
    wts = {21: np.concatenate([t["w21a"], t["w21b"][:5]]),
           43: np.concatenate([t["w43a"], t["w43b"][:11]]),
           87: np.concatenate([t["w87a"], t["w87b"][:22]])}[level]
    centre = {21: t["w21b"][5], 43: t["w43b"][11], 87: t["w87b"][22]}[level]
    xs = np.concatenate(pos)
    return (np.concatenate([xs, -xs, [0.0]]), np.concatenate([wts, wts, [centre]]))


@pytest.mark.parametrize("level,degree", [(10, 19), (21, 31), (43, 64), (87, 130)])
def test_qng_tables_are_the_kronrod_patterson_rules(level, degree):
    t = qng_tables()
    x, w = rule(t, level)
    assert len(x) == level
    for name in ("x1", "x2", "x3", "x4"):
        v = t[name]
        assert np.all(np.diff(v) < 0) and v[0] < 1 and v[-1] > 0  # descending, inside
    gx, gw = np.polynomial.legendre.leggauss(10)
    np.testing.assert_allclose(np.sort(t["x1"]), gx[5:], rtol=1e-15)
    np.testing.assert_allclose(t["w10"][::-1], gw[5:], rtol=1e-14)
    # exact on Legendre polynomials (well conditioned) up to the rule's degree
    for k in range(degree + 1):
        got = float(np.dot(w, np.polynomial.legendre.legval(x, [0] * k + [1])))
        want = 2.0 if k == 0 else 0.0
        assert abs(got - want) < 2e-14, (level, k, got)
    # and not beyond it (a different rule would fail here or above)
    k = degree + 1 if (degree + 1) % 2 == 0 else degree + 2
    got = float(np.dot(w, np.polynomial.legendre.legval(x, [0] * k + [1])))
    assert abs(got) > 1e-10, (level, k, got)


def f1(alpha):
    return lambda x: x ** alpha * math.log(1.0 / x)


def f3(alpha):
    return lambda x: math.cos(2.0 ** alpha * math.sin(x))


# GSL 2.x integration/test.c, the gsl_integration_qng cases:
# (f, a, b, epsabs, epsrel, result, abserr, neval)
GSL_QNG_CASES = [
    (f1(2.6), 0.0, 1.0, 1e-1, 0.0, 7.716049379303083211e-02, 9.424302199601294244e-08, 21),
    (f1(2.6), 0.0, 1.0, 0.0, 1e-9, 7.716049382706505200e-02, 2.666891413688592066e-12, 43),
    (f1(2.6), 0.0, 1.0, 0.0, 1e-13, 7.716049382716029525e-02, 8.566535680046930668e-16, 87),
    (f1(2.6), 1.0, 0.0, 1e-1, 0.0, -7.716049379303083211e-02, 9.424302199601294244e-08, 21),
    (f1(2.6), 1.0, 0.0, 0.0, 1e-9, -7.716049382706505200e-02, 2.666891413688592066e-12, 43),
    (f3(1.3), 0.3, 2.71, 0.0, 1e-12, -7.238969575482959717e-01, 1.277676889520056369e-14, 43),
]


@pytest.mark.parametrize("case", range(len(GSL_QNG_CASES)))
def test_qng_gsl_known_answers(oracle, case):
    f, a, b, ea, er, res, err, nev = GSL_QNG_CASES[case]
    st, r, e, n, pts = oracle.qng(f, a, b, ea, er)
    assert st == 0 and n == nev and len(pts) == nev
    assert abs(r - res) <= 1e-15 * abs(res), (r, res)
    assert abs(e - err) <= 1e-7 * err, (e, err)


def test_qng_error_paths_and_order(oracle):
    t = qng_tables()
    # tolerance below 50 eps with epsabs 0: EBADTOL before any evaluation
    st, r, e, n, pts = oracle.qng(math.exp, 0.0, 1.0, 0.0, 1e-15)
    assert st == 2 and n == 0 and pts == []
    # beyond the 87-point rule: ETOL with the 87-point result
    truth = 1.0 / 2.3 ** 2
    st, r, e, n, pts = oracle.qng(f1(1.3), 0.0, 1.0, 0.0, 1e-10)
    assert st == 1 and n == 87 and abs(r - truth) < e
    # evaluation order: centre, then c + h x1[0], c - h x1[0], ...
    st, r, e, n, pts = oracle.qng(math.exp, 1.0, 3.0, 0.0, 1.0)
    assert n == 21
    assert pts[0] == 2.0 and pts[1] == 2.0 + t["x1"][0] and pts[2] == 2.0 - t["x1"][0]
    assert pts[11] == 2.0 + t["x2"][0]
    assert abs(r - (math.exp(3) - math.exp(1))) < 1e-12


def test_exponential_pdf(oracle):
    lib = oracle.lib()
    assert lib.or_exponential_pdf(-1e-9, 0.3) == 0.0
    assert lib.or_exponential_pdf(0.0, 0.3) == 1 / 0.3
    assert abs(lib.or_exponential_pdf(0.45, 0.3) - math.exp(-1.5) / 0.3) < 1e-15


def _leaf_paths(tree, i):
    """Path lengths from node i to each leaf below it, by walking down."""
    if not tree.children[i]:
        return [0.0]
    out = []
    for c in tree.children[i]:
        out.extend(tree.bl[c] + d for d in _leaf_paths(tree, c))
    return out


def test_prior_means_and_tree_length(oracle):
    """The flat prior mean (tree length / edges) and Core.prior_mean_map:
    lower + bl(i) + the mean path length from node i to its leaves, checked
    against a direct walk; the product's host functions equal the oracle's."""
    from pplacer_amd import placer, synth
    wl = synth.make_workload("gtr50", seed=2, n_queries=1)
    tree = wl.tree
    assert abs(oracle.tree_length(tree) - float(np.sum(tree.bl))) < 1e-12
    assert placer.tree_length(tree) == oracle.tree_length(tree)
    for lower in (0.0, 0.1):
        m = oracle.prior_mean_map(tree, lower)
        for i in range(tree.n_nodes - 1):
            want = lower + tree.bl[i] + float(np.mean(_leaf_paths(tree, i)))
            assert abs(m[i] - want) < 1e-12, (i, m[i], want)
        np.testing.assert_array_equal(placer.prior_mean_map(tree, lower)[:-1], m[:-1])


@pytest.mark.parametrize("prior", ["flat", "uniform", "informative"])
def test_marginal_matches_double_integral(oracle, prior):
    """calc_marg_prob = base + log(cut^-1 \\int_0^cut \\int_0^U exp(ll - base)
    prior(p) dp dd): against scipy's adaptive dblquad over [0, max_pend] of the
    same likelihood surface.  The oracle stops at U, where ll + log prior has
    dropped 10 units (three_tax.ml:109-163), which costs ~e^-10 of the mass;
    a dropped prior, base or 1/cut term moves the result by O(1)."""
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=3, n_queries=4)
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True, prior=prior,
                                             prior_lower=0.05)
    mu_flat = oracle.tree_length(wl.tree) / (wl.tree.n_nodes - 1)
    for q in queries[:2]:
        st, pl, _ = core.place(q)
        assert st == 0
        for loc, _, _, dist, pend, _, marg in pl[:2]:
            s2, m2, calls = core.marg_prob(q, loc, pend, dist)
            assert s2 == 0 and m2 == marg and calls > 21 * 21
            base = core.log_like(q, loc, pend, dist)
            mu = mu_flat if prior == "flat" else core.prior_means[loc]
            pr = (lambda x: 1.0) if prior == "uniform" else (lambda x, mu=mu: math.exp(-x / mu) / mu)
            cut = float(wl.tree.bl[loc])
            val, _ = integrate.dblquad(
                lambda p, d: math.exp(core.log_like(q, loc, p, d) - base) * pr(p),
                0.0, cut, 0.0, 2.0, epsabs=0, epsrel=1e-9)
            ref = base + math.log(val / cut)
            assert abs(marg - ref) < 2e-4, (prior, loc, marg, ref)


def test_marginal_short_edge_integrates_pendant_only(oracle):
    """cut_bl =~ 0 (|cut| < 1e-5, ppatteries.ml:96): distal fixed at cut/2 and
    only the pendant integral, with no division by cut (three_tax.ml:199-201)."""
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=3, n_queries=2)
    loc = 4
    wl.tree.bl[loc] = 4e-6
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True)
    q = queries[0]
    st, marg, calls = core.marg_prob(q, loc, 0.05, 2e-6)
    assert st == 0
    base = core.log_like(q, loc, 0.05, 2e-6)
    mu = oracle.tree_length(wl.tree) / (wl.tree.n_nodes - 1)
    val, _ = integrate.quad(lambda p: math.exp(core.log_like(q, loc, p, 2e-6) - base)
                            * math.exp(-p / mu) / mu, 0.0, 2.0, epsabs=0, epsrel=1e-10)
    assert abs(marg - (base + math.log(val))) < 1e-5


def test_marginal_tolerance_errors(oracle):
    """A relative tolerance QNG cannot meet with epsabs = 0 raises GSL's
    EBADTOL, which calc_marg_prob does not catch."""
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=3, n_queries=2)
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True, pp_rel_err=1e-15)
    st, marg, calls = core.marg_prob(queries[0], 3, 0.05, wl.tree.bl[3] / 2)
    assert st != 0


def test_posterior_probabilities(oracle):
    """-p: kept placements carry post_prob = ll_normalized_prob of their
    marginal likelihoods (core.ml:393-401): they sum to 1 and follow the
    marginals' order; -p's prefs (prefs.ml:198-203) widen the kept set."""
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=5, n_queries=5)
    model, pre, core, queries = oracle_setup(oracle, wl, calc_pp=True)
    assert core.prefs["keep_at_most"] == 20 and core.prefs["keep_factor"] == 0.001
    assert core.prefs["max_strikes"] == 20
    for q in queries:
        st, pl, _ = core.place(q)
        assert st == 0
        post = np.array([p[5] for p in pl])
        marg = np.array([p[6] for p in pl])
        assert abs(post.sum() - 1.0) < 1e-12
        np.testing.assert_allclose(post, np.exp(marg - marg.max()) / np.exp(marg - marg.max()).sum(),
                                   rtol=1e-12)
        # without -p the same query gives the same ML placements
        _, _, core0, _ = oracle_setup(oracle, wl, keep_at_most=20, keep_factor=0.001,
                                      max_strikes=20)
        st0, pl0, _ = core0.place(q)
        assert [p[:5] for p in pl0] == [p[:5] for p in pl]
        assert all(p[5] is None for p in pl0)
