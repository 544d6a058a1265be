"""Pins for the CPU oracle (oracle/pp_oracle.c) against the reference's own
tests and fixtures, closed forms and an independent brute-force likelihood."""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from tests import bruteforce as bf
from tests.helpers import like_fixture, oracle_setup


# ---- tests/pplacer/test_linear.ml:4-29 ------------------------------------

def test_masked_logdot_closed_forms(oracle):
    mask = np.array([1, 0, 1], dtype=np.uint16)
    m1 = np.array([[-1.0, 0.0, 1.0], [1000.0, 1.0, 3.0], [0.0, 0.0, 2.0]])
    m2 = np.array([[-1.0, 1.0, 2.0], [1.0, 3.0, 0.0], [0.0, 1.0, 2.0]])
    assert abs(oracle.mat_masked_logdot(m1, m2, mask) - (math.log(3) + math.log(4))) < 1e-5
    t1 = np.array([m1, m1])
    t2 = np.array([[[-1.0, 1.0, 0.0], [1.0, 3.0, 0.0], [0.0, 1.0, 2.0]],
                   [[-1.0, 1.0, 2.0], [1.0, 3.0, 0.0], [0.0, 1.0, 3.0]]])
    expected = (math.log(1 + 3) + math.log(4 + 6)) - 2 * math.log(2)
    assert abs(oracle.ten_masked_logdot(t1, t2, mask) - expected) < 1e-5


def test_bounded_logdot_equals_full_mask(oracle):
    rng = np.random.default_rng(3)
    x, y = rng.random((4, 50, 4)), rng.random((4, 50, 4))
    a = oracle.ten_bounded_logdot(x, y, 0, 49)
    b = oracle.ten_masked_logdot(x, y, np.ones(50, np.uint16))
    assert a == b
    # a sub-interval equals the mask of that interval
    m = np.zeros(50, np.uint16)
    m[10:31] = 1
    assert oracle.ten_bounded_logdot(x, y, 10, 30) == oracle.ten_masked_logdot(x, y, m)


# ---- tests/pplacer/test_misc.ml:3-7 ----------------------------------------

def test_mask_reindex(oracle):
    assert list(oracle.mask_reindex([False, False])) == []
    assert list(oracle.mask_reindex([True, False, False, True, False, True])) == [0, 3, 5]


# ---- pplacer_src/gamma.ml:46-50 ---------------------------------------------

def test_discrete_gamma_reference_values(oracle):
    want = [0.0333877533835995541, 0.251915917593438454, 0.820268481973647878,
            2.89442784704931411]
    got = oracle.discrete_gamma(4, 0.5)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=0)
    for k, a in [(1, 0.7), (4, 0.3), (8, 1.5), (20, 0.2)]:
        r = oracle.discrete_gamma(k, a)
        assert abs(r.mean() - 1.0) < 1e-12
        assert np.all(np.diff(r) > 0) or k == 1


def test_incomplete_gamma_against_scipy(oracle):
    from scipy import special
    for a in (0.2, 0.5, 1.5, 4.0):
        for x in (1e-4, 0.1, 1.0, 3.0, 20.0):
            assert abs(oracle.lib().or_gamma_inc_P(a, x) - special.gammainc(a, x)) < 1e-13
        for p in (0.01, 0.25, 0.5, 0.75, 0.99):
            x = oracle.lib().or_gamma_Pinv(a, p)
            assert abs(special.gammainc(a, x) - p) < 1e-13


# ---- tests/pplacer/test_diagd.ml:7-47 (transposed, see DESIGN) --------------

def test_diagd_r_verified_exponentials(oracle):
    b = np.array([[-1.0, 0.15], [0.15, -2.0]])
    d = np.array([0.25, 0.75])
    x, l, xit = oracle.of_d_b(d, b)
    exp1 = np.array([[0.77992929, 0.01668155], [0.05004464, 0.22387769]])
    exp4 = np.array([[0.37187386, 0.011053085], [0.03315925, 0.003437709]])
    # Diagd.to_exp returns expm(B D) = expm(D B)^T; the R values are expm(D B)
    np.testing.assert_allclose(oracle.to_exp(x, l, xit, 1.0), exp1.T, atol=1e-8)
    np.testing.assert_allclose(oracle.to_exp(x, l, xit, 4.0), exp4.T, atol=1e-8)
    t = oracle.multi_exp(x, l, xit, [2.0, 8.0], 0.5)
    np.testing.assert_allclose(t[0], exp1.T, atol=1e-8)
    np.testing.assert_allclose(t[1], exp4.T, atol=1e-8)
    # dediagonalize(l) remakes B D (test_diagd.ml:12-15, transposed likewise)
    np.testing.assert_allclose(oracle.dediagonalize(x, l, xit), b @ np.diag(d), atol=1e-14)


@pytest.mark.parametrize("S", [4, 20])
def test_transition_matrices_closed_form(oracle, S):
    rng = np.random.default_rng(S)
    r = np.zeros((S, S))
    iu = np.triu_indices(S, 1)
    r[iu] = rng.exponential(1.0, len(iu[0]))
    r = r + r.T
    pi = rng.dirichlet(np.ones(S))
    x, l, xit = oracle.normed_of_exchangeable_pair(r, pi)
    Q = bf.rate_matrix(r, pi)
    assert abs(-np.sum(pi * np.diag(Q)) - 1) < 1e-14
    np.testing.assert_allclose(oracle.dediagonalize(x, l, xit), Q, atol=1e-12)
    rates = oracle.discrete_gamma(4, 0.5)
    for t in (0.0, 1e-3, 0.05, 0.7, 3.0):
        P = oracle.multi_exp(x, l, xit, rates, t)
        for k in range(4):
            np.testing.assert_allclose(P[k], sl.expm(Q * rates[k] * t), atol=1e-12)
            np.testing.assert_allclose(P[k].sum(axis=1), 1.0, atol=1e-12)
            np.testing.assert_allclose(pi @ P[k], pi, atol=1e-12)
    P1 = oracle.to_exp(x, l, xit, 0.3)
    P2 = oracle.to_exp(x, l, xit, 0.5)
    np.testing.assert_allclose(P1 @ P2, oracle.to_exp(x, l, xit, 0.8), atol=1e-12)
    np.testing.assert_allclose(oracle.to_exp(x, l, xit, 0.0), np.eye(S), atol=1e-13)


def test_symm_eigs_reconstructs(oracle):
    # tests/pplacer/test_matrix.ml:9-14: columns are eigenvectors
    rng = np.random.default_rng(7)
    a = rng.standard_normal((6, 6))
    a = a + a.T
    l, u = oracle.symm_eigs(a)
    np.testing.assert_allclose(u @ np.diag(l) @ u.T, a, atol=1e-12)
    np.testing.assert_allclose(u.T @ u, np.eye(6), atol=1e-12)


# ---- tests/pplacer/test_like.ml:16-97: PhyML tree likelihood at every node --

@pytest.mark.parametrize("name", ["gtr", "jtt", "wag"])
def test_tree_likelihood_matches_phyml_at_every_node(oracle, name):
    fx = like_fixture(oracle, name)
    model = oracle.Model(fx["exch"], fx["statd"], fx["rates"])
    pre = oracle.Precompute(model, fx["tree"], fx["leaf"], fx["N"])
    pre.pull_and_supernodes()
    S, N, R = fx["S"], fx["N"], model.R
    one = (np.ones((R, N, S)), np.zeros(N, dtype=np.int64))
    want, rel = fx["expected"], fx["rel"]
    sn = []
    for i in range(fx["tree"].n_nodes):
        half = fx["tree"].bl[i] / 2.0
        ud = model.evolve(pre.darr[i], pre.darr_e[i], half)
        up = model.evolve(pre.parr[i], pre.parr_e[i], half)
        a = model.slow_log_like3(ud, up, one)
        b = model.logdot((pre.snodes[i], pre.snodes_e[i]), one)
        assert abs(a - want) <= rel * abs(want), (i, a)
        assert abs(b - want) <= rel * abs(want), (i, b)
        sn.append(b)
    # the supernode log-likelihood is the same tree likelihood at every edge
    assert max(sn) - min(sn) < 1e-9 * abs(want)


def test_tree_likelihood_bruteforce(oracle):
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=5)
    S = 4
    ref_seqs = [s for _, s in wl.ref_aln]
    statd = oracle.emper_freq(S, ref_seqs)
    rates = oracle.discrete_gamma(4, wl.alpha)
    model = oracle.Model(wl.exch, statd, rates)
    leaf = wl.leaf_seqs()
    N = len(ref_seqs[0])
    pre = oracle.Precompute(model, wl.tree, leaf, N)
    pre.pull_and_supernodes()
    want = bf.tree_loglike(wl.tree, leaf, wl.exch, statd, rates)
    one = (np.ones((4, N, S)), np.zeros(N, dtype=np.int64))
    for i in range(wl.tree.n_nodes):
        got = model.logdot((pre.snodes[i], pre.snodes_e[i]), one)
        assert abs(got - want) < 1e-9 * abs(want)


# ---- exponent pulling -------------------------------------------------------

def test_pull_exponent_is_exact_power_of_two(oracle):
    rng = np.random.default_rng(11)
    a = rng.random((4, 30, 4)) * np.exp2(-rng.integers(0, 80, size=(1, 30, 1)))
    e = rng.integers(-5, 5, size=30)
    a2, e2 = oracle.perhaps_pull_exponent(-10, a, e)
    # value * 2^e preserved exactly
    np.testing.assert_array_equal(a2 * np.exp2(e2 - e)[None, :, None], a)
    for site in range(30):
        mx = a[:, site, :].max()
        te = math.frexp(mx)[1]
        if te < -10:
            assert e2[site] - e[site] == te
            assert 0.5 <= a2[:, site, :].max() < 1.0
        else:
            assert e2[site] == e[site]


# ---- stage 1 and stage 2 against the brute-force likelihood -----------------

@pytest.fixture(scope="module")
def small(oracle):
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=2)
    model, pre, core, queries = oracle_setup(oracle, wl)
    leaf = {i: s for i, s in zip(wl.tree.leaves(), [None] * 0)}
    return wl, model, pre, core, queries


def _cut_leaf(wl, pre):
    rows = pre.rows
    return {i: rows[i].tobytes().decode() for i in wl.tree.leaves()}


def test_stage1_score_is_midpoint_placement_likelihood(oracle, small):
    wl, model, pre, core, queries = small
    leaf = _cut_leaf(wl, pre)
    for q in queries[:4]:
        scores = core.stage1_scores(q)
        for loc in range(0, wl.tree.n_nodes - 1, 3):
            cut_bl = wl.tree.bl[loc]
            want = bf.placement_loglike(wl.tree, leaf, wl.exch, model.statd, model.rates, loc,
                                        cut_bl / 2.0, 0.1, q)
            assert abs(scores[loc] - want) < 1e-9 * abs(want)


def test_three_tax_likelihood_and_optimum(oracle, small):
    wl, model, pre, core, queries = small
    leaf = _cut_leaf(wl, pre)
    q = queries[1]
    scores = core.stage1_scores(q)
    loc = int(core.rank(scores)[0])
    st, ll, pend, dist, calls = core.optimize(q, loc, 0.1, wl.tree.bl[loc] / 2.0, 1e-5)
    assert st == 0 and calls > 10
    want = bf.placement_loglike(wl.tree, leaf, wl.exch, model.statd, model.rates, loc, dist,
                                pend, q)
    assert abs(ll - want) < 1e-9 * abs(want)
    # the optimum is no worse than a dense grid (within tolerance effects)
    cut_bl = wl.tree.bl[loc]
    grid = max(bf.placement_loglike(wl.tree, leaf, wl.exch, model.statd, model.rates, loc, d, p,
                                    q)
               for d in np.linspace(0, cut_bl, 9) for p in np.geomspace(1e-6, 2.0, 25))
    assert ll >= grid - 1e-3


def test_rank_is_stable_descending_with_nan_last(oracle):
    s = np.array([1.0, 3.0, np.nan, 3.0, -np.inf, 2.0, 1.0])
    assert list(oracle.Core.rank(s)) == [1, 3, 5, 0, 6, 4, 2]


# ---- Brent (GSL restatement) --------------------------------------------------

# The example of the GSL reference manual's "One Dimensional Minimization"
# chapter (GSL 2.x, doc/examples/min.c): Brent on f(x) = cos(x) + 1 from
# x = 2 in [0, 6], iterated until gsl_min_test_interval(a, b, 0.001, 0) holds.
# Each row is the manual's printed line: iteration, [lower, upper], minimum,
# minimum - pi, upper - lower, as "%5d [%.7f, %.7f] %.7f %+.7f %.7f".
GSL_MANUAL_BRENT_ROWS = [
    "    0 [0.0000000, 6.0000000] 2.0000000 -1.1415927 6.0000000",
    "    1 [2.0000000, 6.0000000] 3.5278640 +0.3862713 4.0000000",
    "    2 [2.0000000, 3.5278640] 3.1748217 +0.0332290 1.5278640",
    "    3 [2.0000000, 3.1748217] 3.1264576 -0.0151351 1.1748217",
    "    4 [3.1264576, 3.1748217] 3.1414743 -0.0001183 0.0483641",
    "    5 [3.1414743, 3.1748217] 3.1415930 +0.0000004 0.0333474",
    "    6 [3.1414743, 3.1415930] 3.1415927 +0.0000000 0.0001187",
]


def test_brent_iterates_reproduce_the_gsl_manual_example(oracle):
    """Pins the restated GSL Brent step (golden section, parabolic step and its
    admissibility test, the v/w/x history update) to GSL's own published
    iterates: a wrong parabola, a wrong acceptance test or a wrong history
    update moves a printed digit."""
    b = oracle.GslBrent(lambda x: math.cos(x) + 1.0)
    assert b.set(2.0, 0.0, 6.0) == 0
    rows = []
    converged_at = None
    for it in range(0, 100):
        if it > 0:
            assert b.iterate() == 0
        s = b.s
        m, lo, hi = s.x_minimum, s.x_lower, s.x_upper
        rows.append("%5d [%.7f, %.7f] %.7f %+.7f %.7f" % (it, lo, hi, m, m - math.pi, hi - lo))
        if it > 0 and abs(hi - lo) < 0.001:  # gsl_min_test_interval(a, b, 0.001, 0.0)
            converged_at = it
            break
    assert rows == GSL_MANUAL_BRENT_ROWS
    assert converged_at == 6
    # set evaluates f(lower), f(upper), f(minimum), then init evaluates f(v)
    assert b.points[:4] == [0.0, 6.0, 2.0, 0.0 + 0.3819660 * 6.0]
    # one evaluation per iteration
    assert len(b.points) == 4 + 6


def test_brent_set_errors_and_max_iter(oracle):
    b = oracle.GslBrent(lambda x: math.cos(x) + 1.0)
    assert b.set(2.0, 6.0, 0.0) == 2        # lower > upper: GSL_EINVAL
    assert b.set(0.0, 0.0, 6.0) == 2        # minimum on the boundary
    assert b.set(0.2, 0.0, 6.0) == 2        # f(0.2) > f(6): endpoints do not enclose a minimum
    b = oracle.GslBrent(lambda x: float("nan") if x > 5 else math.cos(x) + 1.0)
    assert b.set(2.0, 0.0, 6.0) == 1        # non-finite f(upper): GSL_EBADFUNC
    # Minimization.brent's ExceededMaxIter (minimization.ml:95-99): the cos
    # example needs 6 iterations at tolerance 1e-3, so 5 are too few; the
    # start finder accepts x = 2 at once (f(2) < min(f(0), f(6)))
    st, x, calls = oracle.brent(lambda v: math.cos(v) + 1.0, 2.0, 0.0, 6.0, 1e-3, max_iters=5)
    assert st == 3
    st, x, calls = oracle.brent(lambda v: math.cos(v) + 1.0, 2.0, 0.0, 6.0, 1e-3, max_iters=6)
    assert st == 0 and abs(x - math.pi) < 1e-6
    # f(right), f(left), f(start), then the GSL set's three and init's one, then 6 steps
    assert calls[:7] == [6.0, 0.0, 2.0, 0.0, 6.0, 2.0, 0.3819660 * 6.0] and len(calls) == 13


# ---- Brent: closed-form checks of the minimization.ml wrapper -----------------

def test_brent_closed_forms(oracle):
    st, x, calls = oracle.brent(lambda v: (v - 0.3) ** 2, 0.9, 0.0, 2.0, 1e-6)
    assert st == 0 and abs(x - 0.3) < 1e-6
    # evaluation order: f(right), f(left), then the start (minimization.ml:14)
    assert calls[:3] == [2.0, 0.0, 0.9]
    st, x, calls = oracle.brent(lambda v: v - math.log(v), 0.5, 1e-8, 2.0, 1e-7)
    assert st == 0 and abs(x - 1.0) < 1e-6
    # monotone objective: the bisection start finder reports the lower boundary
    st, x, calls = oracle.brent(lambda v: v, 1.0, 0.0, 2.0, 0.01)
    assert st == 0 and x == 0.0
    # invalid start
    st, x, calls = oracle.brent(lambda v: v, 0.0, 0.0, 2.0, 0.01)
    assert st == 5
    # non-finite value inside the GSL part
    st, x, calls = oracle.brent(lambda v: (v - 1) ** 2 if v < 1.2 else float("inf"), 0.5, 0.0, 2.0,
                                1e-3)
    assert st == 0 or st == 1


# ---- baseball, keep and like-weight ratios ------------------------------------

def test_ll_normalized_prob(oracle):
    ll = np.array([-10.0, -11.0, -10.5, -30.0])
    p = oracle.ll_normalized_prob(ll)
    np.testing.assert_allclose(p, np.exp(ll - ll.max()) / np.exp(ll - ll.max()).sum(), rtol=1e-14)
    assert abs(p.sum() - 1) < 1e-15


def test_baseball_evaluate_all_finds_exhaustive_best(oracle):
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=4)
    _, _, core_all, queries = oracle_setup(oracle, wl, max_strikes=0)
    _, _, core_bb, _ = oracle_setup(oracle, wl)
    for q in queries[:6]:
        st, pl, stats = core_all.place(q)
        assert st == 0 and stats.n_pitched == wl.tree.n_nodes - 1
        st2, pl2, stats2 = core_bb.place(q)
        assert st2 == 0 and stats2.n_pitched <= stats.n_pitched
        assert abs(sum(p[1] for p in pl) - 1.0) < 1e-12
        assert all(pl[i][1] >= pl[i + 1][1] for i in range(len(pl) - 1))
        # the exhaustive best is at least as good as the baseball best
        assert pl[0][2] >= pl2[0][2] - 1e-6


def test_premask_definitions(oracle):
    ref = ["AC-T", "A--T", "----"[:0] + "AC-?"]
    qs = ["-C-T", "A---"]
    st, mask, keep, n = oracle.premask(4, ref, qs)
    assert st == 0 and list(mask) == [True, True, False, True] and n == 3
    st, mask, keep, n = oracle.premask(4, ref, ["--G-"])
    assert st != 0   # no overlap with any reference column
    st, mask, keep, n = oracle.premask(4, ref, ["--G-", "A---"], discard_nonoverlapped=True)
    assert st == 0 and list(keep) == [False, True] and list(mask) == [True, False, False, False]
    st, mask, keep, n = oracle.premask(4, ref, ["AZ--"])
    assert st != 0   # unknown base


def test_emper_freq(oracle):
    f = oracle.emper_freq(4, ["AACG", "T-?N"])
    # A A C G T + N spread evenly; gaps skipped
    want = np.array([2, 1, 1, 1]) + 0.25
    np.testing.assert_allclose(f, want / want.sum(), rtol=1e-15)
