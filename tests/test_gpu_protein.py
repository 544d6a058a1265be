"""GPU parity of the 20-state (protein) placement path against the oracle,
through the C ABI.

The tree passes and stage 1 use the stubs' arithmetic (GLVs bit-equal,
stage-1 scores within 1e-12 relative).  The FP64 stage 2 evaluates the same
likelihood in the eigenbasis of Q instead of forming P(t), so probe
likelihoods agree to rounding; the FP32 stage 2 evolves with FP32 P(t) on
the matrix cores.  Placements are held to the nucleotide path's bar: the
same best edge (or a near-tie), likelihood within 1e-9 relative in FP64 and
1e-4 absolute in FP32 (refined in FP64), branch lengths within 2e-4.
"""
import numpy as np
import pytest

from tests.helpers import like_fixture, oracle_setup
from tests.test_gpu_parity import _compare_placements

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


@pytest.fixture(scope="module")
def prot10(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload prot10 seed 4, 12 queries")
    wl = synth.make_workload("prot10", seed=4, n_queries=12)
    model, pre, core, oq = oracle_setup(oracle, wl)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    assert queries == oq
    yield dict(wl=wl, model=model, pre=pre, core=core, pl=pl, queries=queries)
    pl.close()


def test_protein_glvs_match_oracle(prot10):
    _print_case("prot10 model + tree GLVs")
    pl, model, pre = prot10["pl"], prot10["model"], prot10["pre"]
    x, l, xit = pl.model_diagd()
    np.testing.assert_allclose(l, model.l, rtol=1e-13, atol=1e-15)
    d, de, p, pe, s, se = pl.download_glvs()
    np.testing.assert_array_equal(de, pre.darr_e)
    np.testing.assert_array_equal(pe, pre.parr_e)
    np.testing.assert_allclose(d, pre.darr, rtol=1e-12, atol=0)
    np.testing.assert_allclose(p, pre.parr, rtol=1e-12, atol=0)
    E = prot10["wl"].tree.n_nodes - 1
    np.testing.assert_array_equal(se[:E], pre.snodes_e[:E])
    np.testing.assert_allclose(s[:E], pre.snodes[:E], rtol=1e-12, atol=0)


def test_protein_stage1(prot10):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("prot10 stage1 fp64/fp32")
    pl, core, queries = prot10["pl"], prot10["core"], prot10["queries"]
    s64 = pl.stage1_scores(queries, 0.1, PP_FP64)
    s32 = pl.stage1_scores(queries, 0.1, PP_FP32)
    for q in range(len(queries)):
        want = core.stage1_scores(queries[q])
        np.testing.assert_allclose(s64[q], want, rtol=1e-12, atol=0)
        assert np.max(np.abs(s32[q] - want)) < 2e-3
        np.testing.assert_array_equal(core.rank(s64[q]), core.rank(want))


def test_protein_three_tax_optimize(prot10):
    from pplacer_amd.placer import PP_FP64
    _print_case("prot10 three_tax optimize")
    pl, core, queries = prot10["pl"], prot10["core"], prot10["queries"]
    E = prot10["wl"].tree.n_nodes - 1
    jobs = [(q, loc) for q in range(4) for loc in (0, E // 2, E - 1)]
    qi = [j[0] for j in jobs]
    locs = [j[1] for j in jobs]
    pend0 = [0.1] * len(jobs)
    dist0 = [float(prot10["wl"].tree.bl[loc]) / 2 for loc in locs]
    ll, pend, dist, calls, st = pl.optimize(queries, qi, locs, pend0, dist0, 0.01, 2.0, PP_FP64)
    for j, (q, loc) in enumerate(jobs):
        ost, oll, opend, odist, ocalls = core.optimize(queries[q], loc, pend0[j], dist0[j], 0.01)
        assert st[j] == ost == 0
        assert abs(ll[j] - oll) <= 1e-9 * abs(oll), (j, ll[j], oll)
        assert abs(pend[j] - opend) < 2e-2 and abs(dist[j] - odist) < 2e-2


def test_protein_fp32_mfma_optimize(prot10):
    """The FP32 form's 20-state engine (state-basis evolves on the matrix
    cores, pp_engine_prot32.h) on optimize jobs at the final tolerance
    (1e-5, core.ml:14).  Its error against FP64 at the same branch lengths
    (the job's last probe, whose lengths it reports) is held to the north
    star's FP32 bar, 1e-4 absolute.  Its Brent path differs from the FP64
    one, so its optimum is held to the FP32 bar plus what the likelihood's
    FP32 noise lets Brent settle on: 5e-4 of the oracle's optimum, lengths
    within 1e-2."""
    from pplacer_amd.placer import PP_FP32
    _print_case("prot10 three_tax optimize fp32 (MFMA)")
    pl, core, queries = prot10["pl"], prot10["core"], prot10["queries"]
    E = prot10["wl"].tree.n_nodes - 1
    jobs = [(q, loc) for q in range(len(queries)) for loc in (0, E // 3, E // 2, E - 1)]
    qi = [j[0] for j in jobs]
    locs = [j[1] for j in jobs]
    pend0 = [0.1] * len(jobs)
    dist0 = [float(prot10["wl"].tree.bl[loc]) / 2 for loc in locs]
    ll, pend, dist, calls, st = pl.optimize(queries, qi, locs, pend0, dist0, 1e-5, 2.0, PP_FP32)
    e_eval, e_opt, e_len = [], [], []
    for j, (q, loc) in enumerate(jobs):
        ost, oll, opend, odist, ocalls = core.optimize(queries[q], loc, pend0[j], dist0[j], 1e-5)
        assert st[j] == ost == 0
        at = core.log_like(queries[q], loc, pend[j], dist[j])
        e_eval.append(abs(ll[j] - at))
        e_opt.append(abs(ll[j] - oll))
        e_len.append(max(abs(pend[j] - opend), abs(dist[j] - odist)))
    print("fp32 MFMA optimize, %d jobs: max |ll - ll_fp64(same lengths)| = %.3g, "
          "max |ll - oracle optimum| = %.3g, max length difference %.3g"
          % (len(jobs), max(e_eval), max(e_opt), max(e_len)), flush=True)
    assert max(e_eval) <= 1e-4
    assert max(e_opt) <= 5e-4
    assert max(e_len) <= 1e-2


def test_protein_place_fp64(prot10):
    from pplacer_amd.placer import PP_FP64
    _print_case("prot10 place fp64")
    pl, core, queries = prot10["pl"], prot10["core"], prot10["queries"]
    got, status, stats = pl.place(queries, precision=PP_FP64)
    n = _compare_placements(core, queries, got, status, 1e-9, True)
    assert n == len(queries)


def test_protein_place_fp32(prot10):
    from pplacer_amd.placer import PP_FP32
    _print_case("prot10 place fp32")
    pl, core, queries = prot10["pl"], prot10["core"], prot10["queries"]
    got, status, stats = pl.place(queries, precision=PP_FP32)
    ties = [0]
    n = _compare_placements(core, queries, got, status, 1e-4, False, ties)
    assert n + ties[0] == len(queries)


@pytest.mark.parametrize("name", ["jtt", "wag"])
def test_protein_reference_fixtures(oracle, name):
    """The reference's own protein likelihood fixtures (test_like.ml): GPU
    tree passes equal the oracle, and the log-likelihood at every edge
    (logdot of a supernode with the unit GLV) is PhyML's."""
    from pplacer_amd import placer
    _print_case("like fixture %s" % name)
    f = like_fixture(oracle, name)
    pl = placer.Placer(0)
    try:
        pl.build_model(f["exch"], f["statd"], f["rates"])
        pl.set_tree(f["tree"], f["leaf"], f["N"])
        model = oracle.Model(f["exch"], f["statd"], f["rates"])
        pre = oracle.Precompute(model, f["tree"], f["leaf"], f["N"])
        pre.pull_and_supernodes()
        d, de, p, pe, s, se = pl.download_glvs()
        np.testing.assert_array_equal(de, pre.darr_e)
        np.testing.assert_allclose(d, pre.darr, rtol=1e-12, atol=0)
        np.testing.assert_allclose(p, pre.parr, rtol=1e-12, atol=0)
        E = f["tree"].n_nodes - 1
        one = (np.ones((model.R, f["N"], 20)), np.zeros(f["N"], dtype=np.int64))
        for e in (0, E // 2, E - 1):
            ll = model.logdot((s[e], se[e]), one)
            assert abs(ll - f["expected"]) <= f["rel"] * abs(f["expected"]), (e, ll, f["expected"])
    finally:
        pl.close()


# ---- the protein parity shape (500 taxa, 400 columns) ------------------------------

@pytest.fixture(scope="module")
def prot500(oracle):
    from pplacer_amd import placer, synth
    _print_case("workload prot500 seed 2, 6 queries")
    wl = synth.make_workload("prot500", seed=2, n_queries=6)
    model, pre, core, oq = oracle_setup(oracle, wl)
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    assert queries == oq
    yield dict(wl=wl, model=model, pre=pre, core=core, pl=pl, queries=queries)
    pl.close()


def test_protein_full_size(prot500):
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("prot500 glvs, stage1, placements")
    pl, pre, core, queries = prot500["pl"], prot500["pre"], prot500["core"], prot500["queries"]
    d, de, p, pe, s, se = pl.download_glvs()
    np.testing.assert_array_equal(de, pre.darr_e)
    np.testing.assert_array_equal(pe, pre.parr_e)
    np.testing.assert_allclose(d, pre.darr, rtol=1e-12, atol=0)
    s64 = pl.stage1_scores(queries, 0.1, PP_FP64)
    np.testing.assert_allclose(s64[0], core.stage1_scores(queries[0]), rtol=1e-12, atol=0)
    sub = queries[:3]
    got, status, stats = pl.place(sub, precision=PP_FP64)
    assert _compare_placements(core, sub, got, status, 1e-9, True) == len(sub)
    got32, status32, _ = pl.place(sub, precision=PP_FP32)
    assert np.all(status32 == 0)
    ties = [0]
    assert _compare_placements(core, sub, got32, status32, 1e-4, False, ties) + ties[0] == len(sub)
