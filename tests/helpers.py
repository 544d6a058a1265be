"""Shared test helpers: build the oracle's view of a workload."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cut(seq, mask):
    return "".join(c for c, m in zip(seq, mask) if m)


def oracle_setup(po, wl, premask=True, **prefs):
    """(model, precompute, core, queries) for a synth.Workload, following
    pplacer_run.ml:167-341: model from the un-masked reference alignment,
    pre-mask, precompute, pull exponents at -10, supernodes."""
    S = wl.n_states
    ref_seqs = [s for _, s in wl.ref_aln]
    statd = po.emper_freq(S, ref_seqs)
    rates = po.discrete_gamma(wl.n_cats, wl.alpha)
    model = po.Model(wl.exch, statd, rates)
    queries = [s for _, s in wl.queries]
    leaf = wl.leaf_seqs()
    if premask:
        st, mask, keep, n = po.premask(S, ref_seqs, queries)
        assert st == 0
        queries = [cut(q, mask) for q in queries]
        leaf = {i: cut(s, mask) for i, s in leaf.items()}
    N = len(next(iter(leaf.values())))
    pre = po.Precompute(model, wl.tree, leaf, N)
    pre.pull_and_supernodes()
    core = po.Core(pre, **prefs)
    return model, pre, core, queries


def oracle_setup_cat(po, wl, cat_rates, site_cat, premask=True, **prefs):
    """oracle_setup for the Price-CAT model: site_cat gives a category per
    column of the un-masked alignment and is cut by the pre-mask as
    Gcat_model.mask_sites does (gcat_model.ml:429-439).  Returns a dict with
    the oracle objects and the pre-masked leaf rows and categories."""
    S = wl.n_states
    ref_seqs = [s for _, s in wl.ref_aln]
    statd = po.emper_freq(S, ref_seqs)
    queries = [s for _, s in wl.queries]
    leaf = wl.leaf_seqs()
    site_cat = np.asarray(site_cat, dtype=np.int32)
    assert site_cat.shape[0] == len(ref_seqs[0])
    if premask:
        st, mask, keep, n = po.premask(S, ref_seqs, queries)
        assert st == 0
        queries = [cut(q, mask) for q in queries]
        leaf = {i: cut(s, mask) for i, s in leaf.items()}
        site_cat = site_cat[np.asarray(mask, dtype=bool)]
    N = len(next(iter(leaf.values())))
    model = po.Model(wl.exch, statd, cat_rates=cat_rates, site_cat=site_cat)
    pre = po.Precompute(model, wl.tree, leaf, N)
    pre.pull_and_supernodes()
    core = po.Core(pre, **prefs)
    return dict(model=model, pre=pre, core=core, queries=queries, leaf=leaf, N=N,
                site_cat=site_cat, statd=statd)


def ar53_price_cat():
    """(cat_rates, site_categories 0-based) of the donovan refpkg's Price-CAT
    WAG model (20 categories over 8062 columns)."""
    from pplacer_amd import refio
    pm = refio.read_phylo_model(os.path.join(GOLDEN, "price_cat", "ar53_phylo_model.json"))
    return pm.cat_rates, np.asarray(pm.site_categories, dtype=np.int32)


def like_fixture(po, name):
    from pplacer_amd import refio

    exp = json.load(open(os.path.join(GOLDEN, "like", "expected.json")))
    d = os.path.join(GOLDEN, "like", name)
    aln = refio.read_fasta(os.path.join(d, exp[name]["fasta"]))
    tree = refio.read_newick(os.path.join(d, exp[name]["tree"]))
    pm = refio.read_phylo_model(os.path.join(d, "phylo_model.jplace"))
    seqs = [s for _, s in aln]
    rates = po.discrete_gamma(pm.n_cats, pm.alpha)
    if pm.subs_model == "GTR":
        S = 4
        exch = refio.gtr_exchangeabilities(pm.subs_rates)
        statd = po.emper_freq(4, seqs)
    else:
        S = 20
        tables = json.load(open(os.path.join(GOLDEN, "prot_models.json")))[pm.subs_model]
        exch = refio.exchangeabilities_from_lower(tables["lower"])
        statd = po.emper_freq(20, seqs) if pm.empirical_frequencies else np.array(tables["statd"])
    names = dict(aln)
    leaf = {i: names[tree.names[i]] for i in tree.leaves()}
    return dict(tree=tree, leaf=leaf, exch=exch, statd=statd, rates=rates, S=S,
                N=len(seqs[0]), expected=exp[name]["log_like"], rel=exp["rel_epsilon"])
