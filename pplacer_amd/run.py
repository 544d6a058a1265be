"""File-level driver: the Pplacer_run.run_file / run_placements flow
(pplacer_src/pplacer_run.ml:167-783) for the gamma mixture and Price-CAT
models, with the placement itself on the GPU.

    python -m pplacer_amd.run --tree t.nwk --aln ref_and_queries.fasta \
        --model phylo_model.json -o out.jplace [--fp32]

Reads the reference tree (edge numbers = pplacer's post-order ids), one
alignment holding reference and query sequences (a sequence whose name is a
leaf of the tree is a reference, pplacer_run.ml:658-680), and a refpkg
phylo_model JSON (gmix_model.ml:386-413, gcat_model.ml:453-478).  Protein models need the
exchangeability table passed as JSON (`--exch`, `{"lower": rows, "statd": [...]}`).
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import jplace, placer, refio


def partition(aln, tree: refio.Tree):
    leaf_names = {tree.names[i] for i in tree.leaves()}
    refs = [(n, s) for n, s in aln if n in leaf_names]
    queries = [(n, s) for n, s in aln if n not in leaf_names]
    missing = leaf_names - {n for n, _ in refs}
    if missing:
        raise ValueError("taxon not found in alignment: '%s'" % sorted(missing)[0])
    return refs, queries


def cat_sites(site_categories, mask):
    """Gcat_model.mask_sites (gcat_model.ml:429-439): keep the categories of
    the columns the pre-mask keeps.  The reference leaves them untouched when
    the lengths differ and its model check then rejects the alignment; here
    that is an error straight away."""
    mask = np.asarray(mask, dtype=bool)
    if len(site_categories) != mask.shape[0]:
        raise ValueError("Price-CAT model has %d site categories for an alignment of %d columns"
                         % (len(site_categories), mask.shape[0]))
    return np.asarray(site_categories, dtype=np.int32)[mask]


def run_placements(tree: refio.Tree, refs, queries, pm: refio.PhyloModel, exch_table=None,
                   prefs=None, precision=placer.PP_FP64, device=0, prior="flat", prior_lower=0.0):
    """Returns (pqueries, status) where pqueries is [(namlom, [PlacementRecord])]."""
    S = 4 if pm.subs_model == "GTR" else 20
    ref_seqs = [s for _, s in refs]
    if S == 4:
        exch = refio.gtr_exchangeabilities(pm.subs_rates)
        statd = placer.emper_freq(4, ref_seqs)
    else:
        if exch_table is None:
            raise ValueError("protein model %s needs its exchangeability table" % pm.subs_model)
        exch = refio.exchangeabilities_from_lower(exch_table["lower"])
        statd = (placer.emper_freq(20, ref_seqs) if pm.empirical_frequencies
                 else np.asarray(exch_table["statd"], dtype=float))
    # pre-mask (pplacer_run.ml:59-144) and de-duplication (:228-253)
    mask, keep, n = placer.premask(S, ref_seqs, [s for _, s in queries])
    cut_q = placer.cut_columns([s for _, s in queries], mask)
    uniq, groups = jplace.dedup([(nm, s) for (nm, _), s in zip(queries, cut_q)])
    name_to_seq = dict(zip((nm for nm, _ in refs), placer.cut_columns(ref_seqs, mask)))
    leaf = {i: name_to_seq[tree.names[i]] for i in tree.leaves()}
    pl = placer.Placer(device)
    try:
        if pm.ras_model == "Price-CAT":
            pl.build_model_cat(exch, statd, pm.cat_rates, cat_sites(pm.site_categories, mask))
        else:
            pl.build_model(exch, statd, placer.discrete_gamma(pm.n_cats, pm.alpha))
        pl.set_tree(tree, leaf, n)
        if prefs is not None and prefs.calc_pp:
            pl.set_prior(prior, prior_lower)  # pplacer_run.ml:380-389
        res, status, stats = pl.place([s.upper() for _, s in uniq], prefs=prefs,
                                      precision=precision)
    finally:
        pl.close()
    pqueries = []
    for (name, _), recs in zip(uniq, res):
        if recs:
            pqueries.append((jplace.redup([(name, 1.0)], groups), recs))
    return pqueries, status


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--tree", required=True)
    ap.add_argument("--aln", required=True)
    ap.add_argument("--model", required=True)
    ap.add_argument("--exch", default=None)
    ap.add_argument("-o", "--out", required=True)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--max-strikes", type=int, default=6)
    ap.add_argument("--strike-box", type=float, default=3.0)
    ap.add_argument("--max-pitches", type=int, default=40)
    ap.add_argument("--keep-at-most", type=int, default=7)
    ap.add_argument("--keep-factor", type=float, default=0.01)
    ap.add_argument("--start-pend", type=float, default=0.1)
    ap.add_argument("--max-pend", type=float, default=2.0)
    ap.add_argument("-p", dest="calc_pp", action="store_true",
                    help="posterior probabilities (sets keep-at-most 20, keep-factor 0.001, "
                         "max-strikes 20 unless given)")
    ap.add_argument("--pp-rel-err", type=float, default=0.01)
    ap.add_argument("--unif-prior", action="store_true")
    ap.add_argument("--inform-prior", action="store_true")
    ap.add_argument("--prior-lower", type=float, default=0.0)
    args = ap.parse_args(argv)
    given = set(a.split("=")[0] for a in (argv if argv is not None else sys.argv[1:]))
    if args.calc_pp:  # prefs.ml:198-203
        if "--keep-at-most" not in given:
            args.keep_at_most = 20
        if "--keep-factor" not in given:
            args.keep_factor = 0.001
        if "--max-strikes" not in given:
            args.max_strikes = 20
    tree = refio.read_newick(args.tree)
    refs, queries = partition(refio.read_fasta(args.aln), tree)
    pm = refio.read_phylo_model(args.model)
    exch_table = json.load(open(args.exch)) if args.exch else None
    prefs = placer.default_prefs(max_strikes=args.max_strikes, strike_box=args.strike_box,
                                 max_pitches=args.max_pitches, keep_at_most=args.keep_at_most,
                                 keep_factor=args.keep_factor, start_pend=args.start_pend,
                                 max_pend=args.max_pend, calc_pp=args.calc_pp,
                                 pp_rel_err=args.pp_rel_err)
    prior = "uniform" if args.unif_prior else "informative" if args.inform_prior else "flat"
    pqueries, status = run_placements(tree, refs, queries, pm, exch_table, prefs,
                                      placer.PP_FP32 if args.fp32 else placer.PP_FP64,
                                      prior=prior, prior_lower=args.prior_lower)
    jplace.write(args.out, tree, pqueries, invocation=" ".join(sys.argv))
    bad = int(np.sum(np.asarray(status) != 0))
    if bad:
        print("%d queries failed" % bad, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
