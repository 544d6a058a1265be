"""File-level driver: the Pplacer_run.run_file / run_placements flow
(pplacer_src/pplacer_run.ml:167-783) for the gamma mixture and Price-CAT
models, with the placement itself on the GPU.

    python -m pplacer_amd.run -c my.refpkg queries.fasta -o out.jplace [--fp32]
    python -m pplacer_amd.run -t t.nwk -r ref.sto --model phylo_model.json queries.sto

The reference package (a directory or zip with CONTENTS.json) supplies the
tree (edge numbers = pplacer's post-order ids), the reference alignment
(aln_fasta or aln_sto) and the phylo_model; -t, -r and --model replace its
entries.  The query file (FASTA or Stockholm, by suffix) holds the queries and,
optionally, every reference sequence, which then replaces the package's
alignment (pplacer_run.ml:658-680).  Protein models (LG, WAG, JTT) use the
built-in exchangeabilities of prot_models.ml; --exch replaces them with a
JSON table `{"lower": rows, "statd": [...]}`.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import jplace, placer, refio


def partition_queries(ref_names, aln):
    """Pplacer_run.partition_queries (pplacer_run.ml:658-680): (queries,
    reference rows found in the query file, or None when there are none)."""
    refs = [(n, s) for n, s in aln if n in ref_names]
    queries = [(n, s) for n, s in aln if n not in ref_names]
    if not refs:
        return queries, None
    if len(refs) != len(ref_names):
        raise ValueError("Some, but not all reference sequences found in provided alignment file. "
                         "Are there overlaps between query and reference sequence IDs?")
    seen = set()
    for n, _ in refs:  # pplacer_run.ml:741-751
        if n in seen:
            raise ValueError("duplicate reference sequence '%s' in query file" % n)
        seen.add(n)
    return queries, refs


def inputs_of_refpkg(rp: refio.Refpkg, query_aln):
    """(tree, reference rows, queries, phylo_model) as run_file assembles them
    (pplacer_run.ml:720-757): the tree's leaf names select the references, the
    query file's reference rows replace the package's alignment."""
    tree = rp.tree()
    names = [tree.names[i] for i in tree.leaves()]
    if len(set(names)) != len(names):
        raise ValueError("Repeated names in reference tree!")
    queries, refs = partition_queries(set(names), query_aln)
    if refs is None:
        refs = [(n, s) for n, s in rp.alignment() if n in set(names)]
    return tree, refs, queries


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
    ref_seqs = [s for _, s in refs]
    if exch_table is None:
        S, exch, statd = refio.model_matrices(pm, ref_seqs, placer.emper_freq)
    else:
        S = 20
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
    ap.add_argument("query", nargs="?", default=None,
                    help="query alignment (FASTA or Stockholm), optionally with the references")
    ap.add_argument("-c", "--refpkg", default=None, help="reference package (directory or zip)")
    ap.add_argument("-t", "--tree", default=None, help="reference tree (replaces the package's)")
    ap.add_argument("-r", "--ref-aln", default=None,
                    help="reference alignment (replaces the package's)")
    ap.add_argument("--aln", default=None,
                    help="one alignment with references and queries (instead of a query file)")
    ap.add_argument("--model", default=None, help="phylo_model JSON (replaces the package's)")
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
    query_file = args.query or args.aln
    if query_file is None:
        ap.error("a query file is needed")
    if args.refpkg is None and args.tree is None:
        ap.error("please specify a reference tree with -t or -c")
    ref_base = (args.ref_aln or "")[:-3] if (args.ref_aln or "").endswith(".gz") else (args.ref_aln or "")
    key = "aln_sto" if ref_base.endswith((".sto", ".sth")) else "aln_fasta"
    rp = refio.Refpkg(args.refpkg, {"tree": args.tree, "phylo_model": args.model,
                                    key: args.ref_aln})
    if args.ref_aln:  # -r replaces whichever alignment the package names
        rp.files.pop("aln_fasta" if key == "aln_sto" else "aln_sto", None)
    tree, refs, queries = inputs_of_refpkg(rp, refio.read_alignment(query_file))
    pm = rp.phylo_model()
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
