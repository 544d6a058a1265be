"""The file driver end to end on the GPU with a reference package, as
`pplacer -c my.refpkg queries.fasta` runs (pplacer_run.ml:682-783): the tree,
the Stockholm reference alignment and the phylo_model come from the refpkg,
the queries from their own file; a protein refpkg names the built-in WAG
model (prot_models.ml:80-105) with its own frequencies.  Every placement in
the jplace equals the oracle's on the same inputs (edge, likelihood within
1e-9 relative, branch lengths within twice final_tolerance)."""
import json
import os

import numpy as np
import pytest

from tests.helpers import cut

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _refpkg(root, wl, model):
    os.makedirs(root, exist_ok=True)
    with open(os.path.join(root, "ref.tre"), "w") as f:
        f.write(wl.tree.to_newick(digits=17))
    with open(os.path.join(root, "model.json"), "w") as f:
        json.dump(model, f)
    with open(os.path.join(root, "ref.sto"), "w") as f:
        f.write("# STOCKHOLM 1.0\n#=GF ID synthetic\n")
        for n, s in wl.ref_aln:  # two interleaved blocks
            f.write("%s %s\n" % (n, s[:40].replace("-", ".")))
        f.write("\n")
        for n, s in wl.ref_aln:
            f.write("%s %s\n" % (n, s[40:]))
        f.write("//\n")
    with open(os.path.join(root, "CONTENTS.json"), "w") as f:
        json.dump({"files": {"tree": "ref.tre", "aln_sto": "ref.sto", "phylo_model": "model.json"},
                   "metadata": {"format_version": "1.1"}}, f)


def _oracle_placements(po, wl, exch, statd, rates):
    ref_seqs = [s for _, s in wl.ref_aln]
    queries = [s for _, s in wl.queries]
    st, mask, keep, n = po.premask(wl.n_states, ref_seqs, queries)
    leaf = {i: cut(s, mask) for i, s in wl.leaf_seqs().items()}
    model = po.Model(exch, statd, rates)
    pre = po.Precompute(model, wl.tree, leaf, n)
    pre.pull_and_supernodes()
    core = po.Core(pre)
    return {name: core.place(cut(q, mask)) for name, q in wl.queries}


@pytest.mark.parametrize("kind", ["gtr", "wag"])
def test_cli_with_refpkg_matches_oracle(tmp_path, oracle, kind):
    from pplacer_amd import jplace, placer, refio, run, synth
    wl = synth.make_workload("gtr12" if kind == "gtr" else "prot10", seed=6, n_queries=6)
    model = wl.phylo_model_json()
    if kind == "wag":
        model.update(subs_model="WAG", empirical_frequencies=False)
    _print_case("run -c refpkg (%s)" % kind)
    root = str(tmp_path / "syn.refpkg")
    _refpkg(root, wl, model)
    qf = str(tmp_path / "queries.fasta")
    refio.write_fasta(qf, wl.queries)
    out = str(tmp_path / "out.jplace")
    assert run.main(["-c", root, qf, "-o", out]) == 0
    tree_s, fields, pqs = jplace.read(out)
    got = {namlom[0][0]: pl for namlom, pl in pqs}
    # the oracle on the same inputs: frequencies as Gstar_support picks them
    ref_seqs = [s for _, s in wl.ref_aln]
    pm = refio.parse_phylo_model(model)
    S, exch, statd = refio.model_matrices(pm, ref_seqs, oracle.emper_freq)
    if kind == "wag":
        assert np.array_equal(statd, refio.prot_model("WAG")[1])
    want = _oracle_placements(oracle, wl, exch, statd, oracle.discrete_gamma(pm.n_cats, pm.alpha))
    assert set(got) == set(want)
    for name, (st, w, _) in want.items():
        assert st == 0
        g = got[name][0]
        assert g["edge_num"] == w[0][0]
        # jplace prints %.12g: compare at that precision
        assert abs(g["likelihood"] - w[0][2]) <= 1e-9 * abs(w[0][2]) + 1e-12 * abs(w[0][2])
        assert abs(g["distal_length"] - w[0][3]) < 2e-5
        assert abs(g["pendant_length"] - w[0][4]) < 2e-5
