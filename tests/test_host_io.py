"""Host I/O without OCaml: Stockholm alignments (pplacer_src/stockholm.ml),
the alignment-file suffix dispatch (common_src/alignment.ml:62-70), reference
packages (pplacer_src/refpkg_parse.ml, refpkg.ml), the query-file partition of
Pplacer_run.run_file (pplacer_run.ml:658-757) and the built-in protein models
(prot_models.ml:70-163)."""
import gzip
import json
import os
import zipfile

import numpy as np
import pytest

from pplacer_amd import refio, run
from tests.helpers import GOLDEN

STO = """# STOCKHOLM 1.0
#=GF ID example

seqB  AC.G
seqA  ac-t
#=GC SS_cons ....

seqB  TT
seqA  G.
//
"""


def test_stockholm_interleaved_blocks_gaps_and_order():
    aln = refio.parse_stockholm(STO)
    # blocks concatenate per name, '.' becomes '-', upper case, names in
    # descending order (the reference's fold over a string map conses)
    assert aln == [("seqB", "AC-GTT"), ("seqA", "AC-TG-")]


@pytest.mark.parametrize("text,msg", [
    ("seqA AC\n//\n", "before header"),
    ("# STOCKHOLM 1.0\nseqA AC\n", "footer"),
    ("# STOCKHOLM 1.0\nseqA AC\nseqB ACG\n//\n", "lengths"),
    ("# STOCKHOLM 1.0\nseqA AC\n//\nseqB AC\n", "after footer"),
    ("# STOCKHOLM 1.0\n# STOCKHOLM 1.0\n//\n", "header unexpected"),
])
def test_stockholm_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        refio.parse_stockholm(text)


def test_alignment_suffix_dispatch(tmp_path):
    (tmp_path / "a.sto").write_text(STO)
    (tmp_path / "a.fasta").write_text(">x\nac-g\n>y\nACTG\n")
    with gzip.open(tmp_path / "b.sth.gz", "wt") as f:
        f.write(STO)
    assert refio.read_alignment(str(tmp_path / "a.sto")) == refio.parse_stockholm(STO)
    assert refio.read_alignment(str(tmp_path / "b.sth.gz")) == refio.parse_stockholm(STO)
    assert refio.read_alignment(str(tmp_path / "a.fasta")) == [("x", "AC-G"), ("y", "ACTG")]
    (tmp_path / "a.txt").write_text(">x\nA\n")
    with pytest.raises(ValueError, match="unfamiliar suffix"):
        refio.read_alignment(str(tmp_path / "a.txt"))


def test_reference_test_refpkgs():
    """tests/data/{simple,tiny}.refpkg of the reference: CONTENTS.json names
    the files; the package name drops .refpkg; format_version is read."""
    for name in ("simple", "tiny"):
        rp = refio.Refpkg(os.path.join(GOLDEN, "refpkg", name + ".refpkg"))
        assert rp.name == name and rp.format_version == "1.1"
        t = rp.tree()
        assert sorted(t.names[i] for i in t.leaves()) == ["C", "D", "F", "G"]
        assert t.n_nodes == 7
        with pytest.raises(ValueError, match="alignment"):
            rp.alignment()
    # the donovan refpkg's CONTENTS.json: keys as pplacer reads them
    c = json.load(open(os.path.join(GOLDEN, "refpkg", "ar53_CONTENTS.json")))
    assert {"phylo_model", "aln_fasta", "tree"} <= set(c["files"])


def _write_refpkg(root, wl, aln_key="aln_sto"):
    os.makedirs(root, exist_ok=True)
    files = {"tree": "ref.tre", "phylo_model": "model.json"}
    with open(os.path.join(root, "ref.tre"), "w") as f:
        f.write(wl.tree.to_newick())
    with open(os.path.join(root, "model.json"), "w") as f:
        json.dump(wl.phylo_model_json(), f)
    if aln_key == "aln_sto":
        files["aln_sto"] = "ref.sto"
        with open(os.path.join(root, "ref.sto"), "w") as f:
            f.write("# STOCKHOLM 1.0\n")
            for n, s in wl.ref_aln:
                f.write("%s %s\n" % (n, s.replace("-", ".")))
            f.write("//\n")
    else:
        files["aln_fasta"] = "ref.fasta"
        refio.write_fasta(os.path.join(root, "ref.fasta"), wl.ref_aln)
    with open(os.path.join(root, "CONTENTS.json"), "w") as f:
        json.dump({"files": files, "metadata": {"format_version": "1.1"}}, f)


def test_refpkg_dir_zip_and_query_partition(tmp_path):
    from pplacer_amd import synth
    wl = synth.make_workload("gtr12", seed=2, n_queries=5)
    root = str(tmp_path / "x.refpkg")
    _write_refpkg(root, wl)
    zpath = str(tmp_path / "x.zip")
    with zipfile.ZipFile(zpath, "w") as z:
        for n in os.listdir(root):
            z.write(os.path.join(root, n), "x.refpkg/" + n)
    want_refs = sorted(wl.ref_aln)
    for path in (root, zpath):
        rp = refio.Refpkg(path)
        # a query file with queries only: the package's alignment is used
        tree, refs, queries = run.inputs_of_refpkg(rp, wl.queries)
        assert sorted(refs) == want_refs and queries == wl.queries
        assert tree.n_nodes == wl.tree.n_nodes
        np.testing.assert_allclose(tree.bl, wl.tree.bl, rtol=1e-9)
        pm = rp.phylo_model()
        assert pm.subs_model == "GTR" and pm.n_cats == wl.n_cats
    # references in the query file replace the package's alignment
    rp = refio.Refpkg(root)
    other = [(n, s.replace("A", "C")) for n, s in wl.ref_aln]
    tree, refs, queries = run.inputs_of_refpkg(rp, other + wl.queries)
    assert refs == other and queries == wl.queries
    with pytest.raises(ValueError, match="Some, but not all"):
        run.inputs_of_refpkg(rp, other[:2] + wl.queries)
    # overrides: -t / -r replace the package's entries
    alt = str(tmp_path / "alt.fasta")
    refio.write_fasta(alt, other)
    rp = refio.Refpkg(root, {"aln_fasta": alt})
    rp.files.pop("aln_sto")
    assert rp.alignment() == [(n, s.upper()) for n, s in other]
    with pytest.raises(ValueError, match="does not contain a tree"):
        refio.Refpkg(None, {"aln_fasta": alt}).tree()


def test_builtin_protein_models_equal_the_reference_tables():
    tables = json.load(open(os.path.join(GOLDEN, "prot_models.json")))
    for name in ("LG", "WAG", "JTT"):
        exch, statd = refio.prot_model(name)
        np.testing.assert_array_equal(exch, refio.exchangeabilities_from_lower(tables[name]["lower"]))
        np.testing.assert_array_equal(statd, tables[name]["statd"])
        assert exch.shape == (20, 20) and np.all(exch == exch.T) and np.all(np.diag(exch) == 0)
        assert abs(statd.sum() - 1) < 1e-4
    with pytest.raises(ValueError, match="unknown model"):
        refio.prot_model("Dayhoff")
    # Gstar_support.seqtype_and_trans_statd_of_info
    pm = refio.PhyloModel(subs_model="WAG", empirical_frequencies=False, alpha=0.5, n_cats=4,
                          subs_rates=None)
    S, exch, statd = refio.model_matrices(pm, ["ARND"], lambda s, seqs: np.full(s, 1.0 / s))
    assert S == 20 and np.array_equal(statd, refio.prot_model("WAG")[1])
    pm.empirical_frequencies = True
    S, exch, statd = refio.model_matrices(pm, ["ARND"], lambda s, seqs: np.full(s, 1.0 / s))
    assert np.all(statd == 0.05)
    pm = refio.PhyloModel(subs_model="GTR", empirical_frequencies=True, alpha=0.5, n_cats=4,
                          subs_rates=None)
    with pytest.raises(ValueError, match="no substitution rates"):
        refio.model_matrices(pm, ["ACGT"], lambda s, seqs: np.full(s, 1.0 / s))
