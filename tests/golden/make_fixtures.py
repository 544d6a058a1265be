"""Writes the data fixtures under tests/golden/ from the reference checkout.

Run once in a container where /root/reference exists:
    python tests/golden/make_fixtures.py

* like/{gtr,jtt,wag}/: the alignment, tree and phylo_model files that the
  reference's tests/pplacer/test_like.ml:50-124 reads (PhyML-verified tree
  log-likelihoods are recorded in like/expected.json from test_like.ml:105-124).
* prot_models.json: the WAG, JTT and LG exchangeabilities and frequencies, the
  numeric tables of pplacer_src/prot_models.ml:70-163, as plain numbers (also
  written to pplacer_amd/data/, where they are the package's built-in models).
* refpkg/: the reference's test refpkgs simple.refpkg and tiny.refpkg as they
  are, and the donovan refpkg's CONTENTS.json.
* price_cat/ar53_phylo_model.json: the Price-CAT phylo_model of the donovan
  regression refpkg (donovan-test/ar53_pplacer_testset/gtdb_r226_ar53.refpkg),
  the same JSON object written without indentation.
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def copy_like():
    src = os.path.join(REF, "tests/data/like")
    files = {
        "gtr": ["hiv.fasta", "hiv.phy_phyml_tree.txt", "phylo_model.jplace"],
        "jtt": ["actin.fasta", "actin.phy_phyml_tree.txt", "phylo_model.jplace"],
        "wag": ["actin.fasta", "actin.phy_phyml_tree.txt", "phylo_model.jplace"],
    }
    for d, names in files.items():
        os.makedirs(os.path.join(HERE, "like", d), exist_ok=True)
        for n in names:
            shutil.copy(os.path.join(src, d, n), os.path.join(HERE, "like", d, n))
    expected = {
        "gtr": {"fasta": "hiv.fasta", "tree": "hiv.phy_phyml_tree.txt", "log_like": -1752.42340},
        "jtt": {"fasta": "actin.fasta", "tree": "actin.phy_phyml_tree.txt", "log_like": -2326.22075},
        "wag": {"fasta": "actin.fasta", "tree": "actin.phy_phyml_tree.txt", "log_like": -2334.79151},
        "rel_epsilon": 1e-7,
    }
    with open(os.path.join(HERE, "like", "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)


def prot_tables():
    text = open(os.path.join(REF, "pplacer_src/prot_models.ml")).read()
    out = {}
    for name in ("wag", "lg", "jtt"):
        m = re.search(r"let %s_trans =\s*parseProtModel \[\|(.*?)\|\]" % name, text, re.S)
        rows = [[float(x) for x in row.split()] for row in re.findall(r'"([^"]*)"', m.group(1))]
        f = re.search(r'let %s_statd = parseFreq "([^"]*)"' % name, text)
        out[name.upper()] = {"lower": rows, "statd": [float(x) for x in f.group(1).split()]}
    # the fixture, and the same numbers as the package's built-in models
    for path in (os.path.join(HERE, "prot_models.json"),
                 os.path.join(HERE, "..", "..", "pplacer_amd", "data", "prot_models.json")):
        with open(path, "w") as fh:
            json.dump(out, fh)


def copy_jplace():
    """Placement files that tests/pplacer/test_placefile.ml:1-24 round-trips."""
    for d, names in {"simple": ["test1.jplace", "test2.jplace", "test3.jplace"],
                     "psbA": ["DCM.jplace", "coastal.jplace", "surface.jplace", "upwelling.jplace"]}.items():
        os.makedirs(os.path.join(HERE, "jplace", d), exist_ok=True)
        for n in names:
            dst = os.path.join(HERE, "jplace", d, n)
            shutil.copy(os.path.join(REF, "tests/data", d, n), dst)
            os.chmod(dst, 0o644)


def copy_price_cat():
    src = os.path.join(REF, "donovan-test/ar53_pplacer_testset/gtdb_r226_ar53.refpkg",
                       "phylo_model59kpwu4z.json")
    os.makedirs(os.path.join(HERE, "price_cat"), exist_ok=True)
    with open(src) as f:
        obj = json.load(f)
    with open(os.path.join(HERE, "price_cat", "ar53_phylo_model.json"), "w") as f:
        json.dump(obj, f, separators=(",", ":"))


def copy_refpkgs():
    """The reference's own test refpkgs (tests/data/{simple,tiny}.refpkg:
    CONTENTS.json and the files it names) and the donovan refpkg's
    CONTENTS.json (whose alignment is not in the checkout)."""
    for name in ("simple.refpkg", "tiny.refpkg"):
        src = os.path.join(REF, "tests/data", name)
        dst = os.path.join(HERE, "refpkg", name)
        os.makedirs(dst, exist_ok=True)
        for n in os.listdir(src):
            shutil.copy(os.path.join(src, n), os.path.join(dst, n))
            os.chmod(os.path.join(dst, n), 0o644)
    shutil.copy(os.path.join(REF, "donovan-test/ar53_pplacer_testset/gtdb_r226_ar53.refpkg/CONTENTS.json"),
                os.path.join(HERE, "refpkg", "ar53_CONTENTS.json"))
    os.chmod(os.path.join(HERE, "refpkg", "ar53_CONTENTS.json"), 0o644)


if __name__ == "__main__":
    copy_like()
    prot_tables()
    copy_jplace()
    copy_price_cat()
    copy_refpkgs()
