"""Host-side readers for the inputs of the placement path.

* Newick trees with pplacer's edge numbering: every node gets the id of its
  position in a post-order walk of the parse, the root gets the largest id and
  a zero branch length (pplacer_src/newick_parser.mly:49-62,104-108,
  pplacer_src/newick_gtree.ml:145-157).
* FASTA alignments, upper-cased (common_src/alignment.ml, ``upper_aln_of_any_file``).
* Stockholm alignments (pplacer_src/stockholm.ml) and the suffix dispatch of
  ``Alignment.list_of_any_file`` (common_src/alignment.ml:62-70).
* refpkg ``phylo_model`` JSON for the gamma mixture model
  (pplacer_src/gmix_model.ml:386-413) and the Price-CAT model.
* reference packages: a directory or zip file with a CONTENTS.json naming the
  tree, the alignment and the phylo_model (pplacer_src/refpkg_parse.ml:62-103,
  pplacer_src/refpkg.ml:27-63).
* the built-in WAG, JTT and LG protein models (pplacer_src/prot_models.ml:70-163),
  shipped as numbers in data/prot_models.json.
"""
from __future__ import annotations

import gzip
import json
import os
import re
import zipfile
from dataclasses import dataclass, field

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@dataclass
class Tree:
    """A rooted tree in pplacer's post-order numbering; root id = n_nodes - 1."""

    children: list  # children[id] -> list of child ids, in Newick order
    bl: np.ndarray  # [G] float64
    names: list  # names[id] -> leaf / node label or None
    _csr: tuple = field(default=None, repr=False)

    @property
    def n_nodes(self) -> int:
        return len(self.children)

    @property
    def root(self) -> int:
        return self.n_nodes - 1

    @property
    def n_edges(self) -> int:
        return self.n_nodes - 1

    def is_leaf(self, i: int) -> bool:
        return not self.children[i]

    def leaves(self):
        return [i for i in range(self.n_nodes) if not self.children[i]]

    def csr(self):
        """(child_start[G+1] int32, children[] int32) in Newick child order."""
        if self._csr is None:
            starts = np.zeros(self.n_nodes + 1, dtype=np.int32)
            flat = []
            for i, ch in enumerate(self.children):
                starts[i + 1] = starts[i] + len(ch)
                flat.extend(ch)
            self._csr = (starts, np.asarray(flat, dtype=np.int32))
        return self._csr

    def parent(self) -> np.ndarray:
        par = np.full(self.n_nodes, -1, dtype=np.int32)
        for i, ch in enumerate(self.children):
            for c in ch:
                par[c] = i
        return par

    def depth(self) -> np.ndarray:
        """Number of edges from the root to each node."""
        par = self.parent()
        d = np.zeros(self.n_nodes, dtype=np.int32)
        for i in range(self.n_nodes - 2, -1, -1):  # parents have larger ids
            d[i] = d[par[i]] + 1
        return d

    def to_newick(self, with_ids: bool = False, digits: int = 10) -> str:
        def rec(i):
            s = ""
            if self.children[i]:
                s = "(" + ",".join(rec(c) for c in self.children[i]) + ")"
            if self.names[i]:
                s += self.names[i]
            if i != self.root:
                s += ":%.*g" % (digits, self.bl[i])
            if with_ids:
                s += "{%d}" % i
            return s

        return rec(self.root) + ";"


_STOP = set("(),:;[]{}")


def parse_newick(text: str) -> Tree:
    """Parse one Newick tree, numbering nodes exactly as pplacer's parser does:
    a leaf takes the next id when it is read, an internal node when its
    closing parenthesis is read (post-order)."""
    s = text.strip()
    pos = 0
    n = len(s)
    children: list = []
    bls: list = []
    names: list = []

    def skip_ws():
        nonlocal pos
        while pos < n and s[pos].isspace():
            pos += 1

    def read_label():
        nonlocal pos
        skip_ws()
        if pos < n and s[pos] == "'":
            end = s.index("'", pos + 1)
            lab = s[pos + 1:end]
            pos = end + 1
            return lab
        start = pos
        while pos < n and s[pos] not in _STOP:
            pos += 1
        lab = s[start:pos].strip()
        return lab if lab else None

    def new_node(ch, name):
        children.append(ch)
        bls.append(None)
        names.append(name)
        return len(children) - 1

    def read_suffix(node):
        nonlocal pos
        skip_ws()
        if pos < n and s[pos] == ":":
            pos += 1
            lab = read_label()
            try:
                bls[node] = float(lab)
            except (TypeError, ValueError):
                raise ValueError("branch lengths must be floats: %r" % lab)
        skip_ws()
        while pos < n and s[pos] in "[{":
            close = "]" if s[pos] == "[" else "}"
            pos = s.index(close, pos) + 1
            skip_ws()

    def subtree():
        nonlocal pos
        skip_ws()
        if pos < n and s[pos] == "(":
            pos += 1
            ch = [subtree()]
            skip_ws()
            while pos < n and s[pos] == ",":
                pos += 1
                ch.append(subtree())
                skip_ws()
            if pos >= n or s[pos] != ")":
                raise ValueError("syntax error parsing Newick at %d" % pos)
            pos += 1
            node = new_node(ch, None)
            lab = read_label()
            if lab:
                names[node] = lab
            read_suffix(node)
            return node
        lab = read_label()
        node = new_node([], lab)
        read_suffix(node)
        return node

    subtree()
    skip_ws()
    if pos < n and s[pos] == ";":
        pos += 1
    skip_ws()
    if pos != n:
        raise ValueError("trailing text after Newick tree at %d" % pos)
    bl = np.array([0.0 if b is None else b for b in bls], dtype=np.float64)
    bl[-1] = 0.0 if bls[-1] is None else bls[-1]  # add_zero_root_bl
    return Tree(children=children, bl=bl, names=names)


def read_newick(path: str) -> Tree:
    with open(path) as f:
        return parse_newick(f.read())


def parse_fasta(text: str):
    """[(name, SEQUENCE)] with sequences upper-cased."""
    out = []
    name = None
    chunks: list = []
    for line in text.splitlines():
        line = line.strip()
        if not line:
            continue
        if line.startswith(">"):
            if name is not None:
                out.append((name, "".join(chunks).upper()))
            name = line[1:].strip()
            chunks = []
        else:
            chunks.append(line)
    if name is not None:
        out.append((name, "".join(chunks).upper()))
    return out


def read_fasta(path: str):
    with open(path) as f:
        return parse_fasta(f.read())


# A Stockholm line is blank, the header, the footer, markup (#=...), or an
# alignment row "name<blanks>sequence" (stockholm.ml:9-26).
_STO_ROW = re.compile(r"^([^ \t\n\r#]+)[ \t]+([^ \t\n\r#]+)[ \t]*$")


def parse_stockholm(text: str):
    """Stockholm.parse (pplacer_src/stockholm.ml:36-82): the rows of each name
    concatenated in file order, '.' read as '-', every sequence of one
    length, and the records listed in descending name order (a fold over a
    string map that conses each binding).  Sequences are upper-cased as
    Alignment.upper_aln_of_any_file does."""
    rows: dict = {}
    state = "header"
    for lineno, raw in enumerate(text.splitlines(), 1):
        line = raw.rstrip(" \t\r")
        if not line:
            continue
        if line == "# STOCKHOLM 1.0":
            if state != "header":
                raise ValueError("Stockholm line %d: header unexpected" % lineno)
            state = "body"
            continue
        if state == "header":
            raise ValueError("Stockholm line %d: something before header" % lineno)
        if state == "done":
            raise ValueError("Stockholm line %d: something after footer" % lineno)
        if line == "//":
            state = "done"
            continue
        if line.startswith("#="):
            continue
        m = _STO_ROW.match(line)
        if not m:
            raise ValueError("Stockholm line %d: not an alignment row: %r" % (lineno, line))
        rows.setdefault(m.group(1), []).append(m.group(2).replace(".", "-"))
    if state != "done":
        raise ValueError("Stockholm: didn't reach footer by EOF")
    out = [(name, "".join(parts).upper()) for name, parts in
           sorted(rows.items(), key=lambda kv: kv[0].encode(), reverse=True)]
    lengths = {len(seq) for _, seq in out}
    if len(lengths) > 1:
        raise ValueError("Stockholm: sequences of different lengths %s" % sorted(lengths))
    return out


def _read_text(path: str) -> str:
    """A file's text, gunzipped when its name ends in .gz (MaybeZipped)."""
    if path.endswith(".gz"):
        with gzip.open(path, "rt") as f:
            return f.read()
    with open(path) as f:
        return f.read()


def parse_alignment(name: str, text: str):
    """Alignment.list_of_any_file by the file name's suffix (a .gz suffix is
    looked through): .fasta / .fa are FASTA, .sth / .sto Stockholm; records
    upper-cased."""
    base = name[:-3] if name.endswith(".gz") else name
    if base.endswith((".fasta", ".fa")):
        return parse_fasta(text)
    if base.endswith((".sth", ".sto")):
        return parse_stockholm(text)
    raise ValueError("unfamiliar suffix on " + name)


def read_alignment(path: str):
    return parse_alignment(path, _read_text(path))


def write_fasta(path: str, records) -> None:
    with open(path, "w") as f:
        for name, seq in records:
            f.write(">%s\n%s\n" % (name, seq))


@dataclass
class PhyloModel:
    """refpkg phylo_model contents: the gamma mixture model
    (gmix_model.ml:386-413) or the Price-CAT model (gcat_model.ml:453-478)."""

    subs_model: str
    empirical_frequencies: bool
    alpha: float | None
    n_cats: int
    subs_rates: list | None  # ac ag at cg ct gt for GTR
    ras_model: str = "gamma"
    cat_rates: list | None = None  # Price-CAT "Rates"
    site_categories: list | None = None  # Price-CAT "SiteCategories", shifted to 0-based


def parse_phylo_model(obj) -> PhyloModel:
    """The ras_model dispatch of Refpkg.get_model (refpkg.ml:38-63) over
    Gmix_model.init_of_json (gmix_model.ml:386-413) and
    Gcat_model.init_of_json (gcat_model.ml:453-478)."""
    if isinstance(obj, str):
        obj = json.loads(obj)
    ras = obj.get("ras_model")
    if ras not in ("gamma", "Price-CAT"):
        raise ValueError("invalid ras_model: %s" % ras)
    rates = None
    if "subs_rates" in obj:
        sr = obj["subs_rates"]
        rates = [float(sr[k]) for k in ("ac", "ag", "at", "cg", "ct", "gt")]
    common = dict(subs_model=obj["subs_model"],
                  empirical_frequencies=bool(obj["empirical_frequencies"]), subs_rates=rates)
    if ras == "gamma":
        return PhyloModel(alpha=float(obj["gamma"]["alpha"]), n_cats=int(obj["gamma"]["n_cats"]),
                          **common)
    pc = obj["Price-CAT"]
    cat_rates = [float(x) for x in pc["Rates"]]
    # FastTree writes 1-based categories (gcat_model.ml:475)
    cats = [int(x) - 1 for x in pc["SiteCategories"]]
    return PhyloModel(alpha=None, n_cats=len(cat_rates), ras_model="Price-CAT",
                      cat_rates=cat_rates, site_categories=cats, **common)


def read_phylo_model(path: str) -> PhyloModel:
    with open(path) as f:
        return parse_phylo_model(json.load(f))


def gtr_exchangeabilities(rates) -> np.ndarray:
    """Symmetric 4x4 from ac ag at cg ct gt with zero diagonal
    (pplacer_src/nuc_models.ml:64-84)."""
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    m = np.zeros((4, 4))
    for (i, j), v in zip(pairs, rates):
        m[i, j] = m[j, i] = v
    return m


def exchangeabilities_from_lower(rows) -> np.ndarray:
    """Symmetric matrix from lower-triangular rows (prot_models.ml:57-68)."""
    n = len(rows) + 1
    m = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                m[i, j] = rows[max(i, j) - 1][min(i, j)]
    return m


# ---- built-in protein models (prot_models.ml:70-163) ------------------------

def prot_model(name: str):
    """Prot_models.trans_and_statd_of_model_name: (exchangeabilities [20][20],
    model frequencies [20]) of LG, WAG or JTT."""
    with open(os.path.join(DATA, "prot_models.json")) as f:
        tables = json.load(f)
    if name not in tables:
        raise ValueError("unknown model: " + name)
    t = tables[name]
    return exchangeabilities_from_lower(t["lower"]), np.asarray(t["statd"], dtype=np.float64)


def model_matrices(pm: PhyloModel, ref_seqs, emper_freq):
    """Gstar_support.seqtype_and_trans_statd_of_info (gstar_support.ml:81-98):
    (n_states, exchangeabilities, stationary frequencies) of a phylo_model;
    emper_freq(n_states, seqs) gives the alignment's empirical frequencies.
    GTR takes its subs_rates and always the alignment's frequencies; a protein
    model its built-in table and, unless empirical_frequencies, its own
    frequencies."""
    if pm.subs_model == "GTR":
        if pm.subs_rates is None:
            raise ValueError("GTR specified but no substitution rates given.")
        return 4, gtr_exchangeabilities(pm.subs_rates), emper_freq(4, ref_seqs)
    exch, statd = prot_model(pm.subs_model)
    return 20, exch, (emper_freq(20, ref_seqs) if pm.empirical_frequencies else statd)


# ---- reference packages (refpkg_parse.ml, refpkg.ml) -------------------------

class Refpkg:
    """A reference package: a directory, or a zip file holding exactly one
    directory, with a CONTENTS.json whose "files" object maps keys (tree,
    aln_fasta, aln_sto, phylo_model, ...) to file names inside it
    (Refpkg_parse.strmap_of_dir / strmap_of_zipfile).  Entries given in
    `override` (key -> path) replace the package's, as pplacer's -t / -r do
    (pplacer_run.ml:701-719)."""

    def __init__(self, path: str | None = None, override: dict | None = None):
        self._zip = None
        self.files: dict = {}
        self.name = None
        self.format_version = None
        if path is not None:
            if os.path.isdir(path):
                root = path.rstrip("/")
                contents = json.load(open(os.path.join(root, "CONTENTS.json")))
                self.files = {k: os.path.join(root, v) for k, v in contents["files"].items()}
            else:
                self._zip = zipfile.ZipFile(path)
                dirs = {n.split("/")[0] + "/" for n in self._zip.namelist() if "/" in n}
                if len(dirs) != 1:
                    raise ValueError("reference package zipfiles must contain exactly one directory")
                d = dirs.pop()
                contents = json.loads(self._zip.read(d + "CONTENTS.json"))
                self.files = {k: "zip:" + d + v for k, v in contents["files"].items()}
            self.format_version = contents.get("metadata", {}).get("format_version")
            base = os.path.basename(path.rstrip("/"))
            self.name = base[:-len(".refpkg")] if base.endswith(".refpkg") else base
        for k, v in (override or {}).items():
            if v:
                self.files[k] = v

    def _text(self, key: str):
        p = self.files[key]
        if p.startswith("zip:"):
            data = self._zip.read(p[4:])
            if p.endswith(".gz"):
                data = gzip.decompress(data)
            return p[4:], data.decode()
        return p, _read_text(p)

    def has(self, key: str) -> bool:
        return key in self.files

    def tree(self) -> Tree:
        """Refpkg.get_ref_tree."""
        if "tree" not in self.files:
            raise ValueError("the reference package provided does not contain a tree")
        return parse_newick(self._text("tree")[1])

    def alignment(self):
        """Refpkg.get_aln_fasta: aln_fasta, else aln_sto; upper-cased."""
        if "aln_fasta" in self.files:
            return parse_fasta(self._text("aln_fasta")[1])
        if "aln_sto" in self.files:
            return parse_stockholm(self._text("aln_sto")[1])
        raise ValueError("the reference package provided does not contain an alignment")

    def phylo_model(self) -> PhyloModel:
        """Refpkg.get_model's phylo_model (refpkg.ml:39-63)."""
        if "phylo_model" not in self.files:
            raise ValueError("the reference package provided does not contain a phylo_model "
                             "(statistics files are not read)")
        return parse_phylo_model(self._text("phylo_model")[1])
