"""Host-side readers for the inputs of the placement path.

* Newick trees with pplacer's edge numbering: every node gets the id of its
  position in a post-order walk of the parse, the root gets the largest id and
  a zero branch length (pplacer_src/newick_parser.mly:49-62,104-108,
  pplacer_src/newick_gtree.ml:145-157).
* FASTA alignments, upper-cased (common_src/alignment.ml, ``upper_aln_of_any_file``).
* refpkg ``phylo_model`` JSON for the gamma mixture model
  (pplacer_src/gmix_model.ml:386-413).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np


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
