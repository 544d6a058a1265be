"""Deterministic synthetic workloads for the placement path.

One seeded generator covers every workload shape the benchmarks and parity
tests use (nucleotide GTR+Gamma4 and 20-state protein models, Yule topologies,
i.i.d. exponential branch lengths, sequences simulated down the tree, queries
simulated from random edge midpoints and cut to contiguous fragments with the
rest gaps).  Sequences are simulated with continuous per-site Gamma(alpha)
rates; the placement model then uses pplacer's discrete Gamma.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import refio

NUC = "ACGT"
PROT = "ARNDCQEGHILKMFPSTWYV"

# name: (states, taxa, mean bl, columns, queries, frag lo, frag hi, query mode, query param)
WORKLOADS = {
    "gtr50": (4, 50, 0.05, 1000, 200, 150, 300, "pendant", 0.02),
    "gtr2000": (4, 2000, 0.03, 1600, 100_000, 100, 500, "subst", 0.02),
    "prot500": (20, 500, 0.05, 400, 20_000, 60, 400, "subst", 0.02),
    "gtr10000": (4, 10_000, 0.03, 1500, 1_000_000, 80, 600, "subst", 0.02),
    # small shapes for unit tests
    "gtr12": (4, 12, 0.08, 120, 24, 30, 100, "pendant", 0.02),
    "prot10": (20, 10, 0.1, 80, 12, 20, 70, "pendant", 0.02),
}


@dataclass
class Workload:
    name: str
    seed: int
    n_states: int
    tree: refio.Tree
    ref_aln: list  # [(name, seq)] for leaves
    queries: list  # [(name, seq)]
    exch: np.ndarray  # [S][S] symmetric, zero diagonal
    statd: np.ndarray  # [S] generating frequencies
    alpha: float
    n_cats: int
    true_locs: np.ndarray  # edge each query was simulated from

    def leaf_seqs(self) -> dict:
        m = dict(self.ref_aln)
        return {i: m[self.tree.names[i]] for i in self.tree.leaves()}

    def phylo_model_json(self) -> dict:
        d = {"ras_model": "gamma", "gamma": {"alpha": self.alpha, "n_cats": self.n_cats},
             "empirical_frequencies": True}
        if self.n_states == 4:
            r = self.exch
            d.update(subs_model="GTR", datatype="DNA",
                     subs_rates={k: float(r[i, j]) for k, (i, j) in zip(
                         ("ac", "ag", "at", "cg", "ct", "gt"),
                         ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)))})
        else:
            d.update(subs_model="custom", datatype="AA")
        return d


def _yule_newick(rng, n_taxa, mean_bl):
    """Random Yule topology on n_taxa leaves with a trifurcating root and
    i.i.d. Exp(mean_bl) branch lengths, as a Newick string."""
    # nodes: children lists; start with root having three leaf children
    children = {0: [1, 2, 3]}
    leaves = [1, 2, 3]
    nxt = 4
    while len(leaves) < n_taxa:
        k = int(rng.integers(len(leaves)))
        leaf = leaves[k]
        children[leaf] = [nxt, nxt + 1]
        leaves[k] = nxt
        leaves.append(nxt + 1)
        nxt += 2
    names = {}
    for i, leaf in enumerate(sorted(leaves)):
        names[leaf] = "t%d" % i

    def rec(v):
        if v in children:
            s = "(" + ",".join(rec(c) for c in children[v]) + ")"
        else:
            s = names[v]
        if v != 0:
            s += ":%.10f" % float(rng.exponential(mean_bl) + 1e-6)
        return s

    return rec(0) + ";"


def _eig(exch, statd):
    """Eigendecomposition of the normalised reversible Q for simulation."""
    Q = exch * statd[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    Q /= -np.sum(statd * np.diag(Q))
    dr = np.sqrt(statd)
    A = (dr[:, None] * Q) / dr[None, :]
    A = 0.5 * (A + A.T)
    lam, U = np.linalg.eigh(A)
    X = U / dr[:, None]
    Xi = U.T * dr[None, :]
    return lam, X, Xi


def _evolve_states(rng, states, t_site, lam, X, Xi):
    """Sample child states given parent states and per-site times t_site."""
    # P[site] = X diag(exp(lam t)) Xi, row for the parent state only
    e = np.exp(lam[None, :] * t_site[:, None])  # [N, S]
    rows = X[states] * e  # [N, S]
    P = rows @ Xi  # [N, S]
    P = np.clip(P, 0.0, None)
    P /= P.sum(axis=1, keepdims=True)
    u = rng.random(len(states))
    c = np.cumsum(P, axis=1)
    return np.minimum((u[:, None] > c).sum(axis=1), P.shape[1] - 1)


def _queries_batched(rng, st, tree, site_rate, lam, X, Xi, abet, locs, lens, starts, qpar, S,
                     ncol, batch=8192):
    """Substitution-mode queries, a batch at a time: the mid-edge states of
    every column of a batch's fragments in one _evolve_states call, then
    i.i.d. substitutions (a different random stream from the per-query loop,
    so the same seed gives different queries)."""
    out = []
    gap = ord("-")
    nq = len(locs)
    for b0 in range(0, nq, batch):
        b1 = min(nq, b0 + batch)
        v = locs[b0:b1]
        ln = lens[b0:b1]
        a = starts[b0:b1]
        qid = np.repeat(np.arange(b1 - b0), ln)
        col = np.concatenate([np.arange(x, x + y) for x, y in zip(a, ln)])
        parent = st[v[qid], col]
        t = site_rate[col] * (tree.bl[v[qid]] / 2.0)
        mid = _evolve_states(rng, parent, t, lam, X, Xi)
        flip = rng.random(len(mid)) < qpar
        mid[flip] = (mid[flip] + rng.integers(1, S, size=int(flip.sum()))) % S
        rows = np.full((b1 - b0, ncol), gap, dtype=np.uint8)
        rows[qid, col] = abet[mid]
        out.extend(("q%d" % (b0 + k), rows[k].tobytes().decode()) for k in range(b1 - b0))
    return out


def make_workload(name: str, seed: int = 1, n_queries: int | None = None,
                  n_taxa: int | None = None, n_columns: int | None = None,
                  batched_queries: bool = False) -> Workload:
    """The named workload; batched_queries generates substitution-mode
    queries a batch at a time (for streams of a million queries)."""
    S, taxa, mean_bl, ncol, nq, flo, fhi, qmode, qpar = WORKLOADS[name]
    taxa = n_taxa or taxa
    ncol = n_columns or ncol
    nq = nq if n_queries is None else n_queries
    rng = np.random.Generator(np.random.PCG64(seed))
    alphabet = NUC if S == 4 else PROT
    tree = refio.parse_newick(_yule_newick(rng, taxa, mean_bl))
    G = tree.n_nodes
    # model
    exch = np.zeros((S, S))
    iu = np.triu_indices(S, 1)
    exch[iu] = rng.exponential(1.0, size=len(iu[0]))
    exch = exch + exch.T
    statd = rng.dirichlet(np.ones(S))
    statd = 0.9 * statd + 0.1 / S  # keep every state reachable
    alpha = 0.5
    lam, X, Xi = _eig(exch, statd)
    site_rate = rng.gamma(alpha, 1.0 / alpha, size=ncol)
    # simulate states at every node (root first: ids descend from the root)
    st = np.empty((G, ncol), dtype=np.int64)
    st[tree.root] = rng.choice(S, size=ncol, p=statd)
    par = tree.parent()
    for v in range(G - 2, -1, -1):
        st[v] = _evolve_states(rng, st[par[v]], site_rate * tree.bl[v], lam, X, Xi)
    abet = np.frombuffer(alphabet.encode(), dtype=np.uint8)
    ref_aln = [(tree.names[v], abet[st[v]].tobytes().decode()) for v in tree.leaves()]
    # queries
    locs = rng.integers(0, G - 1, size=nq)
    lens = rng.integers(flo, min(fhi, ncol) + 1, size=nq)
    starts = (rng.random(nq) * (ncol - lens + 1)).astype(np.int64)
    queries = []
    gap = ord("-")
    if batched_queries and qmode == "subst":
        queries = _queries_batched(rng, st, tree, site_rate, lam, X, Xi, abet, locs, lens, starts,
                                   qpar, S, ncol)
        nq = 0  # skip the per-query loop
    for k in range(nq):
        v = int(locs[k])
        a, b = int(starts[k]), int(starts[k] + lens[k])
        sr = site_rate[a:b]
        mid = _evolve_states(rng, st[v, a:b], sr * (tree.bl[v] / 2.0), lam, X, Xi)
        if qmode == "pendant":
            q = _evolve_states(rng, mid, sr * rng.exponential(qpar), lam, X, Xi)
        else:
            q = mid.copy()
            flip = rng.random(b - a) < qpar
            q[flip] = (q[flip] + rng.integers(1, S, size=int(flip.sum()))) % S
        row = np.full(ncol, gap, dtype=np.uint8)
        row[a:b] = abet[q]
        queries.append(("q%d" % k, row.tobytes().decode()))
    return Workload(name=name, seed=seed, n_states=S, tree=tree, ref_aln=ref_aln,
                    queries=queries, exch=exch, statd=statd, alpha=alpha, n_cats=4,
                    true_locs=locs)
