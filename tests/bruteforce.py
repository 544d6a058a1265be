"""Independent likelihood checks (test helper).

A textbook Felsenstein pruning written directly from the model definition:
Q_ij = B_ij * pi_j off the diagonal, rows summing to zero, scaled so that
-sum_i pi_i Q_ii = 1, P_r(t) = expm(Q * rate_r * t) from scipy (Pade), no
eigendecomposition and no exponent pulling.  It shares no code with the
oracle's diagd / like_stree / three_tax path, so it pins those by a different
route (and is only used on small trees where underflow cannot occur).
"""
import numpy as np
import scipy.linalg as sl

NUC_BITS = {"A": 1, "C": 2, "G": 4, "T": 8, "R": 5, "Y": 10, "M": 3, "K": 12, "S": 6,
            "W": 9, "H": 11, "B": 14, "V": 7, "D": 13, "N": 15, "-": 15, "?": 15, "X": 15}
PROT = "ARNDCQEGHILKMFPSTWYV"


def lv(c, S):
    if S == 4:
        b = NUC_BITS[c]
        return np.array([(b >> i) & 1 for i in range(4)], dtype=float)
    v = np.zeros(20)
    if c in PROT:
        v[PROT.index(c)] = 1
    elif c == "B":
        v[[2, 3]] = 1
    elif c == "J":
        v[[9, 10]] = 1
    elif c == "Z":
        v[[5, 6]] = 1
    else:
        v[:] = 1
    return v


def rate_matrix(exch, statd):
    S = len(statd)
    Q = np.array(exch, dtype=float) * np.asarray(statd)[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mu = -np.sum(np.asarray(statd) * np.diag(Q))
    return Q / mu


def transition(Q, t):
    return sl.expm(Q * t)


class Node:
    def __init__(self, children=(), bl=0.0, seq=None):
        self.children = list(children)
        self.bl = bl
        self.seq = seq


def from_tree(tree, leaf_seqs):
    nodes = [None] * tree.n_nodes
    for i in range(tree.n_nodes):
        nodes[i] = Node([], tree.bl[i], leaf_seqs.get(i))
    for i, ch in enumerate(tree.children):
        nodes[i].children = [nodes[c] for c in ch]
    return nodes


def attach(tree, leaf_seqs, loc, dist, pend, query):
    """Copy of the tree with the query hung off edge `loc`, `dist` above the
    node `loc` and `pend` out, as a three-taxon insertion."""
    nodes = from_tree(tree, leaf_seqs)
    par = tree.parent()
    below = nodes[loc]
    cut = tree.bl[loc]
    q = Node([], pend, query)
    x = Node([below, q], cut - dist)
    below.bl = dist
    p = nodes[par[loc]]
    p.children = [x if c is below else c for c in p.children]
    return nodes[tree.root]


def site_loglikes(root, exch, statd, rates, sites=None):
    Q = rate_matrix(exch, statd)
    S = len(statd)
    R = len(rates)
    n = len(_first_seq(root))
    sites = np.arange(n) if sites is None else np.asarray(sites)

    def partial(node, r):
        if not node.children:
            return np.array([lv(node.seq[k], S) for k in sites])
        out = np.ones((len(sites), S))
        for c in node.children:
            Pm = transition(Q, rates[r] * c.bl)
            out = out * (partial(c, r) @ Pm.T)
        return out

    tot = np.zeros(len(sites))
    for r in range(R):
        tot += partial(root, r) @ np.asarray(statd)
    return np.log(tot / R)


def _first_seq(node):
    while node.children:
        node = node.children[0]
    return node.seq


def tree_loglike(tree, leaf_seqs, exch, statd, rates):
    return float(np.sum(site_loglikes(from_tree(tree, leaf_seqs)[tree.root], exch, statd, rates)))


def placement_loglike(tree, leaf_seqs, exch, statd, rates, loc, dist, pend, query):
    """Log-likelihood over the query's informative sites of the tree with the
    query attached on edge `loc` (what three_tax.ml:30-36 computes)."""
    sites = [k for k, c in enumerate(query) if c not in "-?"]
    root = attach(tree, leaf_seqs, loc, dist, pend, query)
    return float(np.sum(site_loglikes(root, exch, statd, rates, sites)))


def site_loglikes_cat(root, exch, statd, cat_rates, site_cat, sites=None):
    """Price-CAT site log-likelihoods: site k evolves at rate
    cat_rates[site_cat[k]] alone (no mixture, no 1/R)."""
    site_cat = np.asarray(site_cat)
    sites = np.arange(len(site_cat)) if sites is None else np.asarray(sites)
    out = np.zeros(len(sites))
    for c in np.unique(site_cat[sites]):
        pick = np.nonzero(site_cat[sites] == c)[0]
        out[pick] = site_loglikes(root, exch, statd, [cat_rates[c]], sites[pick])
    return out


def tree_loglike_cat(tree, leaf_seqs, exch, statd, cat_rates, site_cat):
    root = from_tree(tree, leaf_seqs)[tree.root]
    return float(np.sum(site_loglikes_cat(root, exch, statd, cat_rates, site_cat)))


def placement_loglike_cat(tree, leaf_seqs, exch, statd, cat_rates, site_cat, loc, dist, pend,
                          query):
    sites = [k for k, c in enumerate(query) if c not in "-?"]
    root = attach(tree, leaf_seqs, loc, dist, pend, query)
    return float(np.sum(site_loglikes_cat(root, exch, statd, cat_rates, site_cat, sites)))
