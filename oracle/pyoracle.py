"""ctypes view of the C oracle (oracle/pp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / the timed CPU baseline.  The
product path (pplacer_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libpporacle.so")

OR_OK = 0
OR_GSL_EBADFUNC = 1
OR_GSL_EINVAL = 2
OR_EXCEEDED_MAX_ITER = 3

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)
_ip = C.POINTER(C.c_int)


class OrModel(C.Structure):
    _fields_ = [("n_rates", C.c_int), ("n_states", C.c_int), ("statd", _dp), ("x", _dp),
                ("l", _dp), ("xit", _dp), ("rates", _dp), ("n_cats", C.c_int),
                ("cat_rates", _dp), ("site_cat", _ip)]


class OrTree(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("child_start", _ip), ("children", _ip), ("bl", _dp)]


class OrPrefs(C.Structure):
    _fields_ = [("start_pend", C.c_double), ("max_pend", C.c_double),
                ("initial_tolerance", C.c_double), ("strike_box", C.c_double),
                ("keep_factor", C.c_double), ("max_strikes", C.c_int),
                ("max_pitches", C.c_int), ("keep_at_most", C.c_int), ("calc_pp", C.c_int),
                ("pp_rel_err", C.c_double), ("prior_kind", C.c_int), ("prior_mean", C.c_double),
                ("prior_means", _dp)]


class OrCore(C.Structure):
    _fields_ = [("model", C.POINTER(OrModel)), ("tree", C.POINTER(OrTree)),
                ("n_sites", C.c_int), ("darr", _dp), ("darr_e", _i64p), ("parr", _dp),
                ("parr_e", _i64p), ("snodes", _dp), ("snodes_e", _i64p), ("prefs", OrPrefs)]


class OrPlacement(C.Structure):
    _fields_ = [("loc", C.c_int), ("ml_ratio", C.c_double), ("log_like", C.c_double),
                ("dist_bl", C.c_double), ("pend_bl", C.c_double), ("post_prob", C.c_double),
                ("marginal_prob", C.c_double)]


class OrStats(C.Structure):
    _fields_ = [("n_logdots", C.c_long), ("n_like_calls", C.c_long), ("n_pitched", C.c_int),
                ("n_skipped", C.c_int), ("n_max_iter", C.c_int), ("t_ranking", C.c_double),
                ("t_ml", C.c_double)]


OR_FUN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class OrGslBrent(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x_lower", "x_upper", "x_minimum", "f_lower", "f_upper",
                                          "f_minimum", "v", "w", "d", "e", "f_v", "f_w")]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, v = C.c_double, C.c_int, C.c_void_p
        sig = {
            "or_gemmish": (None, [_dp, _dp, _dp, i, i]),
            "or_dediagonalize": (None, [_dp, _dp, _dp, _dp, i]),
            "or_ten_log_like3": (d, [_dp, _dp, _dp, _dp, _dp, i, i, i]),
            "or_ten_pairwise_prod": (None, [_dp, _dp, _dp, C.c_long]),
            "or_ten_statd_pairwise_prod": (None, [_dp, _dp, _dp, _dp, i, i, i]),
            "or_ten_masked_logdot": (d, [_dp, _dp, _u16p, _dp, i, i, i]),
            "or_ten_bounded_logdot": (d, [_dp, _dp, i, i, _dp, i, i, i]),
            "or_mat_masked_logdot": (d, [_dp, _dp, _u16p, i, i]),
            "or_mat_log_like3": (d, [_dp, _dp, _dp, _dp, i, i]),
            "or_mat_bounded_logdot": (d, [_dp, _dp, i, i, i]),
            "or_get_twoexp": (i, [d]),
            "or_of_twoexp": (d, [i]),
            "or_total_twoexp": (d, [_i64p, i]),
            "or_masked_total_twoexp": (d, [_i64p, _u16p, i]),
            "or_perhaps_pull_exponent": (None, [i, _dp, _i64p, i, i, i]),
            "or_gamma_inc_P": (d, [d, d]),
            "or_gamma_Pinv": (d, [d, d]),
            "or_discrete_gamma": (i, [i, d, _dp]),
            "or_symm_eigs": (i, [_dp, i, _dp, _dp]),
            "or_b_of_exchangeable_pair": (None, [_dp, _dp, i, _dp]),
            "or_of_d_b": (i, [_dp, _dp, i, _dp, _dp, _dp]),
            "or_normed_of_exchangeable_pair": (i, [_dp, _dp, i, _dp, _dp, _dp]),
            "or_to_exp": (None, [_dp, _dp, _dp, _dp, i, d]),
            "or_multi_exp": (None, [_dp, _dp, _dp, _dp, _dp, i, i, d]),
            "or_lv_of_char": (i, [i, C.c_char, _dp]),
            "or_mask_reindex": (None, [_u8p, i, _ip, _ip]),
            "or_emper_freq": (i, [i, C.c_char_p, i, i, _dp]),
            "or_premask": (i, [i, C.c_char_p, i, C.c_char_p, i, i, i, _u8p, _u8p, _ip]),
            "or_evolve_into": (None, [C.POINTER(OrModel), _dp, _i64p, _dp, _i64p, i, d]),
            "or_log_like3": (d, [C.POINTER(OrModel), _dp, i, _dp, _i64p, _dp, _i64p, _dp, _i64p]),
            "or_masked_logdot": (d, [_dp, i, i, i, _dp, _i64p, _dp, _i64p, _u16p]),
            "or_logdot": (d, [_dp, i, i, i, _dp, _i64p, _dp, _i64p]),
            "or_slow_log_like3": (d, [C.POINTER(OrModel), i, _dp, _i64p, _dp, _i64p, _dp, _i64p]),
            "or_calc_distal_and_proximal": (i, [C.POINTER(OrModel), C.POINTER(OrTree), C.c_char_p,
                                                i, _dp, _i64p, _dp, _i64p, _dp, _i64p]),
            "or_pull_exponents_arr": (None, [i, _dp, _i64p, i, i, i, i]),
            "or_prep_supernodes": (None, [C.POINTER(OrModel), C.POINTER(OrTree), i, _dp, _i64p,
                                          _dp, _i64p, _dp, _i64p]),
            "or_brent": (i, [OR_FUN, v, d, d, d, d, i, _dp]),
            "or_gsl_brent_set": (i, [C.POINTER(OrGslBrent), OR_FUN, v, d, d, d]),
            "or_gsl_brent_iterate": (i, [C.POINTER(OrGslBrent), OR_FUN, v]),
            "or_stage1_scores": (i, [C.POINTER(OrCore), C.c_char_p, _dp]),
            "or_rank": (i, [_dp, i, _ip]),
            "or_three_tax_optimize": (i, [C.POINTER(OrCore), C.c_char_p, i, d, d, d, _dp, _dp,
                                          _dp, C.POINTER(C.c_long)]),
            "or_process_query": (i, [C.POINTER(OrCore), C.c_char_p, C.POINTER(OrPlacement), _ip,
                                     C.POINTER(OrStats)]),
            "or_process_queries": (i, [C.POINTER(OrCore), C.c_char_p, i, i,
                                       C.POINTER(OrPlacement), _ip, _ip, C.POINTER(OrStats)]),
            "or_ll_normalized_prob": (None, [_dp, i, _dp]),
            "or_qng": (i, [OR_FUN, v, d, d, d, d, _dp, _dp, _ip, _ip]),
            "or_exponential_pdf": (d, [d, d]),
            "or_three_tax_log_like": (i, [C.POINTER(OrCore), C.c_char_p, i, d, d, _dp]),
            "or_calc_marg_prob": (i, [C.POINTER(OrCore), C.c_char_p, i, d, d, _dp,
                                      C.POINTER(C.c_long)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def P(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# linear_c.c twins


def gemmish(a, b):
    a, b = f64(a), f64(b)
    n_sites, S = b.shape
    dst = np.empty((n_sites, S))
    lib().or_gemmish(P(dst), P(a), P(b), n_sites, S)
    return dst


def dediagonalize(u, lam, uit):
    u, lam, uit = f64(u), f64(lam), f64(uit)
    n = lam.shape[0]
    dst = np.empty((n, n))
    lib().or_dediagonalize(P(dst), P(u), P(lam), P(uit), n)
    return dst


def ten_log_like3(statd, x, y, z):
    statd, x, y, z = f64(statd), f64(x), f64(y), f64(z)
    R, M, S = x.shape
    util = np.empty(max(M, 1))
    return lib().or_ten_log_like3(P(statd), P(x), P(y), P(z), P(util), R, M, S)


def ten_masked_logdot(x, y, mask):
    x, y = f64(x), f64(y)
    mask = np.ascontiguousarray(mask, dtype=np.uint16)
    R, N, S = x.shape
    util = np.empty(max(N, 1))
    return lib().or_ten_masked_logdot(P(x), P(y), P(mask, C.c_uint16), P(util), R, N, S)


def ten_bounded_logdot(x, y, first, last):
    x, y = f64(x), f64(y)
    R, N, S = x.shape
    util = np.empty(max(N, 1))
    return lib().or_ten_bounded_logdot(P(x), P(y), first, last, P(util), R, N, S)


def mat_masked_logdot(x, y, mask):
    x, y = f64(x), f64(y)
    mask = np.ascontiguousarray(mask, dtype=np.uint16)
    N, S = x.shape
    return lib().or_mat_masked_logdot(P(x), P(y), P(mask, C.c_uint16), N, S)


def mat_log_like3(statd, x, y, z):
    x = f64(x)
    N, S = x.shape
    return lib().or_mat_log_like3(P(f64(statd)), P(x), P(f64(y)), P(f64(z)), N, S)


def mat_bounded_logdot(x, y, first, last):
    x = f64(x)
    return lib().or_mat_bounded_logdot(P(x), P(f64(y)), first, last, x.shape[1])


def ten_statd_pairwise_prod(statd, a, b):
    statd, a, b = f64(statd), f64(a), f64(b)
    dst = np.empty_like(a)
    R, N, S = a.shape
    lib().or_ten_statd_pairwise_prod(P(statd), P(dst), P(a), P(b), R, N, S)
    return dst


def ten_pairwise_prod(x, y):
    x, y = f64(x), f64(y)
    dst = np.empty_like(x)
    lib().or_ten_pairwise_prod(P(dst), P(x), P(y), x.size)
    return dst


def perhaps_pull_exponent(min_twoexp, a, e):
    a = f64(a).copy()
    e = np.ascontiguousarray(e, dtype=np.int64).copy()
    R, N, S = a.shape
    lib().or_perhaps_pull_exponent(min_twoexp, P(a), P(e, C.c_int64), R, N, S)
    return a, e


# --------------------------------------------------------------------------
# model building


def discrete_gamma(n_cat, alpha):
    out = np.empty(n_cat)
    st = lib().or_discrete_gamma(n_cat, alpha, P(out))
    assert st == OR_OK
    return out


def symm_eigs(m):
    m = f64(m)
    n = m.shape[0]
    l, u = np.empty(n), np.empty((n, n))
    lib().or_symm_eigs(P(m), n, P(l), P(u))
    return l, u


def normed_of_exchangeable_pair(r, pi):
    r, pi = f64(r), f64(pi)
    n = pi.shape[0]
    x, l, xit = np.empty((n, n)), np.empty(n), np.empty((n, n))
    st = lib().or_normed_of_exchangeable_pair(P(r), P(pi), n, P(x), P(l), P(xit))
    assert st == OR_OK
    return x, l, xit


def of_d_b(d, b):
    d, b = f64(d), f64(b)
    n = d.shape[0]
    x, l, xit = np.empty((n, n)), np.empty(n), np.empty((n, n))
    st = lib().or_of_d_b(P(d), P(b), n, P(x), P(l), P(xit))
    assert st == OR_OK
    return x, l, xit


def to_exp(x, l, xit, bl):
    x, l, xit = f64(x), f64(l), f64(xit)
    n = l.shape[0]
    dst = np.empty((n, n))
    lib().or_to_exp(P(dst), P(x), P(l), P(xit), n, bl)
    return dst


def multi_exp(x, l, xit, rates, bl):
    x, l, xit, rates = f64(x), f64(l), f64(xit), f64(rates)
    n, R = l.shape[0], rates.shape[0]
    dst = np.empty((R, n, n))
    lib().or_multi_exp(P(dst), P(x), P(l), P(xit), P(rates), R, n, bl)
    return dst


def emper_freq(n_states, seqs):
    n_sites = len(seqs[0])
    buf = "".join(seqs).encode()
    out = np.empty(n_states)
    st = lib().or_emper_freq(n_states, buf, len(seqs), n_sites, P(out))
    assert st == OR_OK
    return out


def lv_of_char(n_states, c):
    out = np.empty(n_states)
    st = lib().or_lv_of_char(n_states, c.encode(), P(out))
    if st != OR_OK:
        raise ValueError("%r not a known character" % c)
    return out


def mask_reindex(mask):
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty(max(len(mask), 1), dtype=np.int32)
    n = C.c_int(0)
    lib().or_mask_reindex(P(mask, C.c_uint8), len(mask), P(out, C.c_int), C.byref(n))
    return out[: n.value]


def premask(n_states, ref_seqs, query_seqs, discard_nonoverlapped=False):
    n_sites = len(ref_seqs[0])
    mask = np.zeros(n_sites, dtype=np.uint8)
    keep = np.zeros(max(len(query_seqs), 1), dtype=np.uint8)
    ml = C.c_int(0)
    st = lib().or_premask(n_states, "".join(ref_seqs).encode(), len(ref_seqs),
                          "".join(query_seqs).encode(), len(query_seqs), n_sites,
                          int(discard_nonoverlapped), P(mask, C.c_uint8), P(keep, C.c_uint8),
                          C.byref(ml))
    return st, mask.astype(bool), keep[: len(query_seqs)].astype(bool), ml.value


# --------------------------------------------------------------------------
# model + tree objects


class Model:
    """Gamma mixture model: statd, rates and the diagonalised Q (x, l, xit).
    With cat_rates and site_cat it is the Price-CAT model (gcat_model.ml):
    rate-free GLVs, site i evolving at rate cat_rates[site_cat[i]] (site_cat
    indexes the pre-masked sites)."""

    def __init__(self, exch, statd, rates=None, cat_rates=None, site_cat=None):
        self.exch = f64(exch)
        self.statd = f64(statd)
        self.x, self.l, self.xit = normed_of_exchangeable_pair(self.exch, self.statd)
        self.S = self.statd.shape[0]
        if cat_rates is not None:
            self.rates = f64([1.0])
            self.cat_rates = f64(cat_rates)
            self.site_cat = np.ascontiguousarray(site_cat, dtype=np.int32)
            assert self.site_cat.min() >= 0 and self.site_cat.max() < len(self.cat_rates)
            n_cats, cr, sc = len(self.cat_rates), P(self.cat_rates), P(self.site_cat, C.c_int)
        else:
            self.rates = f64(rates)
            self.cat_rates = self.site_cat = None
            n_cats, cr, sc = 0, None, None
        self.R = self.rates.shape[0]
        self.c = OrModel(self.R, self.S, P(self.statd), P(self.x), P(self.l), P(self.xit),
                         P(self.rates), n_cats, cr, sc)

    def evolve(self, a, e, bl):
        a = f64(a)
        e = np.ascontiguousarray(e, dtype=np.int64)
        R, N, S = a.shape
        da, de = np.empty_like(a), np.empty_like(e)
        lib().or_evolve_into(C.byref(self.c), P(da), P(de, C.c_int64), P(a), P(e, C.c_int64), N, bl)
        return da, de

    def log_like3(self, x, y, z):
        (xa, xe), (ya, ye), (za, ze) = x, y, z
        N = xa.shape[1]
        util = np.empty(max(N, 1))
        return lib().or_log_like3(C.byref(self.c), P(util), N, P(f64(xa)), P(xe, C.c_int64),
                                  P(f64(ya)), P(ye, C.c_int64), P(f64(za)), P(ze, C.c_int64))

    def slow_log_like3(self, x, y, z):
        (xa, xe), (ya, ye), (za, ze) = x, y, z
        N = xa.shape[1]
        return lib().or_slow_log_like3(C.byref(self.c), N, P(f64(xa)), P(xe, C.c_int64),
                                       P(f64(ya)), P(ye, C.c_int64), P(f64(za)), P(ze, C.c_int64))

    def logdot(self, x, y):
        (xa, xe), (ya, ye) = x, y
        R, N, S = xa.shape
        util = np.empty(max(N, 1))
        return lib().or_logdot(P(util), R, N, S, P(f64(xa)), P(xe, C.c_int64), P(f64(ya)),
                               P(ye, C.c_int64))


class TreeC:
    def __init__(self, tree):
        self.tree = tree
        self.starts, self.children = tree.csr()
        self.bl = f64(tree.bl)
        self.c = OrTree(tree.n_nodes, P(self.starts, C.c_int), P(self.children, C.c_int),
                        P(self.bl))


PRIOR_FLAT_EXP, PRIOR_UNIFORM, PRIOR_INFORMATIVE = 0, 1, 2


def default_prefs(**kw):
    """prefs.ml:65-121 defaults.  calc_pp=True applies -p's changes
    (prefs.ml:198-203: keep_at_most 20, keep_factor 0.001, max_strikes 20)
    unless they are given.  prior: "flat" (exponential, mean = tree length /
    edges, pplacer_run.ml:386-389), "uniform" or "informative" (with
    prior_lower, core.ml:34-52)."""
    p = dict(start_pend=0.1, max_pend=2.0, initial_tolerance=0.01, strike_box=3.0,
             keep_factor=0.01, max_strikes=6, max_pitches=40, keep_at_most=7,
             calc_pp=False, pp_rel_err=0.01, prior="flat", prior_lower=0.0)
    if kw.get("calc_pp"):
        p.update(keep_at_most=20, keep_factor=0.001, max_strikes=20)
    p.update(kw)
    return p


def tree_length(tree) -> float:
    """Gtree.tree_length (gtree.ml:102-114): every non-root branch length,
    summed subtree by subtree as the reference folds them."""
    def aux(i):
        acc = float(tree.bl[i])
        for c in tree.children[i]:
            acc = acc + aux(c)
        return acc
    tot = 0.0
    for c in tree.children[tree.root]:
        tot = tot + aux(c)
    return tot


def prior_mean_map(tree, lower=0.0):
    """Core.prior_mean_map (core.ml:34-52) with top_segment_f = (+ lower):
    per edge id, lower + its branch length + the average distance from its
    lower node to the leaves below, in the reference's list order."""
    bl = [float(x) for x in tree.bl]
    means = np.zeros(tree.n_nodes)

    def aux(i):
        if not tree.children[i]:
            means[i] = lower + bl[i]
            return [bl[i]]
        dist = []
        for c in tree.children[i]:
            dist = dist + aux(c)
        shifted = [d + (lower + bl[i]) for d in dist]
        tot = shifted[0]
        for d in shifted[1:]:
            tot = tot + d
        means[i] = tot / float(len(shifted))
        return [d + bl[i] for d in dist]

    aux(tree.root)
    return means


class Precompute:
    """darr / parr / snodes for a tree, as pplacer_run.ml:286-341 builds them."""

    def __init__(self, model: Model, tree, leaf_seqs: dict, n_sites: int):
        """leaf_seqs: leaf id -> sequence string of length n_sites."""
        self.model, self.tree, self.N = model, tree, n_sites
        self.tc = TreeC(tree)
        G, R, S = tree.n_nodes, model.R, model.S
        rows = np.full((G, n_sites), ord("-"), dtype=np.uint8)
        for i, s in leaf_seqs.items():
            rows[i] = np.frombuffer(s.encode(), dtype=np.uint8)
        self.rows = rows
        shp, eshp = (G, R, n_sites, S), (G, n_sites)
        self.darr, self.parr, self.snodes = np.empty(shp), np.empty(shp), np.empty(shp)
        self.darr_e, self.parr_e, self.snodes_e = (np.zeros(eshp, dtype=np.int64) for _ in range(3))
        st = lib().or_calc_distal_and_proximal(
            C.byref(model.c), C.byref(self.tc.c), rows.tobytes(), n_sites, P(self.darr),
            P(self.darr_e, C.c_int64), P(self.parr), P(self.parr_e, C.c_int64), P(self.snodes),
            P(self.snodes_e, C.c_int64))
        if st != OR_OK:
            raise ValueError("oracle precompute failed: %d" % st)

    def pull_and_supernodes(self):
        G, R, S, N = self.tree.n_nodes, self.model.R, self.model.S, self.N
        for a, e in ((self.darr, self.darr_e), (self.parr, self.parr_e)):
            lib().or_pull_exponents_arr(-10, P(a), P(e, C.c_int64), G, R, N, S)
        lib().or_prep_supernodes(C.byref(self.model.c), C.byref(self.tc.c), N, P(self.darr),
                                 P(self.darr_e, C.c_int64), P(self.parr),
                                 P(self.parr_e, C.c_int64), P(self.snodes),
                                 P(self.snodes_e, C.c_int64))


class Core:
    """Per-query placement on a precomputed tree (core.ml:76-411)."""

    def __init__(self, pre: Precompute, **prefs):
        self.pre = pre
        self.prefs = default_prefs(**prefs)
        p = self.prefs
        tree = pre.tree
        kind = {"flat": PRIOR_FLAT_EXP, "uniform": PRIOR_UNIFORM,
                "informative": PRIOR_INFORMATIVE}[p["prior"]]
        mean = tree_length(tree) / float(tree.n_nodes - 1)
        self.prior_means = f64(prior_mean_map(tree, p["prior_lower"]))
        self.c = OrCore(C.pointer(pre.model.c), C.pointer(pre.tc.c), pre.N, P(pre.darr),
                        P(pre.darr_e, C.c_int64), P(pre.parr), P(pre.parr_e, C.c_int64),
                        P(pre.snodes), P(pre.snodes_e, C.c_int64),
                        OrPrefs(p["start_pend"], p["max_pend"], p["initial_tolerance"],
                                p["strike_box"], p["keep_factor"], p["max_strikes"],
                                p["max_pitches"], p["keep_at_most"], int(bool(p["calc_pp"])),
                                p["pp_rel_err"], kind, mean, P(self.prior_means)))

    def log_like(self, query, loc, pend, distal):
        """Three_tax.log_like of the query at (loc, pend, distal)."""
        ll = C.c_double()
        st = lib().or_three_tax_log_like(C.byref(self.c), query.encode(), loc, pend, distal,
                                         C.byref(ll))
        assert st == 0
        return ll.value

    def marg_prob(self, query, loc, pend, distal):
        """Three_tax.calc_marg_prob at (loc, pend, distal): (status, value, calls)."""
        m = C.c_double()
        calls = C.c_long(0)
        st = lib().or_calc_marg_prob(C.byref(self.c), query.encode(), loc, pend, distal,
                                     C.byref(m), C.byref(calls))
        return st, m.value, calls.value

    def stage1_scores(self, query: str):
        E = self.pre.tree.n_nodes - 1
        out = np.empty(E)
        st = lib().or_stage1_scores(C.byref(self.c), query.encode(), P(out))
        if st != OR_OK:
            raise ValueError("stage1 failed %d" % st)
        return out

    @staticmethod
    def rank(scores):
        scores = f64(scores)
        order = np.empty(len(scores), dtype=np.int32)
        lib().or_rank(P(scores), len(scores), P(order, C.c_int))
        return order

    def optimize(self, query, loc, pend, distal, tol):
        ll, pe, di = C.c_double(), C.c_double(), C.c_double()
        calls = C.c_long(0)
        st = lib().or_three_tax_optimize(C.byref(self.c), query.encode(), loc, pend, distal, tol,
                                         C.byref(ll), C.byref(pe), C.byref(di), C.byref(calls))
        return st, ll.value, pe.value, di.value, calls.value

    def place(self, query: str):
        K = self.prefs["keep_at_most"]
        out = (OrPlacement * K)()
        n = C.c_int(0)
        stats = OrStats()
        st = lib().or_process_query(C.byref(self.c), query.encode(), out, C.byref(n),
                                    C.byref(stats))
        pp = self.prefs["calc_pp"]
        pl = [(out[i].loc, out[i].ml_ratio, out[i].log_like, out[i].dist_bl, out[i].pend_bl,
               out[i].post_prob if pp else None, out[i].marginal_prob if pp else None)
              for i in range(n.value)]
        return st, pl, stats

    def place_many(self, queries, n_threads=1):
        K = self.prefs["keep_at_most"]
        Q = len(queries)
        out = (OrPlacement * (Q * K))()
        n_out = np.zeros(Q, dtype=np.int32)
        status = np.zeros(Q, dtype=np.int32)
        stats = OrStats()
        st = lib().or_process_queries(C.byref(self.c), "".join(queries).encode(), Q, n_threads,
                                      out, P(n_out, C.c_int), P(status, C.c_int), C.byref(stats))
        res = []
        for q in range(Q):
            res.append([(out[q * K + i].loc, out[q * K + i].ml_ratio, out[q * K + i].log_like,
                         out[q * K + i].dist_bl, out[q * K + i].pend_bl)
                        for i in range(n_out[q])])
        return st, res, status, stats


def brent(f, raw_start, left, right, tol, max_iters=100):
    calls = []

    def cb(x, _ctx):
        calls.append(x)
        return f(x)

    fn = OR_FUN(cb)
    res = C.c_double()
    st = lib().or_brent(fn, None, raw_start, left, right, tol, max_iters, C.byref(res))
    return st, res.value, calls


class GslBrent:
    """The oracle's GSL Brent minimiser driven step by step, as a GSL program
    drives gsl_min_fminimizer_set / gsl_min_fminimizer_iterate."""

    def __init__(self, f):
        self.points = []

        def cb(x, _ctx):
            self.points.append(x)
            return f(x)

        self._fn = OR_FUN(cb)
        self.s = OrGslBrent()

    def set(self, x_minimum, x_lower, x_upper):
        return lib().or_gsl_brent_set(C.byref(self.s), self._fn, None, x_minimum, x_lower, x_upper)

    def iterate(self):
        return lib().or_gsl_brent_iterate(C.byref(self.s), self._fn, None)


def qng(f, a, b, epsabs, epsrel):
    """or_qng on a Python function: (status, result, abserr, neval, points)."""
    pts = []

    def cb(x, _ctx):
        pts.append(x)
        return f(x)

    fn = OR_FUN(cb)
    r, e, n = C.c_double(), C.c_double(), C.c_int()
    st = lib().or_qng(fn, None, a, b, epsabs, epsrel, C.byref(r), C.byref(e), C.byref(n), None)
    return st, r.value, e.value, n.value, pts


def ll_normalized_prob(ll):
    ll = f64(ll)
    out = np.empty_like(ll)
    lib().or_ll_normalized_prob(P(ll), len(ll), P(out))
    return out
