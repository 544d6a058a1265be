"""Host-side mirror of pplacer's placement driver over libpplacer_hip.so.

`Placer` follows Pplacer_run.run_placements (pplacer_src/pplacer_run.ml:167-656)
and Core.pplacer_core (pplacer_src/core.ml:76-411) for the gamma mixture model:
build the model from the reference alignment, pre-mask, compute the
reference-tree GLVs, then place batches of queries.  Every numeric step runs
in the C++/HIP library; this module only moves arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import refio
from ._lib import (PP_EQUERY, PP_FP32, PP_FP64, PP_PRIOR_FLAT_EXP, PP_PRIOR_INFORMATIVE,
                   PP_PRIOR_UNIFORM, LaunchConfig, Placement, Prefs, Stats, check, lib)

_dp = C.POINTER(C.c_double)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def default_prefs(**kw) -> Prefs:
    """prefs.ml:65-121 defaults (start_pend 0.1, max_pend 2, initial_tolerance
    0.01, strike_box 3, keep_factor 0.01, max_strikes 6, max_pitches 40,
    keep_at_most 7, pp_rel_err 0.01).  calc_pp=True applies the rest of -p
    (prefs.ml:198-203: keep_at_most 20, keep_factor 0.001, max_strikes 20)
    unless those are given too."""
    p = Prefs()
    lib().pp_prefs_default(C.byref(p))
    if kw.get("calc_pp"):
        p.keep_at_most, p.keep_factor, p.max_strikes = 20, 0.001, 20
    for k, v in kw.items():
        setattr(p, k, int(v) if k == "calc_pp" else v)
    return p


def tree_length(tree: refio.Tree) -> float:
    """Gtree.tree_length (gtree.ml:102-114)."""
    starts, children = tree.csr()
    out = C.c_double()
    check(lib().pp_tree_length(tree.n_nodes, _p(starts, C.c_int), _p(children, C.c_int),
                               _p(_f64(tree.bl)), C.byref(out)), "tree_length")
    return out.value


def prior_mean_map(tree: refio.Tree, lower: float = 0.0) -> np.ndarray:
    """Core.prior_mean_map with top_segment_f = (+ lower) (core.ml:34-52):
    the informative prior's mean for every edge id."""
    starts, children = tree.csr()
    out = np.empty(tree.n_nodes)
    check(lib().pp_prior_mean_map(tree.n_nodes, _p(starts, C.c_int), _p(children, C.c_int),
                                  _p(_f64(tree.bl)), lower, _p(out)), "prior_mean_map")
    return out


def discrete_gamma(n_cat: int, alpha: float) -> np.ndarray:
    out = np.empty(n_cat)
    check(lib().pp_discrete_gamma(n_cat, alpha, _p(out)), "discrete_gamma")
    return out


def emper_freq(n_states: int, seqs) -> np.ndarray:
    n_sites = len(seqs[0])
    out = np.empty(n_states)
    check(lib().pp_emper_freq(n_states, "".join(seqs).encode(), len(seqs), n_sites, _p(out)),
          "emper_freq")
    return out


def premask(n_states: int, ref_seqs, query_seqs, discard_nonoverlapped=False):
    """(mask, query_keep, masklen) as Pplacer_run.premask computes them."""
    n_sites = len(ref_seqs[0])
    mask = np.zeros(n_sites, dtype=np.uint8)
    keep = np.zeros(max(len(query_seqs), 1), dtype=np.uint8)
    ml = C.c_int(0)
    check(lib().pp_premask(n_states, "".join(ref_seqs).encode(), len(ref_seqs),
                           "".join(query_seqs).encode(), len(query_seqs), n_sites,
                           int(discard_nonoverlapped), _p(mask, C.c_uint8), _p(keep, C.c_uint8),
                           C.byref(ml)), "premask")
    return mask.astype(bool), keep[: len(query_seqs)].astype(bool), ml.value


def cut_columns(seqs, mask):
    m = np.asarray(mask, dtype=bool)
    out = []
    for s in seqs:
        a = np.frombuffer(s.encode(), dtype=np.uint8)
        out.append(a[m].tobytes().decode())
    return out


@dataclass
class PlacementRecord:
    """Placement.make_ml fields (placement.ml:50-60); with -p also
    post_prob and marginal_like (placement.ml:62-65, 176-180)."""

    edge_num: int
    like_weight_ratio: float
    likelihood: float
    distal_length: float
    pendant_length: float
    post_prob: float | None = None
    marginal_like: float | None = None


class Placer:
    """One GPU context holding a model, a reference tree and query batches."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        check(lib().pp_ctx_create(device, C.byref(self._ctx)), "pp_ctx_create")
        self.n_states = None
        self.n_rates = None
        self.tree = None
        self.n_sites = None

    def close(self):
        if self._ctx:
            from . import glvm
            glvm.forget_context(self._ctx)  # pp_ctx_destroy frees the context's GLVs
            lib().pp_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- model (gmix_model.ml:25-37) -------------------------------------
    def build_model(self, exch, statd, rates):
        exch, statd, rates = _f64(exch), _f64(statd), _f64(rates)
        S, R = statd.shape[0], rates.shape[0]
        check(lib().pp_model_build(self._ctx, S, R, _p(exch), _p(statd), _p(rates)), "model_build")
        self.n_states, self.n_rates = S, R

    def build_model_cat(self, exch, statd, cat_rates, site_cat):
        """Price-CAT model (gcat_model.ml:33-95): rate-free GLVs, site i
        evolving at cat_rates[site_cat[i]].  site_cat indexes the pre-masked
        sites, 0-based (Gcat_model.mask_sites, gcat_model.ml:429-439)."""
        exch, statd, cat_rates = _f64(exch), _f64(statd), _f64(cat_rates)
        site_cat = np.ascontiguousarray(site_cat, dtype=np.int32)
        S = statd.shape[0]
        check(lib().pp_model_build_cat(self._ctx, S, cat_rates.shape[0], _p(exch), _p(statd),
                                       _p(cat_rates)), "model_build_cat")
        check(lib().pp_site_categories(self._ctx, site_cat.shape[0], _p(site_cat, C.c_int)),
              "site_categories")
        self.n_states, self.n_rates = S, 1

    def model_diagd(self):
        S = self.n_states
        x, l, xit = np.empty((S, S)), np.empty(S), np.empty((S, S))
        check(lib().pp_model_get(self._ctx, _p(x), _p(l), _p(xit)), "model_get")
        return x, l, xit

    # ---- tree ------------------------------------------------------------
    def set_tree(self, tree: refio.Tree, leaf_seqs: dict, n_sites: int):
        """leaf_seqs: leaf id -> (pre-masked) sequence."""
        starts, children = tree.csr()
        G = tree.n_nodes
        rows = np.full((G, n_sites), ord("-"), dtype=np.uint8)
        for i, s in leaf_seqs.items():
            rows[i] = np.frombuffer(s.encode(), dtype=np.uint8)
        bl = _f64(tree.bl)
        check(lib().pp_tree_set(self._ctx, G, _p(starts, C.c_int), _p(children, C.c_int), _p(bl),
                                n_sites, rows.tobytes()), "tree_set")
        self.tree, self.n_sites = tree, n_sites

    def upload_glvs(self, tree, n_sites, darr, darr_e, parr, parr_e, snodes, snodes_e):
        G = tree.n_nodes
        args = [_f64(darr), np.ascontiguousarray(darr_e, np.int64), _f64(parr),
                np.ascontiguousarray(parr_e, np.int64), _f64(snodes),
                np.ascontiguousarray(snodes_e, np.int64)]
        bl = _f64(tree.bl)
        check(lib().pp_tree_upload(self._ctx, G, _p(bl), n_sites, _p(args[0]), _p(args[1], C.c_int64),
                                   _p(args[2]), _p(args[3], C.c_int64), _p(args[4]),
                                   _p(args[5], C.c_int64)), "tree_upload")
        self.tree, self.n_sites = tree, n_sites

    def download_glvs(self):
        G, N, R, S = self.tree.n_nodes, self.n_sites, self.n_rates, self.n_states
        shp, eshp = (G, R, N, S), (G, N)
        d, p, s = np.empty(shp), np.empty(shp), np.empty(shp)
        de, pe, se = (np.empty(eshp, dtype=np.int64) for _ in range(3))
        check(lib().pp_tree_download(self._ctx, _p(d), _p(de, C.c_int64), _p(p), _p(pe, C.c_int64),
                                     _p(s), _p(se, C.c_int64)), "tree_download")
        return d, de, p, pe, s, se

    # ---- stage 1 / ranking / stage 2 -------------------------------------
    def stage1_scores(self, queries, start_pend=0.1, precision=PP_FP64):
        E = self.tree.n_nodes - 1
        Q = len(queries)
        out = np.empty((Q, E))
        check(lib().pp_stage1_scores(self._ctx, Q, "".join(queries).encode(), start_pend, precision,
                                     _p(out)), "stage1_scores")
        return out

    def rank_topk(self, scores, k):
        scores = _f64(scores)
        Q, E = scores.shape
        out = np.empty((Q, k), dtype=np.int32)
        check(lib().pp_rank_topk(self._ctx, Q, E, _p(scores), k, _p(out, C.c_int)), "rank_topk")
        return out

    def optimize(self, queries, query_of_job, locs, pend0, dist0, tolerance, max_pend=2.0,
                 precision=PP_FP64):
        n = len(locs)
        qj = np.ascontiguousarray(query_of_job, np.int32)
        lc = np.ascontiguousarray(locs, np.int32)
        p0, d0 = _f64(pend0), _f64(dist0)
        ll, pe, di = np.empty(n), np.empty(n), np.empty(n)
        calls, st = np.empty(n, np.int32), np.empty(n, np.int32)
        check(lib().pp_optimize_batch(self._ctx, n, "".join(queries).encode(), _p(qj, C.c_int),
                                      _p(lc, C.c_int), _p(p0), _p(d0), tolerance, max_pend,
                                      precision, _p(ll), _p(pe), _p(di), _p(calls, C.c_int),
                                      _p(st, C.c_int)), "optimize_batch")
        return ll, pe, di, calls, st

    def place(self, queries, prefs: Prefs | None = None, precision=PP_FP64, strict=False):
        """Core.process_query for every query.  Returns (placements, status,
        stats): placements[q] is a list of PlacementRecord sorted by
        like_weight_ratio descending."""
        prefs = prefs or default_prefs()
        Q = len(queries)
        K = max(1, prefs.keep_at_most)
        out = (Placement * max(1, Q * K))()
        n_out = np.zeros(max(Q, 1), np.int32)
        status = np.zeros(max(Q, 1), np.int32)
        stats = Stats()
        # pp_place_batch: one pass, and in the FP32 form the FP64 retry of
        # queries that failed on the prox_bl assertion; PP_EQUERY only says
        # that some query failed, and status holds which
        rc = lib().pp_place_batch(self._ctx, Q, "".join(queries).encode(), C.byref(prefs), precision,
                                  out, _p(n_out, C.c_int), _p(status, C.c_int), C.byref(stats))
        if rc != PP_EQUERY:
            check(rc, "place_batch")
        res = []
        for q in range(Q):
            row = []
            for i in range(n_out[q]):
                o = out[q * K + i]
                pp = (o.post_prob, o.marginal_prob) if prefs.calc_pp else (None, None)
                row.append(PlacementRecord(o.loc, o.ml_ratio, o.log_like, o.dist_bl, o.pend_bl,
                                           *pp))
            res.append(row)
        if strict and np.any(status[:Q] != 0):
            bad = int(np.nonzero(status[:Q])[0][0])
            raise RuntimeError("query %d failed with status %d" % (bad, status[bad]))
        return res, status[:Q].copy(), stats

    def set_prior(self, kind: str = "flat", lower: float = 0.0):
        """The -p prior as pplacer_run.ml:380-389 chooses it: "flat"
        (exponential, mean = tree length / edges; the default pp_tree_set
        installs), "uniform" (--unif-prior) or "informative" (--inform-prior,
        with --prior-lower)."""
        if kind == "uniform":
            check(lib().pp_set_prior(self._ctx, PP_PRIOR_UNIFORM, 0, None), "set_prior")
        elif kind == "flat":
            m = np.array([tree_length(self.tree) / float(self.tree.n_nodes - 1)])
            check(lib().pp_set_prior(self._ctx, PP_PRIOR_FLAT_EXP, 1, _p(m)), "set_prior")
        elif kind == "informative":
            m = prior_mean_map(self.tree, lower)
            check(lib().pp_set_prior(self._ctx, PP_PRIOR_INFORMATIVE, m.shape[0], _p(m)),
                  "set_prior")
        else:
            raise ValueError("unknown prior %r" % kind)

    def marginal(self, queries, query_of_job, locs, pend, dist, rel_err=0.01, max_pend=2.0):
        """Three_tax.calc_marg_prob per job: (marginal log-likelihoods, calls, status)."""
        n = len(locs)
        qj = np.ascontiguousarray(query_of_job, np.int32)
        lc = np.ascontiguousarray(locs, np.int32)
        pe, di = _f64(pend), _f64(dist)
        m = np.empty(n)
        calls, st = np.empty(n, np.int32), np.empty(n, np.int32)
        check(lib().pp_marginal_batch(self._ctx, n, "".join(queries).encode(), _p(qj, C.c_int),
                                      _p(lc, C.c_int), _p(pe), _p(di), rel_err, max_pend, _p(m),
                                      _p(calls, C.c_int), _p(st, C.c_int)), "marginal_batch")
        return m, calls, st

    def upload_queries(self, queries):
        check(lib().pp_queries_upload(self._ctx, len(queries), "".join(queries).encode()),
              "queries_upload")

    def place_resident(self, prefs: Prefs, precision=PP_FP64, stats: Stats | None = None):
        stats = stats if stats is not None else Stats()
        check(lib().pp_place_resident(self._ctx, C.byref(prefs), precision, C.byref(stats)),
              "place_resident")
        return stats

    def download_results(self, n_queries: int, keep_at_most: int):
        """pp_results_download after place_resident: (placements, status) as
        place returns them, for the n_queries resident queries."""
        K = max(1, keep_at_most)
        out = (Placement * max(1, n_queries * K))()
        n_out = np.zeros(max(n_queries, 1), np.int32)
        status = np.zeros(max(n_queries, 1), np.int32)
        check(lib().pp_results_download(self._ctx, out, _p(n_out, C.c_int), _p(status, C.c_int)),
              "results_download")
        res = [[PlacementRecord(out[q * K + i].loc, out[q * K + i].ml_ratio, out[q * K + i].log_like,
                                out[q * K + i].dist_bl, out[q * K + i].pend_bl, None, None)
                for i in range(n_out[q])] for q in range(n_queries)]
        return res, status[:n_queries].copy()

    def synchronize(self):
        check(lib().pp_synchronize(self._ctx), "synchronize")

    def launch_config(self) -> dict:
        """The pass's launch shape: streams per stage-2 phase and form
        (index 0 = FP64, 1 = FP32), queries per device chunk, and whether the
        pass's launches are captured into a replayed HIP graph."""
        cfg = LaunchConfig()
        check(lib().pp_launch_config_get(self._ctx, C.byref(cfg)), "launch_config_get")
        return {"pitch_streams": list(cfg.pitch_streams), "refine_streams": list(cfg.refine_streams),
                "chunk": cfg.chunk, "graph": cfg.graph}

    def set_launch_config(self, pitch_streams=None, refine_streams=None, chunk=None, graph=None) -> dict:
        """Changes the given fields of the launch shape (results do not
        depend on it); returns the new shape."""
        cur = self.launch_config()
        cfg = LaunchConfig()
        for k, v in (("pitch_streams", pitch_streams), ("refine_streams", refine_streams)):
            v = cur[k] if v is None else list(v)
            getattr(cfg, k)[0], getattr(cfg, k)[1] = int(v[0]), int(v[1])
        cfg.chunk = int(cur["chunk"] if chunk is None else chunk)
        cfg.graph = int(cur["graph"] if graph is None else graph)
        check(lib().pp_launch_config_set(self._ctx, C.byref(cfg)), "launch_config_set")
        return self.launch_config()


def setup_from_workload(wl, device=0, premasked=True):
    """Model + pre-mask + tree for a synth.Workload, as run_placements does it
    (model from the un-masked reference alignment; pplacer_run.ml:200-205)."""
    S = wl.n_states
    ref_seqs = [s for _, s in wl.ref_aln]
    statd = emper_freq(S, ref_seqs)
    rates = discrete_gamma(wl.n_cats, wl.alpha)
    queries = [s for _, s in wl.queries]
    leaf = wl.leaf_seqs()
    if premasked:
        mask, keep, n = premask(S, ref_seqs, queries)
        queries = cut_columns(queries, mask)
        cut_leaf = cut_columns(list(leaf.values()), mask)
        leaf = dict(zip(leaf.keys(), cut_leaf))
    n_sites = len(next(iter(leaf.values())))
    pl = Placer(device)
    pl.build_model(wl.exch, statd, rates)
    pl.set_tree(wl.tree, leaf, n_sites)
    return pl, queries, statd, rates


__all__ = ["Placer", "PlacementRecord", "default_prefs", "discrete_gamma", "emper_freq",
           "premask", "cut_columns", "setup_from_workload", "PP_FP32", "PP_FP64"]
