"""Python mirror of Glvm.Model for the gamma mixture model
(pplacer_src/glvm.ml:11-55, implemented by gmix_model.ml:58-351) over the
library's device GLVs: a `Glv` owns a pp_glv handle whose data stays in HBM,
and `GmixModel` holds the model functions (log_like3, evolve_into, ...), with
the model of a `Placer`'s context (placer.Placer.build_model).

Host arrays use pplacer's layout: a = [rate][site][state] float64, e =
[site] int64 (the OCaml Bigarrays).
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from ._lib import check, lib

PP_FP_NORMAL, PP_FP_SUBNORMAL, PP_FP_ZERO, PP_FP_INFINITE, PP_FP_NAN = range(5)

# live Glv objects per context: pp_ctx_destroy frees a context's GLVs, so
# their handles are dropped first (forget_context)
_live: dict = {}


def forget_context(ctx) -> None:
    """Drop the handles of every Glv made on ctx (called before the context
    is destroyed, which frees them)."""
    for g in list(_live.pop(ctx.value, ())):
        g._h = C.c_void_p()


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Glv:
    """Gmix_model.Model.Glv.t on the GPU (gmix_model.ml:58-257)."""

    def __init__(self, ctx, handle):
        self._ctx = ctx
        self._h = handle
        _live.setdefault(ctx.value, weakref.WeakSet()).add(self)

    # ---- construction ----------------------------------------------------
    @classmethod
    def make(cls, ctx, n_rates: int, n_sites: int, n_states: int) -> "Glv":
        """Glv.make (gmix_model.ml:76-79); contents zero."""
        h = C.c_void_p()
        check(lib().pp_glv_make(ctx, n_rates, n_sites, n_states, C.byref(h)), "glv_make")
        return cls(ctx, h)

    @classmethod
    def of_arrays(cls, ctx, a, e) -> "Glv":
        a = np.ascontiguousarray(a, dtype=np.float64)
        g = cls.make(ctx, *a.shape)
        g.set(a, e)
        return g

    def mimic(self) -> "Glv":
        h = C.c_void_p()
        check(lib().pp_glv_mimic(self._h, C.byref(h)), "glv_mimic")
        return Glv(self._ctx, h)

    def copy(self) -> "Glv":
        h = C.c_void_p()
        check(lib().pp_glv_copy(self._h, C.byref(h)), "glv_copy")
        return Glv(self._ctx, h)

    def free(self) -> None:
        if self._h:
            lib().pp_glv_free(self._h)
            self._h = C.c_void_p()
            live = _live.get(self._ctx.value)
            if live is not None:
                live.discard(self)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ---- contents ----------------------------------------------------------
    @property
    def dims(self):
        r, n, s = C.c_int(), C.c_int(), C.c_int()
        check(lib().pp_glv_dims(self._h, C.byref(r), C.byref(n), C.byref(s)), "glv_dims")
        return r.value, n.value, s.value

    @property
    def n_sites(self) -> int:
        return self.dims[1]

    def set(self, a, e) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.int64)
        assert a.shape == self.dims and e.shape == (self.dims[1],)
        check(lib().pp_glv_set(self._h, _p(a), _p(e, C.c_int64)), "glv_set")

    def get(self):
        R, N, S = self.dims
        a, e = np.empty((R, N, S)), np.empty(N, np.int64)
        check(lib().pp_glv_get(self._h, _p(a), _p(e, C.c_int64)), "glv_get")
        return a, e

    def memcpy_from(self, src: "Glv") -> None:
        """Glv.memcpy ~dst:self ~src (gmix_model.ml:93-95)."""
        check(lib().pp_glv_memcpy(self._h, src._h), "glv_memcpy")

    def set_unit(self) -> None:
        check(lib().pp_glv_set_unit(self._h), "glv_set_unit")

    def fp_classify(self) -> int:
        c = C.c_int()
        check(lib().pp_glv_fp_classify(self._h, C.byref(c)), "glv_fp_classify")
        return c.value

    def mask_into(self, site_mask, dst: "Glv") -> None:
        """Glv.mask_into site_mask ~src:self ~dst (gmix_model.ml:129-141)."""
        m = np.ascontiguousarray(site_mask, dtype=np.uint8)
        check(lib().pp_glv_mask_into(_p(m, C.c_uint8), m.shape[0], self._h, dst._h), "glv_mask_into")

    def prep_constant_rate_from_lv_arr(self, lv_arr) -> None:
        lv = np.ascontiguousarray(lv_arr, dtype=np.float64)
        check(lib().pp_glv_prep_constant_rate_from_lv_arr(self._h, _p(lv)), "glv_prep_constant_rate")

    def perhaps_pull_exponent(self, min_allowed_twoexp: int) -> None:
        check(lib().pp_glv_perhaps_pull_exponent(min_allowed_twoexp, self._h), "glv_pull_exponent")

    def summarize_post(self, summarize_f, initial=None):
        """Glv.summarize_post (gmix_model.ml:242-256): summarize_f of each
        site's state sums over rates."""
        R, N, S = self.dims
        u = np.empty((N, S))
        check(lib().pp_glv_summarize_post(self._h, _p(u)), "glv_summarize_post")
        return [summarize_f(u[i]) for i in range(N)]

    # ---- likelihoods -------------------------------------------------------
    def masked_logdot(self, y: "Glv", mask) -> float:
        m = np.ascontiguousarray(mask, dtype=np.uint16)
        out = C.c_double()
        check(lib().pp_glv_masked_logdot(self._h, y._h, _p(m, C.c_uint16), C.byref(out)), "glv_masked_logdot")
        return out.value

    def bounded_logdot(self, y: "Glv", start: int, last: int) -> float:
        out = C.c_double()
        check(lib().pp_glv_bounded_logdot(self._h, y._h, start, last, C.byref(out)), "glv_bounded_logdot")
        return out.value

    def logdot(self, y: "Glv") -> float:
        out = C.c_double()
        check(lib().pp_glv_logdot(self._h, y._h, C.byref(out)), "glv_logdot")
        return out.value

    def pairwise_prod(self, x: "Glv", y: "Glv") -> None:
        """Glv.pairwise_prod ~dst:self x y (gmix_model.ml:215-218)."""
        check(lib().pp_glv_pairwise_prod(self._h, x._h, y._h), "glv_pairwise_prod")

    def listwise_prod(self, glvs) -> None:
        """Glv.listwise_prod self glvs (gmix_model.ml:224-233)."""
        arr = (C.c_void_p * len(glvs))(*[g._h.value for g in glvs])
        check(lib().pp_glv_listwise_prod(self._h, arr, len(glvs)), "glv_listwise_prod")


class GmixModel:
    """The model functions of Gmix_model.Model (gmix_model.ml:258-351) on the
    model of a Placer's context."""

    def __init__(self, placer):
        self._ctx = placer._ctx
        self.n_rates, self.n_states = placer.n_rates, placer.n_states

    def make_glv(self, n_sites: int) -> Glv:
        return Glv.make(self._ctx, self.n_rates, n_sites, self.n_states)

    def make_constant_rate_glv_from_lv_arr(self, lv_arr) -> Glv:
        lv = np.ascontiguousarray(lv_arr, dtype=np.float64)
        g = self.make_glv(lv.shape[0])
        g.prep_constant_rate_from_lv_arr(lv)
        return g

    def log_like3(self, x: Glv, y: Glv, z: Glv) -> float:
        out = C.c_double()
        check(lib().pp_gmix_log_like3(self._ctx, x._h, y._h, z._h, C.byref(out)), "gmix_log_like3")
        return out.value

    def site_log_like_arr3(self, x: Glv, y: Glv, z: Glv):
        out = np.empty(max(x.n_sites, 1))
        check(lib().pp_gmix_site_log_like_arr3(self._ctx, x._h, y._h, z._h, _p(out)), "gmix_site_log_like_arr3")
        return out[: x.n_sites]

    def slow_log_like3(self, x: Glv, y: Glv, z: Glv) -> float:
        out = C.c_double()
        check(lib().pp_gmix_slow_log_like3(self._ctx, x._h, y._h, z._h, C.byref(out)), "gmix_slow_log_like3")
        return out.value

    def evolve_into(self, dst: Glv, src: Glv, bl: float) -> None:
        check(lib().pp_gmix_evolve_into(self._ctx, dst._h, src._h, bl), "gmix_evolve_into")

    def statd_pairwise_prod(self, dst: Glv, a: Glv, b: Glv) -> None:
        check(lib().pp_gmix_statd_pairwise_prod(self._ctx, dst._h, a._h, b._h), "gmix_statd_pairwise_prod")
