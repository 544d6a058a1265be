"""Python mirror of pplacer's `Linear` module (pplacer_src/linear.ml:13-89):
the C stubs it binds, re-pointed at the GPU twins in libpplacer_hip.so.

Each function takes numpy arrays in pplacer's layout (tensor = [rate][site]
[state] float64, c_layout) and returns what the OCaml `external` returns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

_dp = C.POINTER(C.c_double)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def gemmish(dst, a, b):
    """dst = a * b^T row-wise (linear_c.c:29-90): dst[site][i] = sum_j a[i][j] b[site][j]."""
    a, b = _f64(a), _f64(b)
    n_sites, S = b.shape
    out = np.empty((n_sites, S))
    check(lib().pp_gemmish(_p(out), _p(a), _p(b), n_sites, S), "gemmish")
    if dst is not None:
        dst[...] = out
    return out


def dediagonalize(dst, u, lam, uit):
    """dst_ij = sum_k lambda_k u_ik uit_jk (linear_c.c:92-146)."""
    u, lam, uit = _f64(u), _f64(lam), _f64(uit)
    n = lam.shape[0]
    out = np.empty((n, n))
    check(lib().pp_dediagonalize(_p(out), _p(u), _p(lam), _p(uit), n), "dediagonalize")
    if dst is not None:
        dst[...] = out
    return out


def ten_log_like3(statd, x, y, z, util=None):
    """linear_c.c:367-424."""
    statd, x, y, z = _f64(statd), _f64(x), _f64(y), _f64(z)
    R, N, S = x.shape
    if y.shape != x.shape or z.shape != x.shape:
        raise ValueError("bad input dimensions")
    u = np.empty(max(N, 1))
    out = C.c_double()
    check(lib().pp_ten_log_like3(_p(statd), _p(x), _p(y), _p(z), _p(u), R, N, S, C.byref(out)),
          "ten_log_like3")
    if util is not None:
        util[:N] = u[:N]
    return out.value


def ten_pairwise_prod(dst, x, y):
    x, y = _f64(x), _f64(y)
    R, N, S = x.shape
    out = np.empty_like(x)
    check(lib().pp_ten_pairwise_prod(_p(out), _p(x), _p(y), R, N, S), "ten_pairwise_prod")
    if dst is not None:
        dst[...] = out
    return out


def ten_statd_pairwise_prod(statd, dst, a, b):
    statd, a, b = _f64(statd), _f64(a), _f64(b)
    R, N, S = a.shape
    out = np.empty_like(a)
    check(lib().pp_ten_statd_pairwise_prod(_p(statd), _p(out), _p(a), _p(b), R, N, S),
          "ten_statd_pairwise_prod")
    if dst is not None:
        dst[...] = out
    return out


def ten_masked_logdot(x, y, mask, util=None):
    """linear_c.c:465-539."""
    x, y = _f64(x), _f64(y)
    mask = np.ascontiguousarray(mask, dtype=np.uint16)
    R, N, S = x.shape
    if mask.shape[0] != N:
        raise ValueError("ten_masked_logdot: Mask length doesn't match!")
    u = np.empty(max(N, 1))
    out = C.c_double()
    check(lib().pp_ten_masked_logdot(_p(x), _p(y), _p(mask, C.c_uint16), _p(u), R, N, S,
                                     C.byref(out)), "ten_masked_logdot")
    if util is not None:
        util[:N] = u[:N]
    return out.value


def ten_bounded_logdot(x, y, first, last, util=None):
    """linear_c.c:541-605."""
    x, y = _f64(x), _f64(y)
    R, N, S = x.shape
    u = np.empty(max(N, 1))
    out = C.c_double()
    check(lib().pp_ten_bounded_logdot(_p(x), _p(y), int(first), int(last), _p(u), R, N, S,
                                      C.byref(out)), "ten_bounded_logdot")
    if util is not None:
        util[: last - first + 1] = u[: last - first + 1]
    return out.value


# ---- mat_* (the Price-CAT model's rate-free [site][state] GLVs) --------------

def mat_log_like3(statd, x, y, z):
    """linear_c.c:168-213 (no util vector, no n log R term)."""
    statd, x, y, z = _f64(statd), _f64(x), _f64(y), _f64(z)
    N, S = x.shape
    if y.shape != x.shape or z.shape != x.shape:
        raise ValueError("bad input dimensions")
    out = C.c_double()
    check(lib().pp_mat_log_like3(_p(statd), _p(x), _p(y), _p(z), N, S, C.byref(out)), "mat_log_like3")
    return out.value


def mat_pairwise_prod(dst, x, y):
    x, y = _f64(x), _f64(y)
    N, S = x.shape
    out = np.empty_like(x)
    check(lib().pp_mat_pairwise_prod(_p(out), _p(x), _p(y), N, S), "mat_pairwise_prod")
    if dst is not None:
        dst[...] = out
    return out


def mat_statd_pairwise_prod(statd, dst, a, b):
    statd, a, b = _f64(statd), _f64(a), _f64(b)
    N, S = a.shape
    out = np.empty_like(a)
    check(lib().pp_mat_statd_pairwise_prod(_p(statd), _p(out), _p(a), _p(b), N, S),
          "mat_statd_pairwise_prod")
    if dst is not None:
        dst[...] = out
    return out


def mat_masked_logdot(x, y, mask):
    """linear_c.c:250-293."""
    x, y = _f64(x), _f64(y)
    mask = np.ascontiguousarray(mask, dtype=np.uint16)
    N, S = x.shape
    if mask.shape[0] != N:
        raise ValueError("mat_masked_logdot: Mask length doesn't match!")
    out = C.c_double()
    check(lib().pp_mat_masked_logdot(_p(x), _p(y), _p(mask, C.c_uint16), N, S, C.byref(out)),
          "mat_masked_logdot")
    return out.value


def mat_bounded_logdot(x, y, first, last):
    """linear_c.c:295-340."""
    x, y = _f64(x), _f64(y)
    N, S = x.shape
    out = C.c_double()
    check(lib().pp_mat_bounded_logdot(_p(x), _p(y), int(first), int(last), N, S, C.byref(out)),
          "mat_bounded_logdot")
    return out.value
