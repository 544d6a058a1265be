"""The correctly rounded FP64 exp and log (pplacer_amd/csrc/pp_crmath.h),
built for the host through the library's test hook: every result checked
against mpmath at 50 digits (rounded to nearest), on random arguments over
the ranges the kernels use, near 1 and near the table boundaries, and on
every argument where glibc's exp or log (the stubs' libm) gives another
double -- there glibc is the one that misrounds.  The device build of the
same source is held bit-equal to this one in tests/test_gpu_crmath.py."""
import ctypes as C

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")


def _cr(x):
    from pplacer_amd._lib import lib
    dp = C.POINTER(C.c_double)
    x = np.ascontiguousarray(x, dtype=np.float64)
    e, l = np.empty_like(x), np.empty_like(x)
    assert lib().pp_test_cr_math_host(len(x), x.ctypes.data_as(dp), e.ctypes.data_as(dp),
                                      l.ctypes.data_as(dp)) == 0
    return e, l


def _args(seed, n):
    rng = np.random.default_rng(seed)
    xe = np.concatenate([rng.uniform(-60, 0, n), rng.uniform(-708, 709.6, n), rng.normal(0, 1e-3, n),
                         # around multiples of ln2/128, where the table index changes
                         np.arange(-200, 200) * (np.log(2) / 128) + rng.normal(0, 1e-12, 400)])
    xl = np.concatenate([np.exp(rng.uniform(-700, 700, n)), rng.uniform(0.98, 1.02, n), 10 ** rng.uniform(-30, 0, n),
                         # interval edges of m in [0.75, 1.5)
                         (0.75 + np.arange(97) / 128.) * (1 + rng.normal(0, 1e-15, 97)),
                         np.nextafter(1.0, 0.0) - np.arange(20) * 2.0 ** -53, 1.0 + np.arange(20) * 2.0 ** -52])
    return xe, xl


def test_exp_log_correctly_rounded_against_mpmath():
    mp.mp.dps = 50
    xe, xl = _args(7, 1500)
    e, _ = _cr(xe)
    _, l = _cr(xl)
    bad_e = [x for x, y in zip(xe, e) if float(mp.exp(mp.mpf(float(x)))) != y]
    bad_l = [x for x, y in zip(xl, l) if float(mp.log(mp.mpf(float(x)))) != y]
    assert not bad_e, [float(v).hex() for v in bad_e[:5]]
    assert not bad_l, [float(v).hex() for v in bad_l[:5]]


def test_glibc_disagreements_are_glibc_misroundings():
    mp.mp.dps = 50
    libm = C.CDLL("libm.so.6")
    libm.exp.restype = libm.log.restype = C.c_double
    libm.exp.argtypes = libm.log.argtypes = [C.c_double]
    xe, xl = _args(11, 40000)
    e, _ = _cr(xe)
    _, l = _cr(xl)
    de = [i for i in range(len(xe)) if libm.exp(float(xe[i])) != e[i]]
    dl = [i for i in range(len(xl)) if libm.log(float(xl[i])) != l[i]]
    # rare: glibc's exp and log are within 0.52 ulp, almost always correctly rounded
    assert len(de) < 1e-2 * len(xe) and len(dl) < 1e-2 * len(xl)
    for i in de:
        assert float(mp.exp(mp.mpf(float(xe[i])))) == e[i], float(xe[i]).hex()
    for i in dl:
        assert float(mp.log(mp.mpf(float(xl[i])))) == l[i], float(xl[i]).hex()


def test_special_arguments_follow_the_library():
    import math
    x = np.array([0.0, -0.0, 1.0, -800.0, 800.0, np.inf, -np.inf, np.nan, 5e-324, -1.0, 710.0, -708.4])
    e, l = _cr(x)
    for xi, ei, li in zip(x, e, l):
        want_e = math.nan if np.isnan(xi) else math.exp(xi) if xi < 709.8 else math.inf
        assert (np.isnan(ei) and np.isnan(want_e)) or ei == want_e, (xi, ei)
        if xi > 0 and np.isfinite(xi):
            assert li == math.log(xi), (xi, li)
        elif xi == 0:
            assert li == -np.inf
        elif xi < 0 or np.isnan(xi):
            assert np.isnan(li)
