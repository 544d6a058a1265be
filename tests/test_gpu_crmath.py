"""The device build of the correctly rounded FP64 exp and log
(pplacer_amd/csrc/pp_crmath.h) against its host build, which
tests/test_crmath_cpu.py checks against mpmath: bit-equal on a million
arguments over the ranges the kernels use (transition-matrix exponents,
site likelihoods) and near the table boundaries."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(fn, x):
    dp = C.POINTER(C.c_double)
    e, l = np.empty_like(x), np.empty_like(x)
    assert fn(len(x), x.ctypes.data_as(dp), e.ctypes.data_as(dp), l.ctypes.data_as(dp)) == 0
    return e, l


def test_device_exp_log_equal_the_host_build():
    from pplacer_amd._lib import lib
    print("\n[gpu case] correctly rounded exp/log, device vs host", flush=True)
    rng = np.random.default_rng(5)
    n = 250000
    x = np.concatenate([rng.uniform(-60, 0, n), rng.uniform(-708, 709.6, n // 4), rng.normal(0, 1e-3, n // 4),
                        np.exp(rng.uniform(-700, 700, n // 2)), rng.uniform(0.98, 1.02, n // 2),
                        10 ** rng.uniform(-30, 0, n),
                        np.arange(-200, 200) * (np.log(2) / 128),
                        np.array([0.0, 1.0, -800.0, 800.0, np.inf, -np.inf, np.nan, 5e-324, -1.0])])
    dev = _run(lib().pp_test_cr_math, x)
    host = _run(lib().pp_test_cr_math_host, x)
    for d, h in zip(dev, host):
        same = (d == h) | (np.isnan(d) & np.isnan(h))
        assert same.all(), [(float(a).hex(), float(b).hex(), float(c).hex()) for a, b, c in
                            zip(x[~same][:5], d[~same][:5], h[~same][:5])]
