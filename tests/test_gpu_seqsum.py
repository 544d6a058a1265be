"""The FP64 engine's per-probe log sum is the stub's sequential sum
(linear_c.c:416-419: ll_tot += log(util[i]) in site order) computed by a whole
wave (wave_sequential_sum in pp_exact_sum.h).  Through the library's test
hook, every sum here must equal Python's left-to-right float sum -- the same
IEEE double additions in the same order -- bit for bit, for sequences shaped
like per-site logs and for adversarial ones: dyadic terms full of exact ties,
partial sums on and near powers of two, mixed signs and scales, zeros and -0,
non-finite terms, and lengths around the 64-lane blocking and past the
wave's 4,096-term limit.  The hook also reports whether the wave took its
parallel path, so the test checks that realistic sums do."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _seq(x):
    s = 0.0
    for v in x:
        s += float(v)
    return s


def _bits(v):
    return struct.pack("<d", v)


def _run(L, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(3)
    dp = C.POINTER(C.c_double)
    rc = L.pp_test_sequential_sum(x.ctypes.data_as(dp), len(x), out.ctypes.data_as(dp))
    assert rc == 0, L.pp_last_error()
    return out[0], out[1], int(out[2])


def _cases(rng):
    for n in (0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 300, 448, 960, 1600, 4096):
        yield "logs of utils in (0,1)", np.log(rng.uniform(1e-9, 1.0, n))
    for n in (200, 700):
        yield "site-log-like", np.log(rng.uniform(0.05, 0.95, n) ** rng.integers(1, 30, n))
        yield "utils up to 4", np.log(rng.uniform(1e-6, 4.0, n))
        yield "dyadic ties", -rng.integers(1, 9, n) * 0.125
        yield "powers of two", -np.ldexp(1.0, rng.integers(0, 6, n)) * np.where(rng.random(n) < 0.3, 1.0,
                                                                                rng.uniform(0.5, 1.5, n))
        yield "mixed signs and scales", (rng.random(n) - 0.7) * np.ldexp(1.0, rng.integers(-20, 20, n))
        x = np.log(rng.uniform(0.1, 0.9, n))
        x[n // 3] = 0.0
        x[n // 2] = -0.0
        yield "zeros", x
    # partial sums landing exactly on powers of two
    yield "exact binade edges", np.array([-0.5, -0.5, -1.0, -2.0, -4.0, -8.0, -16.0] * 40)
    x = np.log(rng.uniform(0.2, 0.8, 500))
    x[100] = -np.inf
    yield "minus infinity", x
    x = np.log(rng.uniform(0.2, 0.8, 500))
    x[250] = np.nan
    yield "nan", x
    yield "past the wave's limit", np.log(rng.uniform(1e-3, 1.0, 5000))


def test_wave_sequential_sum_is_the_sequential_sum():
    from pplacer_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(17)
    parallel = 0
    for name, x in _cases(rng):
        print("[gpu case] sequential sum: %s, n=%d" % (name, len(x)), flush=True)
        want = _seq(x)
        got, chain, n_ev = _run(L, x)
        if math.isnan(want):
            assert math.isnan(got) and math.isnan(chain), name
            continue
        assert _bits(chain) == _bits(want), (name, len(x), chain, want)
        assert _bits(got) == _bits(want), (name, len(x), got, want, n_ev)
        if name in ("logs of utils in (0,1)", "site-log-like") and 0 < len(x) <= 4096:
            assert n_ev > 0, (name, len(x))  # the parallel path, not the fallback
            parallel += 1
        if name == "past the wave's limit":
            assert n_ev == -1
    assert parallel >= 15


def test_wave_sequential_sum_random_sweep():
    """4,000 sums of per-site-log-like sequences, lengths 1-1,600."""
    from pplacer_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(23)
    print("[gpu case] sequential sum sweep", flush=True)
    fell_back = 0
    for t in range(4000):
        n = int(rng.integers(1, 1601))
        x = np.log(rng.uniform(1e-4, 1.0, n) * rng.uniform(0.0, 1.0, n) + 1e-300)
        want = _seq(x)
        got, _, n_ev = _run(L, x)
        assert _bits(got) == _bits(want), (t, n, got, want, n_ev)
        fell_back += n_ev < 0
    assert fell_back <= 4
