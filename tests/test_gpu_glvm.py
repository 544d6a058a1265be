"""The device Glv module (pp_glv_*, pp_gmix_*: Glvm.Model's Glv and model
functions of Gmix_model, gmix_model.ml:58-351) against the oracle's
restatement of the same functions (oracle/pp_oracle.c: or_evolve_into,
or_log_like3, or_slow_log_like3, or_masked_logdot, or_logdot,
or_perhaps_pull_exponent, or_ten_*_pairwise_prod), bit for bit where the
arithmetic has no transcendental (products, evolves, pulls, copies) and to
the last bits where device exp/log enter (P(t), logs).  Each operation
chains on GLVs that stay on the device.  Models: GTR+G4 (4 states) and a
20-state model with 4 rates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _model(oracle, S, rng):
    exch = rng.uniform(0.2, 2.0, (S, S))
    exch = (exch + exch.T) / 2
    np.fill_diagonal(exch, 0.0)
    statd = rng.uniform(0.5, 1.5, S)
    statd /= statd.sum()
    rates = oracle.discrete_gamma(4, 0.7)
    return exch, statd, rates


def _glv(rng, R, N, S, lo=1e-3):
    a = rng.uniform(lo, 1.0, (R, N, S))
    e = rng.integers(-40, 1, N).astype(np.int64)
    return a, e


@pytest.mark.parametrize("S,N", [(4, 300), (20, 120)])
def test_glv_module_matches_oracle(oracle, S, N):
    from pplacer_amd import placer
    from pplacer_amd.glvm import PP_FP_INFINITE, PP_FP_NAN, PP_FP_NORMAL, PP_FP_SUBNORMAL, PP_FP_ZERO
    from pplacer_amd.glvm import Glv, GmixModel
    _print_case("device Glv module, %d states, %d sites" % (S, N))
    rng = np.random.default_rng(31 + S)
    exch, statd, rates = _model(oracle, S, rng)
    om = oracle.Model(exch, statd, rates)
    pl = placer.Placer()
    try:
        pl.build_model(exch, statd, rates)
        m = GmixModel(pl)
        R = m.n_rates
        xa, xe = _glv(rng, R, N, S)
        ya, ye = _glv(rng, R, N, S)
        za, ze = _glv(rng, R, N, S)
        x, y, z = (Glv.of_arrays(pl._ctx, a, e) for a, e in ((xa, xe), (ya, ye), (za, ze)))
        assert x.dims == (R, N, S)
        # set / get / copy / mimic / memcpy / set_unit
        a, e = x.get()
        assert np.array_equal(a, xa) and np.array_equal(e, xe)
        c = x.copy()
        assert all(np.array_equal(u, v) for u, v in zip(c.get(), (xa, xe)))
        mm = x.mimic()
        assert mm.dims == x.dims and not mm.get()[0].any()
        mm.memcpy_from(y)
        assert all(np.array_equal(u, v) for u, v in zip(mm.get(), (ya, ye)))
        mm.set_unit()
        a, e = mm.get()
        assert (a == 1.0).all() and (e == 0).all()
        # Model.evolve_into, a chain of two on the device
        d1, d2 = x.mimic(), x.mimic()
        m.evolve_into(d1, x, 0.137)
        m.evolve_into(d2, d1, 0.021)
        wa, we = om.evolve(xa, xe, 0.137)
        wa, we = om.evolve(wa, we, 0.021)
        ga, ge = d2.get()
        assert np.array_equal(ge, we)
        assert np.max(np.abs(ga - wa) / np.abs(wa)) < 1e-14
        # Model.statd_pairwise_prod and Glv.pairwise_prod / listwise_prod
        sp = x.mimic()
        m.statd_pairwise_prod(sp, x, y)
        ga, ge = sp.get()
        assert np.array_equal(ga, oracle.ten_statd_pairwise_prod(statd, xa, ya)) and np.array_equal(ge, xe + ye)
        lp = x.mimic()
        lp.listwise_prod([x, y, z])
        ga, ge = lp.get()
        assert np.array_equal(ga, oracle.ten_pairwise_prod(oracle.ten_pairwise_prod(xa, ya), za))
        assert np.array_equal(ge, xe + ye + ze)
        lp.listwise_prod([y])
        assert all(np.array_equal(u, v) for u, v in zip(lp.get(), (ya, ye)))
        # Glv.perhaps_pull_exponent on GLVs small enough to pull at some sites
        sa = xa * np.exp2(-rng.integers(0, 80, N))[None, :, None]
        small = Glv.of_arrays(pl._ctx, sa, xe)
        small.perhaps_pull_exponent(-50)
        wa, we = oracle.perhaps_pull_exponent(-50, sa, xe)
        ga, ge = small.get()
        assert np.array_equal(ga, wa) and np.array_equal(ge, we) and (ge != xe).any()
        # Glv.mask_into
        keep = rng.random(N) < 0.6
        dst = Glv.make(pl._ctx, R, int(keep.sum()), S)
        x.mask_into(keep, dst)
        ga, ge = dst.get()
        assert np.array_equal(ga, xa[:, keep, :]) and np.array_equal(ge, xe[keep])
        # likelihoods
        util = np.empty(N)
        got = m.log_like3(x, y, z)
        want = om.log_like3((xa, xe), (ya, ye), (za, ze))
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        got = m.slow_log_like3(x, y, z)
        want = om.slow_log_like3((xa, xe), (ya, ye), (za, ze))
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        arr = m.site_log_like_arr3(x, y, z)
        assert arr.shape == (N,) and abs(arr.sum() - want) <= 1e-10 * abs(want)
        got = x.logdot(y)
        want = om.logdot((xa, xe), (ya, ye))
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        mask = (rng.random(N) < 0.7).astype(np.uint16)
        got = x.masked_logdot(y, mask)
        want = (oracle.ten_masked_logdot(xa, ya, mask)
                + np.log(2.0) * (float(xe[mask > 0].sum()) + float(ye[mask > 0].sum())))
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        got = x.bounded_logdot(y, 5, N - 3)
        want = (oracle.ten_bounded_logdot(xa, ya, 5, N - 3)
                + np.log(2.0) * (float(xe[5:N - 2].sum()) + float(ye[5:N - 2].sum())))
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        # summarize_post: per-site state sums over rates, then the caller's function
        post = x.summarize_post(lambda u: int(np.argmax(u)))
        assert post == [int(i) for i in np.argmax(xa.sum(axis=0), axis=1)]
        # make_constant_rate_glv_from_lv_arr
        lv = rng.uniform(0.0, 1.0, (N, S))
        cr = m.make_constant_rate_glv_from_lv_arr(lv)
        ga, ge = cr.get()
        assert np.array_equal(ga, np.broadcast_to(lv, (R, N, S))) and (ge == 0).all()
        # fp_classify: the worst class present, in OCaml's order
        assert x.fp_classify() == PP_FP_NORMAL
        for v, cls in ((5e-320, PP_FP_SUBNORMAL), (0.0, PP_FP_ZERO), (np.inf, PP_FP_INFINITE),
                       (np.nan, PP_FP_NAN)):
            a2 = xa.copy()
            a2[1, 7, 2] = v
            g2 = Glv.of_arrays(pl._ctx, a2, xe)
            assert g2.fp_classify() == cls
        # a site of zero likelihood is Gmix_model's failure
        z0 = za.copy()
        z0[:, 11, :] = 0.0
        with pytest.raises(Exception, match="Site 11 has zero likelihood"):
            m.site_log_like_arr3(x, y, Glv.of_arrays(pl._ctx, z0, ze))
    finally:
        pl.close()
