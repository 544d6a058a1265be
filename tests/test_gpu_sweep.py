"""Parity on a wide sample: 2,000 queries at 50 taxa and 300 at 2,000 taxa,
every one compared with the oracle, in both forms.

The strict bar (1e-9 relative in FP64) is what the other GPU tests hold on
their cases.  Over a wide sample a few Brent runs part from the oracle's after
a last-bit difference in one probe: the GPU sums a probe's per-site logs in
wave order, the stub sequentially, and at a flat optimum one ulp can flip a
comparison.  Such a run still ends within the Brent tolerance.  This test
states how often that happens, from a measured sweep:

* every status and every best edge equals the oracle's (at 50 taxa the oracle
  fails 5 queries on the prox_bl assertion, and so does the GPU);
* FP64: at least 99.5% of placements within 1e-9 relative (measured: 1995 of
  2000 at 50 taxa, all 300 at 2,000 taxa), all within 1e-8 relative, branch
  lengths within 2e-4;
* FP32: all within 2e-4 absolute (measured maximum 1.02e-4), at least 99.9%
  within 1e-4, branch lengths within 2e-4.
"""
import pytest

from tests.helpers import oracle_setup

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _sweep(oracle, name, seed, nq):
    from pplacer_amd import placer, synth
    from pplacer_amd.placer import PP_FP32, PP_FP64
    _print_case("sweep workload %s seed %d, %d queries" % (name, seed, nq))
    wl = synth.make_workload(name, seed=seed, n_queries=nq)
    model, pre, core, queries = oracle_setup(oracle, wl)
    pl, q2, statd, rates = placer.setup_from_workload(wl)
    try:
        assert q2 == queries
        want = [core.place(q) for q in queries]
        out = {}
        for form, prec in (("fp64", PP_FP64), ("fp32", PP_FP32)):
            _print_case("sweep %s %s" % (name, form))
            got, status, _ = pl.place(queries, precision=prec)
            rel, ab, bl = [], [], []
            for q in range(nq):
                st, w, _ = want[q]
                assert int(status[q]) == st, (form, q)
                if st != 0:
                    continue
                g = got[q][0]
                assert g.edge_num == w[0][0], (form, q)
                ab.append(abs(g.likelihood - w[0][2]))
                rel.append(ab[-1] / abs(w[0][2]))
                bl.append(max(abs(g.pendant_length - w[0][4]), abs(g.distal_length - w[0][3])))
            out[form] = (rel, ab, bl)
        return out
    finally:
        pl.close()


@pytest.mark.parametrize("name,seed,nq", [("gtr50", 21, 2000), ("gtr2000", 22, 300)])
def test_wide_sample_against_the_oracle(oracle, name, seed, nq):
    out = _sweep(oracle, name, seed, nq)
    rel, ab, bl = out["fp64"]
    assert sum(r <= 1e-9 for r in rel) >= 0.995 * len(rel)
    assert max(rel) <= 1e-8 and max(bl) <= 2e-4
    rel, ab, bl = out["fp32"]
    assert max(ab) <= 2e-4 and max(bl) <= 2e-4
    assert sum(a <= 1e-4 for a in ab) >= 0.999 * len(ab)
