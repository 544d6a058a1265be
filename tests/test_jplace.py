"""jplace output and query de-duplication (host side, no GPU)."""
import glob
import json
import os
import re

import pytest

from tests.helpers import GOLDEN

FILES = sorted(glob.glob(os.path.join(GOLDEN, "jplace", "*", "*.jplace")))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_jplace_roundtrip(tmp_path, path):
    # tests/pplacer/test_placefile.ml:1-24: write then read equals the original
    # (placement_equal compares floats with =~, absolute 1e-5)
    from pplacer_amd import jplace
    from pplacer_amd.placer import PlacementRecord
    tree_s, fields, pqs = jplace.read(path)
    recs = []
    for namlom, pl in pqs:
        recs.append((namlom, [PlacementRecord(int(p["edge_num"]), p["like_weight_ratio"],
                                              p["likelihood"], p["distal_length"],
                                              p["pendant_length"]) for p in pl]))
    from pplacer_amd import refio
    # version-1 trees mark edges with [n]; parse the topology and ids
    tree = refio.parse_newick(re.sub(r"\[(\d+)\]", r"{\1}", tree_s))
    out = tmp_path / "rt.jplace"
    jplace.write(str(out), tree, recs)
    tree2, fields2, pqs2 = jplace.read(str(out))
    assert fields2 == jplace.FIELDS
    assert len(pqs2) == len(pqs)
    for (nl1, p1), (nl2, p2) in zip(pqs, pqs2):
        assert [n for n, _ in nl1] == [n for n, _ in nl2]
        assert all(abs(a - b) < 1e-5 for (_, a), (_, b) in zip(nl1, nl2))
        assert len(p1) == len(p2)
        for a, b in zip(p1, p2):
            assert a["edge_num"] == b["edge_num"]
            for f in ("likelihood", "like_weight_ratio", "distal_length", "pendant_length"):
                assert abs(a[f] - b[f]) < 1e-5
    # the written tree keeps every node id
    ids = sorted(int(x) for x in re.findall(r"\{(\d+)\}", tree2))
    assert ids == list(range(tree.n_nodes))


def test_float_format_is_12g(tmp_path):
    from pplacer_amd import jplace, refio
    from pplacer_amd.placer import PlacementRecord
    tree = refio.parse_newick("((A:0.1,B:0.2):0.3,C:0.4);")
    rec = PlacementRecord(1, 0.123456789012345, -1234.56789012345, 1e-7 / 3, 0.5)
    out = tmp_path / "f.jplace"
    jplace.write(str(out), tree, [([("q1", 1.0)], [rec])])
    text = out.read_text()
    assert "0.123456789012" in text and "-1234.56789012" in text and "3.33333333333e-08" in text
    obj = json.loads(text)
    assert obj["version"] == 3 and obj["fields"] == sorted(obj["fields"])
    assert obj["tree"] == "((A:0.1{0},B:0.2{1}):0.3{2},C:0.4{3}):0{4};"


def test_dedup_redup():
    from pplacer_amd import jplace
    qs = [("a", "ACGT"), ("b", "AC-T"), ("c", "ACGT"), ("d", "AC-T"), ("e", "TTTT")]
    uniq, groups = jplace.dedup(qs)
    assert uniq == [("a", "ACGT"), ("b", "AC-T"), ("e", "TTTT")]
    assert groups == {"a": ["a", "c"], "b": ["b", "d"], "e": ["e"]}
    assert jplace.redup([("a", 1.0)], groups) == [("a", 1.0), ("c", 1.0)]
