"""Query de-duplication and jplace output.

* dedup / redup: identical query sequences are placed once and their names
  re-expanded afterwards (pplacer_src/pplacer_run.ml:228-253,
  pplacer_src/placerun.ml:95-110).
* write / read: the jplace format of pplacer_src/placerun_io.ml:58-89 and
  pplacer_src/pquery_io.ml:28-45 -- placements sorted by like_weight_ratio
  (pquery.ml:86-92), fields in sorted order, floats printed with %.12g
  (json_src/json.ml:42), the reference tree with {edge} numbers.
"""
from __future__ import annotations

import json

from . import refio

FIELDS = ["distal_length", "edge_num", "like_weight_ratio", "likelihood", "pendant_length"]


def dedup(queries):
    """queries: [(name, seq)] -> (unique [(rep_name, seq)], groups {rep_name: [names]}).
    Keys on the exact sequence string; the representative is the first name
    seen for that sequence."""
    first = {}
    groups = {}
    unique = []
    for name, seq in queries:
        if seq in first:
            groups[first[seq]].append(name)
        else:
            first[seq] = name
            groups[name] = [name]
            unique.append((name, seq))
    return unique, groups


def redup(namlom, groups):
    """[(name, mass)] of a representative -> every duplicate's name with that mass."""
    out = []
    for n, m in namlom:
        if n in groups:
            out.extend((g, m) for g in groups[n])
        else:
            out.append((n, m))
    return out


def _fmt(x) -> str:
    if isinstance(x, bool):
        return "true" if x else "false"
    if isinstance(x, int):
        return "%d" % x
    if isinstance(x, float):
        if x != x:
            return "nan"
        if x in (float("inf"), float("-inf")):
            return "inf" if x > 0 else "-inf"
        return "%.12g" % x
    if isinstance(x, str):
        return json.dumps(x)
    if isinstance(x, (list, tuple)):
        return "[" + ", ".join(_fmt(v) for v in x) + "]"
    if isinstance(x, dict):
        return "{" + ", ".join("%s: %s" % (json.dumps(k), _fmt(v)) for k, v in x.items()) + "}"
    if x is None:
        return "null"
    raise TypeError(type(x))


def tree_string(tree: refio.Tree) -> str:
    """Newick_gtree.to_string ~with_node_numbers:true: name, :%g length, {id}."""
    def rec(i):
        s = ""
        if tree.children[i]:
            s = "(" + ",".join(rec(c) for c in tree.children[i]) + ")"
        if tree.names[i]:
            s += tree.names[i]
        s += ":%g" % tree.bl[i]
        s += "{%d}" % i
        return s

    return rec(tree.root) + ";"


PP_FIELDS = ["marginal_like", "post_prob"]


def to_json(tree: refio.Tree, pqueries, invocation: str = "pplacer_amd") -> dict:
    """pqueries: [(namlom, [PlacementRecord sorted by like_weight_ratio])].
    Fields are the sorted union of the placements' fields
    (placerun_io.ml:77-83): post_prob and marginal_like join them when any
    placement carries posterior probabilities (placement.ml:176-180)."""
    has_pp = any(r.post_prob is not None for _, pl in pqueries for r in pl)
    fields = sorted(FIELDS + (PP_FIELDS if has_pp else []))
    placements = []
    for namlom, pl in pqueries:
        p = [[getattr(r, f) for f in fields] for r in pl]
        placements.append({"p": p, "nm": [[n, float(m)] for n, m in namlom]})
    return {"tree": tree_string(tree), "placements": placements,
            "metadata": {"invocation": invocation}, "version": 3, "fields": fields}


def write(path: str, tree: refio.Tree, pqueries, invocation: str = "pplacer_amd") -> None:
    with open(path, "w") as f:
        f.write(_fmt(to_json(tree, pqueries, invocation)))
        f.write("\n")


def read(path: str):
    """Returns (tree_string, fields, [(namlom, [dict(field -> value)])]) for
    jplace versions 1-3 (version 1 names are "n", later "nm" with masses)."""
    with open(path) as f:
        obj = json.load(f)
    fields = obj["fields"]
    out = []
    for pq in obj["placements"]:
        if "nm" in pq:
            namlom = [(n, float(m)) for n, m in pq["nm"]]
        else:
            names = pq["n"] if isinstance(pq["n"], list) else [pq["n"]]
            namlom = [(n, 1.0) for n in names]
        out.append((namlom, [dict(zip(fields, p)) for p in pq["p"]]))
    return obj["tree"], fields, out
