"""Shared helpers of the test suite: fixture databases, listings, brute force."""
from __future__ import annotations

import hashlib
import itertools
import json
from pathlib import Path
from typing import Dict, List, Sequence, Tuple

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl
from oracle import py_oracle as po

GOLDEN = Path(__file__).resolve().parent / "golden"


def listing(patterns) -> List[str]:
    """Sorted "code<TAB>support" lines, code in the sequential gSpan's -1
    convention (what the reference CLI prints, src/globals/graph_output.cpp:
    145-150, plus the support)."""
    return sorted(f"{po.code_str(po.normalize(c))}\t{s}" for c, s, *_ in patterns)


def listing_sha(lines: Sequence[str]) -> str:
    blob = ("\n".join(lines) + "\n").encode() if lines else b""
    return hashlib.sha256(blob).hexdigest()


def check_listing(lines: Sequence[str], case: dict) -> None:
    assert len(lines) == case["n_patterns"], (len(lines), case["n_patterns"], case["name"])
    if "lines" in case:
        assert list(lines) == case["lines"], case["name"]
    assert listing_sha(lines) == case["sha256"], case["name"]


def case_db(case: dict) -> gm.GraphDB:
    spec = case["db"]
    if "file" in spec:
        return gm.GraphDB.load(GOLDEN / spec["file"], spec["filetype"])
    if "codes" in spec:
        codes = json.loads((GOLDEN / "test_support_dbs.json").read_text())[spec["codes"]]
        return gm.GraphDB.from_codes(codes)
    g = spec["gen"]
    if g == "config0":
        db, _ = wl.config0(seed=spec["seed"], n_graphs=spec["n_graphs"])
    elif g == "config1":
        db, _ = wl.config1(seed=spec["seed"], n_graphs=spec["n_graphs"])
    elif g == "config2":
        db, _ = wl.config2(seed=spec["seed"], n_graphs=spec["n_graphs"])
    elif g == "config4":
        db = wl.config4(seed=spec["seed"], n_graphs=spec["n_graphs"])
    elif g == "config3":
        db = wl.config3_jobs(seed=spec["seed"])[spec["job"]].database()
    else:
        raise KeyError(g)
    if "db_sha256" in case:
        assert db.sha256() == case["db_sha256"], "generator output changed"
    return db


# ------------------------------------------------------------ brute force
def monomorphisms(pattern_labels, pattern_edges, glabels, gadj) -> int:
    """Number of injective maps of the pattern into the graph that keep
    vertex labels and map every pattern edge onto a graph edge with the same
    edge label.  gadj: dict (u, v) -> elabel for both orientations."""
    n = len(pattern_labels)
    nbrs = [[] for _ in range(n)]
    for a, b, l in pattern_edges:
        nbrs[a].append((b, l))
        nbrs[b].append((a, l))
    order = list(range(n))
    count = 0
    assign = [-1] * n
    used = set()

    def rec(i):
        nonlocal count
        if i == n:
            count += 1
            return
        v = order[i]
        for x in range(len(glabels)):
            if x in used or glabels[x] != pattern_labels[v]:
                continue
            ok = True
            for w, l in nbrs[v]:
                if assign[w] >= 0 and gadj.get((x, assign[w])) != l:
                    ok = False
                    break
            if ok:
                assign[v] = x
                used.add(x)
                rec(i + 1)
                used.discard(x)
                assign[v] = -1

    rec(0)
    return count


def code_graph(code) -> Tuple[List[int], List[Tuple[int, int, int]]]:
    nv = max(max(d[0], d[1]) for d in code) + 1
    labels = [0] * nv
    for d in code:
        if d[2] != -1:
            labels[d[0]] = d[2]
        if d[4] != -1:
            labels[d[1]] = d[4]
    return labels, [(d[0], d[1], d[3]) for d in code]


def canonical_form(labels, edges) -> Tuple:
    """Isomorphism-invariant form of a small labelled graph: the smallest
    (vertex label sequence, sorted labelled edge list) over all vertex
    permutations."""
    n = len(labels)
    best = None
    for perm in itertools.permutations(range(n)):
        inv = [0] * n
        for i, p in enumerate(perm):
            inv[p] = i
        lab = tuple(labels[perm[i]] for i in range(n))
        if best is not None and lab > best[0]:
            continue
        es = tuple(sorted((min(inv[a], inv[b]), max(inv[a], inv[b]), l) for a, b, l in edges))
        cand = (lab, es)
        if best is None or cand < best:
            best = cand
    return best


def brute_force_frequent(graphs, minsup: int, max_edges: int) -> Dict[Tuple, int]:
    """All connected labelled subgraphs with <= max_edges edges and support
    >= minsup, by enumerating connected edge subsets of every graph:
    canonical form -> support."""
    seen_in: Dict[Tuple, set] = {}
    for gi, (labels, edges) in enumerate(graphs):
        m = len(edges)
        inc = [[] for _ in labels]
        for k, (a, b, _) in enumerate(edges):
            inc[a].append(k)
            inc[b].append(k)
        found = set()
        frontier = {frozenset([k]) for k in range(m)}
        for size in range(1, max_edges + 1):
            for s in frontier:
                verts = sorted({x for k in s for x in edges[k][:2]})
                idx = {v: i for i, v in enumerate(verts)}
                sub_e = [(idx[edges[k][0]], idx[edges[k][1]], edges[k][2]) for k in s]
                found.add(canonical_form([labels[v] for v in verts], sub_e))
            if size == max_edges:
                break
            nxt = set()
            for s in frontier:
                verts = {x for k in s for x in edges[k][:2]}
                for v in verts:
                    for k in inc[v]:
                        if k not in s:
                            nxt.add(s | {k})
            frontier = nxt
        for f in found:
            seen_in.setdefault(f, set()).add(gi)
    return {f: len(g) for f, g in seen_in.items() if len(g) >= minsup}
