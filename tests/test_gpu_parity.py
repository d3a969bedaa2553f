"""GPU parity tests: the HIP path (through the C ABI) against the reference's
golden output and the CPU oracle.  Bit-exact: same set of frequent DFS
codes, same supports, same |Proj(P)| per pattern, same E_child / E_cand."""
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl
from oracle import py_oracle as po

from helpers import case_db, check_listing, listing

pytestmark = pytest.mark.gpu


def _by_code(patterns):
    return {po.normalize(c): (s, n) for c, s, n in patterns}


def test_gpu_matches_reference_on_every_golden_case(golden):
    for name, case in golden.items():
        if "oracle" in case:  # the full-size workloads have tests of their own
            continue
        print("case", name, flush=True)
        db = case_db(case)
        pats, st = gm.mine(db, case["minsup"])
        check_listing(listing(pats), case)
        assert st["n_patterns"] == case["n_patterns"]


@pytest.mark.parametrize("name", ["chem_34", "chem_17", "config1_1k", "config4_2k", "get_labeled_test_database_2",
                                  "labeled_2"])
def test_gpu_embedding_counts_match_oracle(golden, name):
    """|Proj(P)| of every pattern, E_child and E_cand equal the oracle's."""
    case = golden[name]
    db = case_db(case)
    pats, st = gm.mine(db, case["minsup"], collect_proj=True)
    opats, ost = po.mine_arrays(db.arrays(), case["minsup"])
    assert _by_code(pats) == _by_code(opats)
    assert st["e_child"] == ost["e_child"]
    assert st["e_root"] == ost["e_root"]
    assert st["e_cand"] == ost["e_cand"]


def test_gpu_wildcard_labels_equal_oracle_strings(golden):
    case = golden["chem_68"]
    db = case_db(case)
    pats, _ = gm.mine(db, case["minsup"], label_style=gm.GM_LABELS_WILDCARD)
    opats, _ = po.mine_arrays(db.arrays(), case["minsup"])
    assert sorted(po.code_str(c) for c, _, _ in pats) == sorted(po.code_str(c) for c, _, _ in opats)


def test_gpu_reported_codes_are_minimal_and_anti_monotone(golden):
    case = golden["chem_17"]
    db = case_db(case)
    pats, _ = gm.mine(db, case["minsup"])
    sup = {tuple(c): s for c, s, _ in pats}
    for c, s, _ in pats:
        assert s >= case["minsup"]
        if len(c) > 1:
            assert tuple(c[:-1]) in sup and sup[tuple(c[:-1])] >= s  # the parent was reported
    for c, _, _ in pats[:300]:
        assert gm.DFSCode(c).is_min()


def test_gpu_batch_equals_individual_runs(golden):
    names = [n for n in golden if n.startswith("config3_job")] + ["chem_68", "config0"]
    dbs = [case_db(golden[n]) for n in names]
    mins = [golden[n]["minsup"] for n in names]
    res, st = gm.mine_batch(dbs, mins, collect_proj=True)
    for n, db, ms, got in zip(names, dbs, mins, res):
        print("job", n, flush=True)
        check_listing(listing(got), golden[n])
        alone, _ = gm.mine(db, ms, collect_proj=True)
        assert _by_code(got) == _by_code(alone)
    assert st["n_patterns"] == sum(golden[n]["n_patterns"] for n in names)


def test_gpu_many_graphs_equal_oracle():
    """20,000 graphs (more than one placement-scan workgroup covers): the
    root level's children are counted and written over vertex waves too;
    patterns, supports and embedding counts equal the oracle's."""
    db, _ = wl.config0(seed=5, n_graphs=20000)
    ms = 200
    pats, st = gm.mine(db, ms, collect_proj=True)
    opats, ost = po.mine_arrays(db.arrays(), ms)
    assert len(opats) > 4000
    assert _by_code(pats) == _by_code(opats)
    assert (st["e_child"], st["e_root"], st["e_cand"]) == (ost["e_child"], ost["e_root"], ost["e_cand"])


def test_gpu_deterministic_across_runs(golden):
    case = golden["config1_1k"]
    db = case_db(case)
    a, sa = gm.mine(db, case["minsup"], collect_proj=True)
    b, sb = gm.mine(db, case["minsup"], collect_proj=True)
    assert a == b
    assert (sa["e_child"], sa["e_cand"]) == (sb["e_child"], sb["e_cand"])


def test_gpu_edge_cases():
    empty = gm.GraphDB.from_arrays([0], [], [0], [], [], [])
    assert gm.mine(empty, 1)[0] == []
    isolated = gm.GraphDB.from_graphs([([0, 1, 2], []), ([1], [])])
    assert gm.mine(isolated, 1)[0] == []
    tri = ([0, 0, 0], [(0, 1, 0), (1, 2, 0), (2, 0, 0)])
    db = gm.GraphDB.from_graphs([tri, tri, ([0, 0], [(0, 1, 0)])])
    assert len(gm.mine(db, 4)[0]) == 0  # min support above |D|
    got, _ = gm.mine(db, 3)  # min support == |D|: only the single edge
    assert listing(got) == ["(0 1 0 0 0)\t3"]
    got, _ = gm.mine(db, 1)
    opats, _ = po.mine_arrays(db.arrays(), 1)
    assert listing(got) == listing(opats)


def test_gpu_largest_graph_and_max_edges():
    """A 256-vertex graph (the largest the byte-wide local ids allow) with
    patterns capped at 4 edges equals the oracle's patterns of <= 4 edges."""
    import random
    rng = random.Random(5)
    n = 256
    edges = {(rng.randrange(v), v): rng.randrange(2) for v in range(1, n)}
    labels = [rng.randrange(3) for _ in range(n)]
    graphs = [(labels, [(a, b, l) for (a, b), l in edges.items()])]
    graphs.append(([rng.randrange(3) for _ in range(40)], [(rng.randrange(v), v, 1) for v in range(1, 40)]))
    db = gm.GraphDB.from_graphs(graphs)
    got, _ = gm.mine(db, 1, max_edges=4, collect_proj=True)
    opats, _ = po.mine_arrays(db.arrays(), 1, max_edges=4)
    assert _by_code(got) == _by_code(opats)


def test_gpu_callback_can_stop_mining(golden):
    case = golden["chem_68"]
    db = case_db(case)
    seen = []

    def stop_after_five(code, sup, nproj):
        seen.append(code)
        return len(seen) < 5  # False on the fifth pattern stops mining

    with pytest.raises(gm.GMError) as e:
        gm.mine(db, case["minsup"], on_pattern=stop_after_five)
    assert e.value.code == gm.GM_E_CALLBACK
    assert len(seen) == 5


def test_gpu_driver_class_api(golden, tmp_path):
    """The reference call sequence (common_main.cpp:91-108) on the GPU."""
    case = golden["chem_68"]
    g = gm.gspan_cuda()
    g.set_database(case_db(case))
    g.set_min_support(case["minsup"])
    out = gm.graph_file_output(tmp_path / "out.txt")
    g.set_graph_output(out)
    g.set_exec_configurator({"label_style": gm.GM_LABELS_WILDCARD})
    g.run()
    out.close()
    lines = (tmp_path / "out.txt").read_text().splitlines()
    assert sorted(lines) == sorted(l.split("\t")[0] for l in case["lines"])
    assert out.get_count() == case["n_patterns"]


@pytest.mark.parametrize("name", ["config1_1k", "chem_34", "config4_2k"])
def test_gpu_write_paths_agree(golden, name):
    """The write kernel's two sources of children -- the candidate lists the
    support kernel stored, and a fresh enumeration -- give the same patterns,
    supports and |Proj|; so do the fixed workgroup sizes.  The first call on
    a database runs with a cold candidate store (some graphs fall back to
    enumeration), the later calls with a warm one."""
    case = golden[name]
    db = case_db(case)
    runs = [gm.mine(db, case["minsup"], collect_proj=True),  # cold store
            gm.mine(db, case["minsup"], collect_proj=True),  # warm store
            gm.mine(db, case["minsup"], collect_proj=True, candidate_lists=False)]
    runs += [gm.mine(db, case["minsup"], collect_proj=True, block_threads=b) for b in (64, 128, 256)]
    want = _by_code(runs[0][0])
    check_listing(listing(runs[0][0]), case)
    for pats, st in runs[1:]:
        assert _by_code(pats) == want
        assert (st["e_child"], st["e_cand"], st["e_root"]) == (runs[0][1]["e_child"], runs[0][1]["e_cand"],
                                                               runs[0][1]["e_root"])


@pytest.mark.parametrize("name", ["config1_1k", "chem_34", "config4_2k", "labeled_2"])
def test_gpu_path_pruning_same_results(golden, name):
    """Dropping the candidates whose new edge forms an infrequent 2-edge path
    with an existing edge changes no pattern, support or |Proj|, and the
    dropped candidates still count in E_cand (the oracle's push count)."""
    case = golden[name]
    db = case_db(case)
    on, st_on = gm.mine(db, case["minsup"], collect_proj=True)
    off, st_off = gm.mine(db, case["minsup"], collect_proj=True, path_pruning=False)
    assert _by_code(on) == _by_code(off)
    check_listing(listing(on), case)
    assert (st_on["e_child"], st_on["e_cand"], st_on["e_root"]) == (st_off["e_child"], st_off["e_cand"],
                                                                    st_off["e_root"])
    assert st_off["pruned"] == 0
    assert 0 <= st_on["pruned"] <= st_on["e_cand"]


@pytest.mark.parametrize("name", ["config1_1k", "chem_34", "config4_2k"])
def test_gpu_speculation_modes_agree(golden, name):
    """Children written before their canonical check (non-minimal ones
    dropped at the next level), written after it, or chosen per level: the
    same patterns, supports, |Proj|, E_cand and E_child."""
    case = golden[name]
    db = case_db(case)
    runs = [gm.mine(db, case["minsup"], collect_proj=True, speculate=m) for m in (0, 1, 2)]
    want = _by_code(runs[0][0])
    check_listing(listing(runs[0][0]), case)
    for pats, st in runs[1:]:
        assert _by_code(pats) == want
        assert (st["e_child"], st["e_cand"], st["e_root"], st["n_patterns"]) == (
            runs[0][1]["e_child"], runs[0][1]["e_cand"], runs[0][1]["e_root"], runs[0][1]["n_patterns"])


def test_gpu_bad_block_threads_rejected(golden):
    db = case_db(golden["chem_17"])
    with pytest.raises(gm.GMError) as e:
        gm.mine(db, 5, block_threads=96)
    assert e.value.code == gm.GM_E_ARG


def test_gpu_callback_exception_propagates(golden):
    """An exception raised by a pattern sink stops mining and reaches the
    caller (ctypes would otherwise print it and carry on)."""
    case = golden["chem_68"]
    db = case_db(case)
    seen = []

    def sink(code, sup, nproj):
        seen.append(code)
        if len(seen) == 3:
            raise KeyError("sink failed")

    with pytest.raises(KeyError):
        gm.mine(db, case["minsup"], on_pattern=sink)
    assert len(seen) == 3


def test_gpu_recount_of_full_count_tables():
    """Stars of 150 leaves with distinct labels: a center row lists 150
    distinct keys, more than a 128-entry count table holds, so root chunks
    and the 2-edge level's chunks overflow and are recounted (k_level_fix
    with key-residue split passes; k_write counts their children itself).
    Patterns, supports, |Proj|, E_child and E_cand equal the oracle's.
    (Recounts of chunks longer than one part run on config2's full
    workload, whose listing is checked against the reference's.)"""
    vb, vl, eb, ef, et, el = [0], [], [0], [], [], []
    for g in range(40):
        leaves = 150 - (g % 3)
        vl += [0] + list(range(1, leaves + 1))
        vb.append(len(vl))
        ef += [0] * leaves
        et += list(range(1, leaves + 1))
        el += [v % 2 for v in range(leaves)]
        eb.append(len(ef))
    db = gm.GraphDB.from_arrays(vb, vl, eb, ef, et, el)
    pats, st = gm.mine(db, 30, collect_proj=True, max_edges=2, table_bits=7)
    opats, ost = po.mine_arrays(db.arrays(), 30, max_edges=2)
    assert st["split_passes"] > 0  # chunks were recounted
    assert _by_code(pats) == _by_code(opats)
    assert st["e_child"] == ost["e_child"] and st["e_cand"] == ost["e_cand"]


def test_gpu_candidate_counts_closed_form():
    """The device enumerates the hand-counted candidates: none for an edge
    a-b with labels 0 and 1; 18 for a triangle with equal labels (12 at
    the root's children, 6 at the path's), with E_child 12 and E_root 6."""
    db = gm.GraphDB.from_arrays([0, 2], [0, 1], [0, 1], [0], [1], [0])
    pats, st = gm.mine(db, 1)
    assert len(pats) == 1 and (st["e_cand"], st["e_child"], st["e_root"]) == (0, 0, 1)
    db = gm.GraphDB.from_arrays([0, 3], [0, 0, 0], [0, 3], [0, 1, 0], [1, 2, 2], [0, 0, 0])
    pats, st = gm.mine(db, 1)
    assert len(pats) == 3 and (st["e_cand"], st["e_child"], st["e_root"]) == (18, 12, 6)
