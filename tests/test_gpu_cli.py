"""The command-line driver on the GPU: the reference's contract end to end
(src/gspan_cuda/common_main.cpp:85-117).  Its pattern file holds the
reference's patterns (the golden listing of the same database and
support), and its [VAL] lines read back with the field rule of the
reference's log parser (scripts/parse_log.awk:3-17: split on '[', field 7,
then on '=' / ']')."""
import subprocess

import pytest

from gpu_graph_mining_amd import build as gbuild
from oracle import py_oracle as po

from helpers import GOLDEN

pytestmark = pytest.mark.gpu


def _val(stdout, name):
    for line in stdout.splitlines():
        fields = line.split("[")
        if "[VAL]" in line and len(fields) > 6:
            parts = fields[6].replace("]", "=").split("=")
            if parts[0] == name:
                return parts[1]
    return None


@pytest.mark.parametrize("name", ["chem_68", "labeled_3"])
def test_cli_mines_the_reference_patterns(tmp_path, golden, name):
    case = golden[name]
    out = tmp_path / "patterns.txt"
    r = subprocess.run([str(gbuild.CLI), case["db"]["filetype"], str(GOLDEN / case["db"]["file"]), str(case["minsup"]),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert _val(r.stdout, "total_frequent_graphs") == str(case["n_patterns"])
    assert float(_val(r.stdout, "total_time")) >= 0
    assert _val(r.stdout, "absolute_min_support") == str(case["minsup"])
    # fully labelled to_string(false) lines; in the sequential gSpan's -1
    # convention they are the golden codes
    codes = []
    for line in out.read_text().splitlines():
        code = tuple(tuple(int(x) for x in t.strip("()").split()) for t in line.split(";"))
        codes.append(po.code_str(po.normalize(code)))
    assert sorted(codes) == sorted(l.split("\t")[0] for l in case["lines"])


def test_cli_pattern_file_in_the_sequential_gspan_order(tmp_path, golden):
    """The CLI writes the patterns in the sequential gSpan's report order:
    in its -1 label convention the file lists the oracle's codes in the
    oracle's emission order (the order of src/gspan's output file)."""
    import gpu_graph_mining_amd as gm
    case = golden["chem_34"]
    out = tmp_path / "patterns.txt"
    r = subprocess.run([str(gbuild.CLI), "txt", str(GOLDEN / "Chemical_340.txt"), str(case["minsup"]), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = []
    for line in out.read_text().splitlines():
        code = tuple(tuple(int(x) for x in t.strip("()").split()) for t in line.split(";"))
        got.append(po.code_str(po.normalize(code)))
    db = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    opats, _ = po.mine_arrays(db.arrays(), case["minsup"])
    assert got == [po.code_str(c) for c, _, _ in opats]


def test_python_file_sink_in_order_is_the_oracle_file(tmp_path, golden):
    """gspan_cuda + graph_file_output(dfs_order=True) with wildcard labels:
    byte for byte the sequential gSpan's output file (the oracle's lines in
    its emission order)."""
    import gpu_graph_mining_amd as gm
    case = golden["chem_68"]
    db = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    path = tmp_path / "p.txt"
    sink = gm.graph_file_output(path, dfs_order=True)
    g = gm.gspan_cuda()
    g.set_database(db)
    g.set_min_support(case["minsup"])
    g.set_graph_output(sink)
    g.set_exec_configurator({"label_style": gm.GM_LABELS_WILDCARD})
    g.run()
    sink.close()
    opats, _ = po.mine_arrays(db.arrays(), case["minsup"])
    assert path.read_text() == "".join(gm.DFSCode(c).to_string(False) + "\n" for c, _, _ in opats)
