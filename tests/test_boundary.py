"""The drop-in boundary on the CPU: the C structs the ctypes mirror
declares have the header's layout (checked by a C program compiled
against include/gspan_amd.h), the C++ driver surface (include/
gspan_amd.hpp) compiles, and the command-line driver keeps the reference's
usage contract and exit codes (src/gspan_cuda/common_main.cpp:33-70)."""
import ctypes
import subprocess
from pathlib import Path

import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import build as gbuild

from helpers import GOLDEN

ROOT = Path(__file__).resolve().parent.parent


def _c_layout(tmp_path, struct, fields):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gspan_amd.h"', "int main(void) {",
             f'  printf("size %zu\\n", sizeof({struct}));']
    lines += [f'  printf("{f} %zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.splitlines())}


@pytest.mark.parametrize("struct,mirror", [("gm_cfg", gm.gm_cfg), ("gm_stats", gm.gm_stats)])
def test_ctypes_mirror_matches_the_header_layout(tmp_path, struct, mirror):
    fields = [f for f, _ in mirror._fields_]
    c = _c_layout(tmp_path, struct, fields)
    assert c["size"] == ctypes.sizeof(mirror)
    for f in fields:
        assert c[f] == getattr(mirror, f).offset, f


def test_cpp_driver_surface_compiles(tmp_path):
    """A reference-style driver (common_main.cpp:85-117 with namespaces
    changed) compiles against the header."""
    src = tmp_path / "driver.cpp"
    src.write_text('''#include "gspan_amd.hpp"
using namespace gspan_amd;
int main(int argc, char **argv) {
  graph_database database;
  database.load(argv[1], argv[2]);
  graph_counting_output out;
  cuda_configurator exec_config;
  exec_config.set_parameter("compute_support.boundaries", 128);
  gspan_cuda *gspan = get_gspan_cuda_instance();
  gspan->set_database(database, true);
  gspan->set_min_support(atoi(argv[3]));
  gspan->set_graph_output(&out);
  gspan->set_exec_configurator(&exec_config);
  gspan->run();
  gspan->delete_database_from_device();
  delete gspan;
  return (int)out.get_count() > 0 ? 0 : 1;
}
''')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-include", "cstdlib", f"-I{ROOT / 'include'}", str(src)],
                   check=True)


def test_cpp_in_memory_database_hand_over(tmp_path):
    """The in-memory hand-over a maintainer's binding uses: a database held
    as the reference's types::graph_database_t (graph_types.hpp:21-64:
    vertices with labels, each undirected edge in both endpoints' lists),
    flattened into gm_db_create's arrays with one entry per edge, reads back
    with the same counts and labels (host side only: no device copy is made
    before mining)."""
    src = tmp_path / "handover.cpp"
    src.write_text('''#include <cstdio>
#include "gspan_amd.hpp"
struct Edge { int from, to, elabel; };
struct Vertex { int label; std::vector<Edge> edge; };
using Graph = std::vector<Vertex>;
int main() {
  std::vector<Graph> db(2);
  db[0] = {{1, {{0, 1, 5}, {0, 2, 6}}}, {2, {{1, 0, 5}}}, {3, {{2, 0, 6}}}};
  db[1] = {{4, {{0, 1, 7}}}, {4, {{1, 0, 7}}}};
  std::vector<int32_t> vb{0}, vl, eb{0}, ef, et, el;
  for (const Graph &g : db) {
    for (const Vertex &v : g) {
      vl.push_back(v.label);
      for (const Edge &e : v.edge)
        if (e.from < e.to) { ef.push_back(e.from); et.push_back(e.to); el.push_back(e.elabel); }
    }
    vb.push_back((int32_t)vl.size());
    eb.push_back((int32_t)ef.size());
  }
  gspan_amd::graph_database gdb;
  gdb.create(vb, vl, eb, ef, et, el);
  int64_t c[6];
  gspan_amd::check(gm_db_counts(gdb.handle(), c), "counts");
  std::printf("%zu %lld %lld %lld\\n", gdb.size(), (long long)c[1], (long long)c[2], (long long)c[3]);
  try {
    gdb.create({0, 2}, {1, 1}, {0}, {}, {}, {});
  } catch (const std::runtime_error &) {
    std::printf("rejected\\n");
  }
  return 0;
}
''')
    exe = tmp_path / "handover"
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{gbuild.LIB_DIR}", "-lgspan_amd",
                    f"-Wl,-rpath,{gbuild.LIB_DIR}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    assert tuple(map(int, out[0].split())) == (2, 5, 3, 3)  # graphs, vertices, edges, largest graph
    assert out[1] == "rejected"


def _cli(*args):
    exe = gbuild.CLI
    if not exe.exists():
        gbuild.build_cli()
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_cli_usage_and_exit_codes():
    chem = GOLDEN / "Chemical_340.txt"
    assert _cli().returncode == 1
    assert _cli("txt", chem).returncode == 1
    assert _cli("nosuchtype", chem, 5).returncode == 1
    r = _cli("txt", chem, 340)  # support >= |D|
    assert r.returncode == 2 and "greater then the database size" in r.stderr
    r = _cli("-txt", chem, 0)
    assert r.returncode == 3 and "absolute_min_support < 1" in r.stderr


def test_tuned_configuration_file_reads_alike_in_python_and_cpp(tmp_path):
    """The autotuner's configuration file (the layout of the reference's
    cuda_configurator: header `id,<name>,...`, one row per id) reads back the
    same through tune.read_config and gspan_amd::cuda_configurator."""
    from gpu_graph_mining_amd import tune
    path = tmp_path / "tuned.csv"
    tune.write_config(path, "config2", {"chunk_rows": 128, "table_bits": 9, "canon_host": 1})
    tune.write_config(path, "config3", {"chunk_rows": 64, "table_bits": 8, "canon_host": 2})
    tune.write_config(path, "config2", {"chunk_rows": 256, "table_bits": 8, "canon_host": 0})  # replaced
    assert tune.read_config(path, "config2") == {"chunk_rows": 256, "table_bits": 8, "canon_host": 0}
    assert tune.read_config(path, "config3") == {"chunk_rows": 64, "table_bits": 8, "canon_host": 2}
    src = tmp_path / "cfg.cpp"
    src.write_text('''#include <cstdio>
#include "gspan_amd.hpp"
int main(int argc, char **argv) {
  gspan_amd::cuda_configurator cc(argv[1], argv[2]);
  cc.parse_config_file();
  gm_cfg cfg;
  cc.fill(cfg);
  std::printf("%d %d %d\\n", cfg.chunk_rows, cfg.table_bits, cfg.canon_host);
  try {
    gspan_amd::cuda_configurator(argv[1], "nosuch").parse_config_file();
  } catch (const std::runtime_error &) {
    std::printf("missing\\n");
  }
  return 0;
}
''')
    exe = tmp_path / "cfg"
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{gbuild.LIB_DIR}", "-lgspan_amd",
                    f"-Wl,-rpath,{gbuild.LIB_DIR}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(path), "config3"], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["64", "8", "2", "missing"]


def test_dfs_report_order_matches_the_sequential_gspan(tmp_path):
    """The sort key (Python) and comparator (C++) of the sequential gSpan's
    report order reproduce the oracle's emission order -- the order of
    src/gspan's output file -- from a shuffled list of its patterns."""
    import random
    from oracle import py_oracle as po
    db = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    pats, _ = po.mine_arrays(db.arrays(), 34)
    codes = [tuple(map(tuple, c)) for c, _, _ in pats]
    shuffled = codes[:]
    random.Random(3).shuffle(shuffled)
    assert sorted(shuffled, key=gm.dfs_order_key) == codes
    # the same with the C++ comparator
    src = tmp_path / "order.cpp"
    src.write_text('''#include <cstdio>
#include <iostream>
#include "gspan_amd.hpp"
int main() {
  std::vector<gspan_amd::DFSCode> codes;
  int n;
  while (std::cin >> n) {
    gspan_amd::DFSCode c;
    for (int i = 0; i < n; ++i) {
      gspan_amd::DFS d;
      std::cin >> d.from >> d.to >> d.fromlabel >> d.elabel >> d.tolabel;
      c.push_back(d);
    }
    codes.push_back(c);
  }
  std::stable_sort(codes.begin(), codes.end(), gspan_amd::dfs_less);
  for (const auto &c : codes) std::printf("%s\\n", c.to_string(false).c_str());
  return 0;
}
''')
    exe = tmp_path / "order"
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    inp = "\n".join(f"{len(c)} " + " ".join(" ".join(map(str, t)) for t in c) for c in shuffled) + "\n"
    out = subprocess.run([str(exe)], input=inp, check=True, capture_output=True, text=True).stdout.splitlines()
    assert out == [gm.DFSCode(c).to_string(False) for c in codes]
