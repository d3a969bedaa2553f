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
