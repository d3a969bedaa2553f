"""Write the golden fixtures of tests/golden/ from the reference build.

Runs the reference's own sequential gSpan (src/gspan + src/globals compiled
in place by oracle/Makefile.ref into oracle/_ref/gspan_ref, with the
support-recording sink of oracle/ref_harness.cpp) on:
  * testdata/Chemical_340 (copied here as Chemical_340.txt) at the supports
    the survey measured;
  * the reference's .bin fixtures labeled.bin / unlabeled.bin;
  * the three 18-graph databases of src/globals/test_support.cpp:10-153
    (their code strings are in test_support_dbs.json);
  * seeded instances of the benchmark configurations built by this
    repository's generators (their database sha256 is recorded).
Each case stores the sorted "code<TAB>support" lines the reference emits
(code in its -1 convention) in full when small, else their sha256 and count.

The full-size planted-pattern case (100,000 transactions, 1 %) also records
the oracle's embedding and byte counts (E_child, E_cand, E_root, B_ext,
B_sup, B_child): the GPU tests check E_child / E_cand against them and the
benchmark prices its roofline with the byte counts.

Usage (needs /root/reference; run from the repo root):
    make -f oracle/Makefile.ref && python tests/golden/make_golden.py
    python tests/golden/make_golden.py --only config2   # add / replace cases
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import gpu_graph_mining_amd as gm  # noqa: E402
from gpu_graph_mining_amd import workloads as wl  # noqa: E402
from oracle import py_oracle as po  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "gspan_ref"
FULL_LIMIT = 1200  # store listings up to this many lines in full


def run_ref(path: Path, ftype: str, minsup: int) -> list:
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "out.tsv"
        subprocess.run([str(REF), ftype, str(path), str(minsup), str(out)], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return out.read_text().splitlines()


def case_record(name: str, lines: list, **meta) -> dict:
    srt = sorted(lines)
    blob = ("\n".join(srt) + "\n").encode() if srt else b""
    rec = dict(name=name, n_patterns=len(srt), sha256=hashlib.sha256(blob).hexdigest(), **meta)
    if len(srt) <= FULL_LIMIT:
        rec["lines"] = srt
    return rec


ORACLE_KEYS = ("e_child", "e_cand", "e_root", "b_ext", "b_sup", "b_child")


def oracle_counts(db, minsup: int) -> dict:
    """Embedding and byte counts of the whole run from the pinned oracle."""
    _, st = po.mine_arrays(db.arrays(), minsup, collect=False)
    return {k: int(st[k]) for k in ORACLE_KEYS}


def big_cases():
    """Cases built on demand with --only (minutes of reference time each)."""
    db, ms = wl.config2()
    yield "config2", db, ms, {"gen": "config2", "seed": 3, "n_graphs": 100000}


def add_only(names: set) -> None:
    path = HERE / "golden.json"
    doc = json.loads(path.read_text())
    cases = [c for c in doc["cases"] if c["name"] not in names]
    with tempfile.TemporaryDirectory() as td:
        for name, db, ms, spec in big_cases():
            if name not in names:
                continue
            p = Path(td) / f"{name}.txt"
            db.write_txt(p)
            rec = case_record(name, run_ref(p, "txt", ms), db=spec, db_sha256=db.sha256(), minsup=ms)
            rec["oracle"] = oracle_counts(db, ms)
            cases.append(rec)
            print(name, rec["n_patterns"], rec["oracle"], flush=True)
    doc["cases"] = cases
    path.write_text(json.dumps(doc, indent=0))


def main() -> None:
    if not REF.exists():
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref")
    only = set(sys.argv[sys.argv.index("--only") + 1].split(",")) if "--only" in sys.argv else None
    if only:
        return add_only(only)
    cases = []
    chem = HERE / "Chemical_340.txt"
    for s in (170, 68, 34, 17, 10):
        cases.append(case_record(f"chem_{s}", run_ref(chem, "txt", s), db={"file": "Chemical_340.txt",
                                                                          "filetype": "txt"}, minsup=s))
    for f in ("labeled.bin", "unlabeled.bin"):
        for s in (2, 3, 5):
            cases.append(case_record(f"{f.split('.')[0]}_{s}", run_ref(HERE / f, "bin", s),
                                     db={"file": f, "filetype": "bin"}, minsup=s))
    dbs = json.loads((HERE / "test_support_dbs.json").read_text())
    with tempfile.TemporaryDirectory() as td:
        for name, codes in dbs.items():
            db = gm.GraphDB.from_codes(codes)
            p = Path(td) / f"{name}.txt"
            db.write_txt(p)
            for s in (1, 2, 3, 5, 9, 17):
                cases.append(case_record(f"{name}_{s}", run_ref(p, "txt", s),
                                         db={"codes": name}, minsup=s))
        gens = []
        db, ms = wl.config0()
        gens.append(("config0", db, ms, {"gen": "config0", "seed": 0, "n_graphs": 1000}))
        db, ms = wl.config1(n_graphs=1000)
        gens.append(("config1_1k", db, ms, {"gen": "config1", "seed": 2, "n_graphs": 1000}))
        db, ms = wl.config1()
        gens.append(("config1", db, ms, {"gen": "config1", "seed": 2, "n_graphs": 10000}))
        db = wl.config4(n_graphs=2000)
        gens.append(("config4_2k", db, wl.abs_support(0.10, 2000), {"gen": "config4", "seed": 4, "n_graphs": 2000}))
        for j in wl.config3_jobs()[:6]:
            if j.n_graphs > 6000:
                continue
            gens.append((f"config3_job{j.index}", j.database(), j.minsup,
                         {"gen": "config3", "seed": 1, "job": j.index}))
        for name, db, ms, spec in gens:
            p = Path(td) / f"{name}.txt"
            db.write_txt(p)
            cases.append(case_record(name, run_ref(p, "txt", ms), db=spec, db_sha256=db.sha256(), minsup=ms))
            print(name, cases[-1]["n_patterns"], flush=True)
    (HERE / "golden.json").write_text(json.dumps({"source": "reference src/gspan via oracle/_ref/gspan_ref",
                                                  "cases": cases}, indent=0))
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
