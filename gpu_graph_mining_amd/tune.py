"""Autotuner for the level kernels' launch shapes and the placement of the
canonical check.

For a workload (one database at a min support, or a batch of jobs) it times
every candidate setting of

  chunk_rows   rows per wave of k_level / k_write (64 .. 512)
  table_bits   log2 entries of a wave's LDS count table (7 .. 10)
  canon_host   0 adaptive, 1 host pool, 2 device (k_canon_level)

checks that each reports exactly the patterns of the default setting, and
keeps the fastest (median of `runs` timed runs).  The result is written as
a configuration file in the layout cudapp::cuda_configurator reads
(src/cuda_pp/cuda_configurator.cpp:26-76): a header row `id,<parameter>,...`
then one row per configuration id, values as integers.  The same file is
read back by read_config() here and by gspan_amd::cuda_configurator(file,
id).parse_config_file() in include/gspan_amd.hpp.

    python -m gpu_graph_mining_amd.tune --workload config3 --out tuned.csv
"""
from __future__ import annotations

import argparse
import csv
import os
import statistics
import sys
import time
from pathlib import Path
from typing import Callable, Dict, List, Sequence, Tuple

from . import GraphDB, Batch, mine

PARAMS = ("chunk_rows", "table_bits", "canon_host")
SHAPES: Tuple[Tuple[int, int], ...] = ((64, 7), (64, 8), (128, 8), (128, 9), (256, 8), (256, 9), (512, 9), (512, 10))
CANON = (0, 1, 2)


def _median_time(fn: Callable[[], object], runs: int) -> float:
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def _sweep(run: Callable[..., object], same: Callable[[object, object], bool], runs: int,
           shapes: Sequence[Tuple[int, int]], canon: Sequence[int]) -> Dict:
    ref = run(count_only=False)
    table = []
    for cr, tb in shapes:
        for ch in canon:
            kw = {"chunk_rows": cr, "table_bits": tb, "canon_host": ch}
            if not same(run(count_only=False, **kw), ref):
                raise RuntimeError(f"setting {kw} changed the results")
            run(count_only=True, **kw)  # warm
            table.append({**kw, "seconds": _median_time(lambda: run(count_only=True, **kw), runs)})
    best = min(table, key=lambda r: r["seconds"])
    default = next((r for r in table if (r["chunk_rows"], r["table_bits"], r["canon_host"]) == (256, 8, 0)), None)
    return {"best": {k: best[k] for k in PARAMS}, "best_seconds": best["seconds"],
            "default_seconds": default["seconds"] if default else None, "table": table}


def tune_mine(db: GraphDB, minsup: int, runs: int = 3, shapes=SHAPES, canon=CANON) -> Dict:
    """Tune plain mining of one database."""
    db.upload()

    def run(count_only: bool, **kw):
        pats, _ = mine(db, minsup, count_only=count_only, **kw)
        return sorted((tuple(c), s) for c, s, _ in pats)

    return _sweep(run, lambda a, b: a == b, runs, shapes, canon)


def tune_batch(batch: Batch, minsups: Sequence[int], runs: int = 3, shapes=SHAPES, canon=CANON) -> Dict:
    """Tune one batched schedule of many jobs."""

    def run(count_only: bool, **kw):
        res, _ = batch.mine(minsups, count_only=count_only, **kw)
        return [sorted((tuple(c), s) for c, s, _ in r) for r in res]

    return _sweep(run, lambda a, b: a == b, runs, shapes, canon)


def write_config(path: os.PathLike, config_id: str, params: Dict[str, int]) -> None:
    """Add or replace one configuration row (header: id + PARAMS)."""
    path = Path(path)
    rows: Dict[str, Dict[str, int]] = {}
    if path.exists():
        for cid, p in _read_rows(path).items():
            rows[cid] = p
    rows[config_id] = {k: int(params[k]) for k in PARAMS}
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["id", *PARAMS])
        for cid, p in rows.items():
            w.writerow([cid, *(p[k] for k in PARAMS)])


def _read_rows(path: Path) -> Dict[str, Dict[str, int]]:
    with open(path, newline="") as f:
        r = list(csv.reader(f))
    if not r:
        raise ValueError(f"{path}: empty configuration file")
    header = [h.strip() for h in r[0][1:]]
    return {row[0].strip(): {h: int(v) for h, v in zip(header, row[1:])} for row in r[1:] if row}


def read_config(path: os.PathLike, config_id: str) -> Dict[str, int]:
    """The parameters of one configuration id (mine() keyword arguments)."""
    rows = _read_rows(Path(path))
    if config_id not in rows:
        raise KeyError(f"{path}: no configuration {config_id!r}")
    return rows[config_id]


def main(argv: List[str] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=("config1", "config2", "config3"), required=True)
    ap.add_argument("--out", required=True, help="configuration file to add the row to")
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--jobs", type=int, default=256, help="config3: the first N jobs of the mix")
    args = ap.parse_args(argv)
    from . import workloads as wl
    if args.workload == "config3":
        jobs = wl.config3_jobs()[:args.jobs]
        b = Batch([j.database() for j in jobs])
        res = tune_batch(b, [j.minsup for j in jobs], runs=args.runs)
    else:
        db, ms = wl.config1() if args.workload == "config1" else wl.config2()
        res = tune_mine(db, ms, runs=args.runs)
    write_config(args.out, args.workload, res["best"])
    for row in sorted(res["table"], key=lambda r: r["seconds"]):
        print(f"# {row['chunk_rows']:4d} {row['table_bits']:2d} {row['canon_host']} {row['seconds'] * 1e3:9.3f} ms",
              flush=True)
    print(f"{args.workload}: best {res['best']} {res['best_seconds'] * 1e3:.3f} ms, default "
          f"{(res['default_seconds'] or 0) * 1e3:.3f} ms", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
