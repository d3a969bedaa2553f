"""Benchmark of the GPU gSpan miner on one MI355X.

A step is one complete mining run (every level step, every pattern) of the
largest single-GPU workload -- 100,000 planted-pattern transactions (seed 3,
Poisson(30) edges, 50 seed patterns, 5 vertex / 5 edge labels), absolute
min support 1,000 (1 %) -- with the database already resident in HBM.
The line reports extended embeddings/s (E_child: embedding rows of every
reported pattern with >= 2 edges, per step), frequent patterns/s and the
per-step mining seconds, plus the roofline of the dominant stage (the
extension + support kernels, timed with HIP events on the mining stream)
and the CPU baseline (the oracle restatement of the reference's sequential
gSpan, one core, on a bounded sample of the same workload).  Parity: the
first run's pattern listing must equal the reference's own (sha256 of its
sorted "code<TAB>support" lines in tests/golden/golden.json), and E_child /
E_cand must equal the oracle's.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1|2|4] [--no-cpu] [--jobs N]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _workload(cfg: int):
    from gpu_graph_mining_amd import workloads as wl
    if cfg == 1:
        db, ms = wl.config1()
        return db, ms, {"workload": "config1: planted-pattern transactions", "n_graphs": 10000, "seed": 2,
                        "min_support": ms, "min_support_frac": 0.05}
    if cfg == 2:
        db, ms = wl.config2()
        return db, ms, {"workload": "config2: planted-pattern transactions", "n_graphs": 100000, "seed": 3,
                        "min_support": ms, "min_support_frac": 0.01}
    if cfg == 4:
        db = wl.config4()
        ms = wl.abs_support(0.05, 50000)
        return db, ms, {"workload": "config4: molecule-like graphs (mining request)", "n_graphs": 50000, "seed": 4,
                        "min_support": ms, "min_support_frac": 0.05}
    raise SystemExit(f"unknown config {cfg}")


def _traffic(kernel: str, workload: str):
    """HBM bytes per launch of the kernel measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), from
    the committed summary, or None."""
    p = ROOT / "profiles" / "pmc_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get(workload, {}).get(kernel)
    except Exception:
        return None


def _golden_case(cfg: int):
    """The reference's listing of the workload (tests/golden/golden.json):
    sha256 and count of its sorted "code<TAB>support" lines, plus the
    oracle's embedding and byte counts for the full-size workload."""
    name = {1: "config1", 2: "config2"}.get(cfg)
    if name is None:
        return None
    cases = json.loads((ROOT / "tests" / "golden" / "golden.json").read_text())["cases"]
    return next((c for c in cases if c["name"] == name), None)


def _code_str(code) -> str:
    """A code in the sequential gSpan's -1 convention, as to_string(false)
    prints it (src/globals/dfs_code.cpp:199-230)."""
    return ";".join("(%d %d %d %d %d)" % (d[0], d[1], d[2] if i == 0 else -1, d[3], d[4] if d[0] < d[1] else -1)
                    for i, d in enumerate(code))


def _sha(lines) -> str:
    import hashlib
    return hashlib.sha256((("\n".join(lines) + "\n") if lines else "").encode()).hexdigest()


def _cpu_model() -> str:
    try:
        for l in open("/proc/cpuinfo"):
            if l.startswith("model name"):
                return l.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical cpus visible)"
    except OSError:
        pass
    return platform.processor()


def _tuned(config_id: str) -> dict:
    """The autotuner's launch shapes for a workload (gpu_graph_mining_amd/
    tuned.csv, written by `python -m gpu_graph_mining_amd.tune`), or {}."""
    from gpu_graph_mining_amd import tune
    p = ROOT / "gpu_graph_mining_amd" / "tuned.csv"
    try:
        return tune.read_config(p, config_id) if p.exists() else {}
    except KeyError:
        return {}


def _batch_section(n_jobs: int, out: dict) -> None:
    """Config 3: n_jobs independent jobs, batched in one schedule (with the
    autotuned launch shapes) vs one job at a time with the defaults on the
    same GPU (same library, same results)."""
    import gpu_graph_mining_amd as gm
    from gpu_graph_mining_amd import workloads as wl
    jobs = wl.config3_jobs()[:n_jobs]
    dbs = [j.database() for j in jobs]
    mins = [j.minsup for j in jobs]
    for d in dbs:
        d.upload()
    batch = gm.Batch(dbs)  # packed and uploaded once, like the per-job databases
    tuned = _tuned("config3")
    # parity: the batched schedule reports exactly each job's own patterns
    res, _ = batch.mine(mins, **tuned)
    alone = [gm.mine(d, m)[0] for d, m in zip(dbs, mins)]
    same = all(sorted((tuple(c), s) for c, s, _ in r) == sorted((tuple(c), s) for c, s, _ in a)
               for r, a in zip(res, alone))
    t0 = time.perf_counter()
    seq = []
    for d, m in zip(dbs, mins):
        _, st = gm.mine(d, m, count_only=True)
        seq.append(st["n_patterns"])
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, st = batch.mine(mins, count_only=True, **tuned)
    t_bat = time.perf_counter() - t0
    out["batch"] = {"jobs": n_jobs, "jobs_per_s_batched": n_jobs / t_bat, "jobs_per_s_sequential": n_jobs / t_seq,
                    "speedup": t_seq / t_bat, "patterns": int(sum(seq)), "results_identical": same,
                    "batched_patterns": int(st["n_patterns"]), "tuned": tuned}


def _batched_only(n_jobs: int, runs: int, out: dict) -> None:
    """The 256-job mix as one batched schedule only (with the autotuned
    launch shapes), `runs` times: the workload of the batch profile."""
    import gpu_graph_mining_amd as gm
    from gpu_graph_mining_amd import workloads as wl
    jobs = wl.config3_jobs()[:n_jobs]
    batch = gm.Batch([j.database() for j in jobs])
    mins = [j.minsup for j in jobs]
    tuned = _tuned("config3")
    batch.mine(mins, count_only=True, **tuned)  # warm
    t0 = time.perf_counter()
    for _ in range(runs):
        _, st = batch.mine(mins, count_only=True, **tuned)
    dt = (time.perf_counter() - t0) / max(1, runs)
    out["batched_only"] = {"jobs": n_jobs, "runs": runs, "seconds_per_run": dt, "jobs_per_s": n_jobs / dt,
                           "patterns": int(st["n_patterns"]), "tuned": tuned}


def _serving_section(n_requests: int, rate: float, out: dict) -> None:
    """Config 4: Poisson arrivals of 90 % pattern-support queries (3-12
    edges, random-walk subgraphs) and 10 % mining jobs (1k-10k graph
    subsets at 5-20 %) against the resident molecule database, served with
    continuous batching (concurrent slots, memory admission); latency =
    completion - arrival.  At the stated rate and at 2x and 4x it."""
    from gpu_graph_mining_amd import serving
    from gpu_graph_mining_amd import workloads as wl
    db = wl.config4()
    db.upload()
    serving.serve(db, rate=rate, n_requests=min(20, n_requests), seed=40)  # warm the handles and the library
    runs = []
    for mult in (1, 2, 4):
        res = serving.serve(db, rate=rate * mult, n_requests=n_requests, seed=4)
        res.pop("_requests")
        runs.append(res)
    out["serving"] = {"workload": "config4: 50,000 molecule-like graphs, 90% support queries / 10% mining",
                      **runs[0]}
    out["serving_rates"] = [{"rate_per_s": r["rate_per_s"], "throughput_per_s": r["throughput_per_s"],
                             "query_p50_ms": r["query"]["p50_ms"], "query_p99_ms": r["query"]["p99_ms"],
                             "mining_p50_ms": r["mining"]["p50_ms"], "mining_p99_ms": r["mining"]["p99_ms"],
                             "all_p99_ms": r["all"]["p99_ms"]} for r in runs]


def _calibration(kinds: str, mib: int) -> None:
    """Kernels of known traffic (gpu_graph_mining_amd/csrc/calibrate.hip),
    each touching every byte of an mib-MiB buffer once: stream16 (16-byte
    loads and stores, reads and writes mib MiB), gather4 (scattered 4-byte
    loads, reads mib MiB), atomic8 (one 64-bit atomic add per word, in a
    scattered order).  Their FETCH_SIZE / WRITE_SIZE readings scale the
    counters to bytes for the access widths the level kernels use."""
    import ctypes
    from gpu_graph_mining_amd import build as gbuild
    lib = ctypes.CDLL(str(gbuild.CALIB))
    lib.gmc_run.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    for kind in kinds.split(","):
        ms = ctypes.c_double(0)
        rc = lib.gmc_run(kind.encode(), ctypes.c_uint64(mib << 20), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"calibration kernel {kind} failed ({rc})")
        print(f"# calibration {kind}: {mib} MiB in {ms.value:.3f} ms ({(mib << 20) / ms.value / 1e6:.1f} GB/s)",
              file=sys.stderr, flush=True)


def _main_section(args) -> dict:
    """The headline workload: timed mining runs, parity against the
    reference listing and the oracle, the CPU baseline and the roofline."""
    import gpu_graph_mining_amd as gm
    db, minsup, config = _workload(args.config)
    t0 = time.perf_counter()
    db.upload()
    upload_s = time.perf_counter() - t0
    counts = db.counts()
    config.update({"n_vertices": counts["n_vertices"], "n_edges": counts["n_edges"]})
    golden = _golden_case(args.config)

    # the autotuner's launch shapes for this workload, if it has a row
    tuned = _tuned(f"config{args.config}")
    config["tuned"] = tuned
    # reference-checked first run (outside the timed region)
    pats, st0 = gm.mine(db, minsup, **tuned)
    print(f"# first run: {st0['n_patterns']} patterns, E_child={st0['e_child']}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        gm.mine(db, minsup, count_only=True, stage_events=True, **tuned)
    ext_s = ev_s = 0.0
    launches = 0
    canon_s = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st = gm.mine(db, minsup, count_only=True, stage_events=True, **tuned)
        ext_s += st["seconds_ext_kernel"]
        ev_s += st["seconds_stage_events"]
        launches += st["ext_kernel_launches"]
        canon_s += st["seconds_canon"]
    elapsed = time.perf_counter() - t0
    step_s = elapsed / args.steps
    e_child = st["e_child"]
    value = e_child / step_s

    line = {
        "metric": "extended embeddings/s",
        "value": value,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": config,
        "patterns": st["n_patterns"],
        "patterns_per_s": st["n_patterns"] / step_s,
        "e_child": e_child,
        "e_cand": st["e_cand"],
        "mine_seconds": step_s,
        "mine_seconds_with_upload": step_s + upload_s,
        "levels": st["n_levels"],
        "canon_seconds_per_step": canon_s / args.steps,
        "stage_seconds_per_step": ev_s / args.steps,
        "stage_seconds_per_step_stamps": ext_s / args.steps,
        "split_passes": st["split_passes"],
        "max_graph_rows": st["max_graph_rows"],
        "max_table_fill": st["max_table_fill"],
    }

    roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
            "traffic": _traffic("stage", config["workload"]),
            "kernel": "k_level (+ k_level_fix): extension enumeration and support counting of one level step",
            "traffic_unit": "HBM bytes per launch (one level step's k_level + k_level_fix)",
            "launches_per_step": launches / args.steps,
            "avg_launch_ms": (ev_s / launches * 1e3) if launches else None}
    # parity: the reference's listing (sha256 + count), the oracle's counts
    ok = True
    if golden is not None:
        lines = sorted(f"{_code_str(c)}\t{s}" for c, s, _ in pats)
        ok = len(lines) == golden["n_patterns"] and _sha(lines) == golden["sha256"]
    oracle = golden.get("oracle") if golden else None
    cpu = None
    if not args.no_cpu:
        from oracle import py_oracle as po
        arrays = db.arrays()
        stride = args.cpu_root_stride or 1
        if stride == 1:
            opats, ost = po.mine_arrays(arrays, minsup)
            sample = "the whole workload"
            if oracle is None:
                got = sorted((po.normalize(c), s) for c, s, _ in pats)
                want = sorted((c, s) for c, s, _ in opats)
                ok = ok and got == want
                oracle = ost
        else:
            ost = po.mine_sampled(arrays, minsup, stride)
            sample = (f"every {stride}th 1-edge root of the oracle's pattern tree, each subtree in full, on the "
                      f"whole database: {ost['n_patterns']} patterns, E_child {ost['e_child']}")
        cpu = {"value": ost["e_child"] / ost["seconds"], "unit": "embeddings/s", "cores": 1, "kind": "port",
               "sample": sample, "seconds": ost["seconds"], "patterns": ost["n_patterns"], "cpu": _cpu_model()}
    if oracle is not None:
        ok = ok and oracle["e_child"] == e_child and oracle["e_cand"] == st["e_cand"]
        # algorithmic bytes of the extension + support stage over the whole
        # run (formula in oracle/cpu_gspan.cpp), per launch = per level step
        b = oracle["b_ext"] + oracle["b_sup"]
        per_launch = b / (launches / args.steps)
        roof["algorithmic_bytes_per_step"] = b
        roof["algorithmic_bytes_per_launch"] = per_launch
        roof["achieved"] = b / (ev_s / args.steps) / 1e9
        roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    line["parity"] = ("identical" if ok else "MISMATCH") if (golden is not None or oracle is not None) else None
    if cpu is not None:
        line["cpu_baseline"] = cpu
    line["roofline"] = roof
    return line


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-root-stride", type=int, default=1,
                    help="CPU baseline sample: grow every Nth 1-edge root of the oracle's pattern tree, each subtree "
                         "in full (1: the whole workload, about 40 s on one EPYC core for config 2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--jobs", type=int, default=256,
                    help="batching: the first N jobs of the 256-job mix, batched vs one at a time (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=400,
                    help="serving: requests against the resident 50,000-graph molecule database (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="serving: Poisson arrival rate, requests/s")
    ap.add_argument("--calibrate", default="", metavar="KINDS",
                    help="after the run, the counter-calibration kernels (comma-separated: stream16, gather4, "
                         "atomic8) over --calibrate-mib MiB each, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--calibrate-mib", type=int, default=256)
    ap.add_argument("--sections", default="main,batch,serving",
                    help="comma-separated parts to run: main (the headline workload), batch, serving, "
                         "batchonly (the batched schedule alone, --steps times)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on a single GPU; "
              f"multi-GPU sharding by graph id is a design note only)", file=sys.stderr)
        return 2

    import gpu_graph_mining_amd as gm
    if gm.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 3

    sections = set(args.sections.split(","))
    if "main" in sections:
        line = _main_section(args)
    else:
        line = {"metric": "extended embeddings/s", "value": None, "unit": "embeddings/s", "n_gpus": 1,
                "sections": sorted(sections)}
    if args.jobs and "batch" in sections:
        _batch_section(args.jobs, line)
    if args.jobs and "batchonly" in sections:
        _batched_only(args.jobs, args.steps, line)
    if args.serve_requests and "serving" in sections:
        _serving_section(args.serve_requests, args.serve_rate, line)
    if args.calibrate:
        _calibration(args.calibrate, args.calibrate_mib)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
