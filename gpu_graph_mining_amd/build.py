"""Build the native library of the miner in-tree with hipcc (gfx950).

Produces gpu_graph_mining_amd/_lib/libgspan_amd.so from csrc/.  Object files
go to gpu_graph_mining_amd/_build/ (git-ignored).  Run as a module:

    python -m gpu_graph_mining_amd.build
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libgspan_amd.so"
ARCH = os.environ.get("GM_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter", f"-I{PKG.parent / 'include'}"]

DEVICE_SOURCES = ["kernels.hip", "canon.hip"]
HOST_SOURCES = ["capi.cpp", "miner.cpp", "canonical.cpp", "host_db.cpp", "generators.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> str:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def check_no_scratch(remarks: str) -> None:
    """Fail when a kernel of the gfx950 code object uses scratch memory
    (from the -Rpass-analysis=kernel-resource-usage remarks)."""
    kernel = None
    bad = []
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            kernel = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and int(m.group(1)) != 0:
            bad.append(f"{kernel}: {m.group(1)} bytes/lane")
    if bad:
        raise RuntimeError("kernels use scratch memory:\n  " + "\n  ".join(bad))


def build(verbose: bool = False, out: Path = LIB, build_dir: Path = BUILD, defines: tuple = ()) -> Path:
    """Compile csrc/ into the shared library `out` (default _lib/)."""
    hipcc = _hipcc()
    build_dir.mkdir(parents=True, exist_ok=True)
    out.parent.mkdir(parents=True, exist_ok=True)
    extra = [f"-D{d}" for d in defines]
    jobs = []
    objs = []
    for src in DEVICE_SOURCES:
        obj = build_dir / (src + ".o")
        # the resource-usage remarks let the build fail on any kernel that
        # uses scratch memory (per-lane arrays the compiler could not keep in
        # registers)
        jobs.append([hipcc, f"--offload-arch={ARCH}", *COMMON, *extra, *DEVICE_FLAGS,
                     "-Rpass-analysis=kernel-resource-usage", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = build_dir / (src + ".o")
        # host-only translation units; hipcc supplies the HIP runtime headers
        jobs.append([hipcc, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd in jobs:
            if verbose:
                print(" ".join(cmd), flush=True)
        outs = list(ex.map(_run, jobs))
    for out_text in outs[:len(DEVICE_SOURCES)]:
        check_no_scratch(out_text)
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    if out == LIB:
        build_cli()
        build_calibration()
    return out


CALIB = LIB_DIR / "libgm_calib.so"


def build_calibration() -> Path:
    """The benchmark's counter-calibration kernels (csrc/calibrate.hip) in a
    library of their own: measurement support, not part of the miner."""
    tmp = CALIB.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, "-shared", str(CSRC / "calibrate.hip"), "-o",
          str(tmp)])
    os.replace(tmp, CALIB)
    return CALIB


CLI_SRC = PKG / "cli" / "gspan_amd_main.cpp"
CLI = LIB_DIR / "gspan_amd"


def build_cli() -> Path:
    """The command-line driver (cli/gspan_amd_main.cpp), plain C++ against
    include/gspan_amd.hpp, linked to the library next to it."""
    tmp = CLI.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{PKG.parent / 'include'}", str(CLI_SRC), "-o", str(tmp),
          f"-L{LIB_DIR}", "-lgspan_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle() -> None:
    """Compile the CPU oracle (test infrastructure) and, where the reference
    sources are present, the reference build used to pin it."""
    root = PKG.parent
    _run(["make", "-s", "-C", str(root), "-f", "oracle/Makefile"])
    if Path("/root/reference/src/gspan/gspan.cpp").exists():
        _run(["make", "-s", "-C", str(root), "-f", "oracle/Makefile.ref"])


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    build_oracle()
    print(p)
