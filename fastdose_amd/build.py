"""Build libfastdose_hip.so (gfx950) in-tree, and the CPU oracle used by the tests.

    python -m fastdose_amd.build            # library + oracle
    python -m fastdose_amd.build --lib-only

The HIP sources are compiled with hipcc for gfx950 only; host-only C++ with g++.  No
contraction anywhere (-ffp-contract=off): the FP32 decision arithmetic of the kernels must match
the host code bit for bit; kernels spell out their FMAs.  Packed FP32 instructions are disabled
(`-packed-fp32-ops`).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libfastdose_hip.so"
CLI = PKG / "fastdose_imrt"
ORACLE_DIR = ROOT / "oracle"

# Per-file code generation options.  kernels_cccs.hip: coalescing the copies on split edges keeps
# more of the step loop's X states in one register set across the rotation branches, and the ILP
# scheduling strategy interleaves the independent per-slot chains better (CCCS kernel 1.52 -> 1.48
# -> 1.46 ms on the 128^3 slab beam).
EXTRA_FLAGS = {"kernels_cccs.hip": ["-mllvm", "-join-splitedges=true", "-mllvm", "-amdgpu-sched-strategy=max-ilp",
                                    "-mllvm", "-structurizecfg-skip-uniform-regions=true"]}

HIP_SOURCES = ["api.hip", "engine.hip", "serve.hip", "noalloc.hip", "kernels_terma.hip", "kernels_cccs.hip", "kernels_resample.hip", "kernels_csr.hip", "kernels_flags.hip", "kernels_spmm.hip", "spmm.hip"]
CXX_SOURCES = ["geometry.cpp", "dosemat_files.cpp"]
ARCH = os.environ.get("FASTDOSE_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def build_library(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([hipcc, "-c", *common, f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                     "-packed-fp32-ops", "-Wno-unused-command-line-argument", *EXTRA_FLAGS.get(src, []),
                     str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append(["g++", "-c", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
                     f"-I{ROOT / 'include'}", f"-I{CSRC}", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    # command-line drop-in for the dose mode, linked against the library
    cli_tmp = CLI.with_name(CLI.name + ".tmp")
    _run([hipcc, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(CSRC / "fastdose_imrt.cpp"), f"-L{PKG}",
          "-lfastdose_hip", "-Wl,-rpath,$ORIGIN", "-o", str(cli_tmp)])
    os.replace(cli_tmp, CLI)
    if verbose:
        print(f"built {CLI}")
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"])
    if verbose:
        print(f"built {ORACLE_DIR / 'liboracle.so'}")
    return ORACLE_DIR / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    args = ap.parse_args()
    build_library(verbose=True)
    if not args.lib_only:
        build_oracle(verbose=True)


if __name__ == "__main__":
    main()
