"""Builds libfpgm_ci.so (HIP for gfx950) in-tree with hipcc.

    python -m fastpgm_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfpgm_ci.so")
GEN_LIB = os.path.join(HERE, "libfpgm_gen.so")
PROBE_LIB = os.path.join(HERE, "libfpgm_probe.so")
HOST_LIB = os.path.join(HERE, "libfpgm_host.so")
HOST_CLI = os.path.join(HERE, "fpgm_pcstable")
HOST_SOURCES = ["dataset.cpp", "network.cpp", "pc_stable.cpp", "bnsl_comparison.cpp", "param_learning.cpp", "host_c_api.cpp"]

SOURCES = ["ci_kernels.hip", "gram_kernels.hip", "mle_kernels.hip", "round_kernels.hip", "fpgm_ci.hip"]
HEADERS = ["ci_kernels.h", "gram_kernels.h", "mle_kernels.h", "round_kernels.h", "chisq.h", "onehot.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-fopenmp",  # host threads for building large rounds
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_gen(verbose: bool = False) -> str:
    """Host-only synthetic-data sampler (csrc/gen.c)."""
    tmp = GEN_LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O3", "-fopenmp", "-fPIC", "-shared", "-o", tmp,
           os.path.join(CSRC, "gen.c")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


def build_probe(verbose: bool = False) -> str:
    """Throughput probes of the resources the counting kernels are bound by (csrc/probe.hip),
    loaded by bench.py only."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "probe.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


def build(verbose: bool = False) -> str:
    build_gen(verbose)
    build_probe(verbose)
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_host(verbose: bool = False) -> str:
    """C++ host mirror (host/: Dataset, Network, PCStable) over libfpgm_ci.so, and its
    command-line driver.  Host code only: g++, no device code."""
    cxx = os.environ.get("CXX", "g++")
    hdir = os.path.join(HERE, "host")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", hdir]
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]
    link = ["-L", HERE, "-Wl,-rpath,$ORIGIN", "-lfpgm_ci"]
    tmp = HOST_LIB + ".tmp"
    cmd = [cxx, *flags, *inc, "-shared", "-o", tmp, *[os.path.join(hdir, f) for f in HOST_SOURCES], *link]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_LIB)
    tmp = HOST_CLI + ".tmp"
    cmd = [cxx, *flags, *inc, "-o", tmp, os.path.join(hdir, "pcstable_cli.cpp"), "-L", HERE,
           "-Wl,-rpath,$ORIGIN", "-lfpgm_host", "-lfpgm_ci"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_CLI)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host(verbose="-v" in sys.argv))
