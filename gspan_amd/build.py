"""Builds the in-tree native artefacts.

* ``gspan_amd/lib/libgspan_amd.so``  HIP kernels + host driver + C-ABI, gfx950 only
* ``gspan_amd/bin/fsm``, ``gspan_amd/bin/convert``  reference-compatible CLIs

Run as ``python -m gspan_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

# hipcc flags: packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def _run(cmd: list[str], quiet: bool) -> None:
    if not quiet:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and not quiet:
        print(out)


def build_library(quiet: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libgspan_amd.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, "gs_kernels.hip"), os.path.join(CSRC, "gs_mlp.hip"),
            os.path.join(CSRC, "gs_ingest.hip"), os.path.join(CSRC, "gs_noalloc.hip"),
            os.path.join(CSRC, "gs_twoedge.hip"),
            os.path.join(CSRC, "gs_miner.cpp")]
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", tmp], quiet)
    os.replace(tmp, out)
    return out


def build_cli(quiet: bool = False) -> list[str]:
    os.makedirs(BIN_DIR, exist_ok=True)
    host = [os.path.join(HOST, "graph.cc")]
    common = ["-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{HOST}"]
    outs = []
    fsm = os.path.join(BIN_DIR, "fsm")
    _run([CXX, *common, *host, os.path.join(HOST, "gspan.cc"), os.path.join(HOST, "fsm_main.cc"),
          f"-L{LIB_DIR}", "-lgspan_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", fsm], quiet)
    outs.append(fsm)
    conv = os.path.join(BIN_DIR, "convert")
    _run([CXX, *common, *host, os.path.join(HOST, "convert_main.cc"), "-o", conv], quiet)
    outs.append(conv)
    return outs


def build(quiet: bool = False) -> None:
    build_library(quiet)
    build_cli(quiet)


if __name__ == "__main__":
    build(quiet="-q" in sys.argv)
