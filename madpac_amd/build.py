"""Build liblbm_hip.so (all HIP sources of madpac_amd/csrc) for gfx950 with hipcc.

Every translation unit is compiled as HIP with:
  * --offload-arch=gfx950 (MI355X, CDNA4);
  * -ffp-contract=off: no fused multiply-add, so the FP64 collision rounds every operation like
    the reference's x86-64 build and stays bit-identical to it;
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 instructions.
The shared library is written next to this file (in-tree) so that it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm_hip.so")
OBJDIR = os.path.join(ROOT, "_build", "lbm_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, file=sys.stderr)
    return LIB


CPP_TESTS = (("test_peano_api", True), ("test_peano_events", True), ("test_peano_amr", True), ("test_peano_blockops", True),
             ("test_geometry", False))


def build_cpp_test(verbose: bool = False) -> str:
    """Compile the C++ programs of tests/cpp: test_peano_api, test_peano_events, test_peano_amr and
    test_peano_blockops (the reference-shaped API, linked against the library) and test_geometry (the block store's geometry
    flags, host only)."""
    exe = None
    for name, links in CPP_TESTS:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        exe = os.path.join(ROOT, "tests", "cpp", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, src]
        if links:
            cmd += ["-L" + HERE, "-llbm_hip", "-Wl,-rpath,$ORIGIN/../../madpac_amd"]
        r = subprocess.run(cmd + ["-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test {name} build failed:\n{r.stderr}")
        if verbose:
            print("built", exe, file=sys.stderr)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_test(verbose=True)
