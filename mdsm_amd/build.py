"""Build libmdsm_hip.so (gfx950) in-tree with hipcc.

    python -m mdsm_amd.build        # or mdsm_amd.build.build()

Every source under mdsm_amd/csrc is compiled with hipcc for gfx950 and linked
into mdsm_amd/libmdsm_hip.so.  Objects go to mdsm_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.environ.get("MDSM_BUILD_DIR", os.path.join(HERE, "_build"))
LIB = os.environ.get("MDSM_BUILD_LIB", os.path.join(HERE, "libmdsm_hip.so"))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
    "-ffp-contract=off",  # one IEEE operation per written operation
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
] + os.environ.get("MDSM_EXTRA_FLAGS", "").split()
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


DRIVER_SRC = os.path.join(HERE, "driver", "mdsm_file.cpp")
# the driver sits next to the library it links (found through $ORIGIN)
DRIVER = os.path.join(os.path.dirname(LIB), "mdsm_file")


def build_driver(verbose: bool = False) -> str:
    """The standalone file pipeline (driver/mdsm_file.cpp), a plain C++
    program linked against libmdsm_hip.so (found next to it at run time)."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall",
           "-I" + os.path.join(HERE, "..", "include"), DRIVER_SRC, "-o", DRIVER + ".tmp",
           "-L" + os.path.dirname(LIB), "-l:" + os.path.basename(LIB),
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
