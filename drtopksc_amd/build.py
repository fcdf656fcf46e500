"""Build libdrtopk.so (the HIP library behind include/drtopk.h) for gfx950, in-tree.

    python -m drtopksc_amd.build        # or: python drtopksc_amd/build.py

Every .hip file under csrc/ is compiled with hipcc into an object, then linked into
drtopksc_amd/lib/libdrtopk.so.  Always rebuilds (the build check must compile).
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libdrtopk.so")
CLI_SRC = os.path.join(HERE, "cli", "topk_main.cpp")
CLI = os.path.join(LIB_DIR, "drtopk_topk")
# GPU check of the C++ mirror (include/drtopk.hpp), run by tests/test_gpu_hpp.py
HPP_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hpp_mirror.cpp")
HPP_TEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "hpp_mirror")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("DRTOPK_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 VALU ops are avoided on this target (no kernel here uses fp32 math anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # command-line driver mirroring the reference's main.cu, linked against the library
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", CLI_SRC, f"-L{LIB_DIR}", "-ldrtopk",
           "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed:\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    if os.path.exists(HPP_TEST_SRC):
        cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-Wno-unused-result", f"-I{INCLUDE}", HPP_TEST_SRC, f"-L{LIB_DIR}", "-ldrtopk",
               "-Wl,-rpath,$ORIGIN/../../drtopksc_amd/lib", "-o", HPP_TEST + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hpp mirror test build failed:\n{r.stderr}")
        os.replace(HPP_TEST + ".tmp", HPP_TEST)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
