"""Build libdbie.so (gfx950) in-tree: host C++ with g++, device code with hipcc.

Usage: python dielectricbie.jl_amd/build.py [-j JOBS] [--force]
Objects go to dielectricbie.jl_amd/_build/, the library to dielectricbie.jl_amd/dbie/libdbie.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "dbie", "libdbie.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
              "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INCLUDE}", "-pthread"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-parameter", f"-I{INCLUDE}", "-munsafe-fp-atomics"]


def _sources():
    files = sorted(os.listdir(SRC))
    return [f for f in files if f.endswith((".cpp", ".hip"))]


def _compile(src: str, force: bool) -> str:
    path = os.path.join(SRC, src)
    obj = os.path.join(OUT_DIR, src + ".o")
    deps = [path] + [os.path.join(SRC, h) for h in os.listdir(SRC) if h.endswith(".hpp")] + [os.path.join(INCLUDE, "dbie.h")]
    if not force and os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
        return obj
    if src.endswith(".cpp"):
        cmd = ["g++", *HOST_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = [f"{ROCM}/bin/hipcc", *HIP_FLAGS, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = True, jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [f"{ROCM}/bin/hipcc", "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs, f"-L{ROCM}/lib",
           "-lrocsolver", "-lrocblas", "-lamdhip64", "-pthread", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    try:
        build(force=a.force, jobs=a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
