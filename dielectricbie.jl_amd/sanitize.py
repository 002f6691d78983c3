"""Sanitizer builds of libdbie's host code, run against the CPU test suite.

    python dielectricbie.jl_amd/sanitize.py address    # AddressSanitizer + UndefinedBehaviorSanitizer
    python dielectricbie.jl_amd/sanitize.py thread     # ThreadSanitizer

The host C++ sources (mesher, near-pair lists and correction blocks on host threads, target-map
refinement and cubatures on host threads, request server, C ABI) are compiled with g++ and the
sanitizer; the device objects of the normal build are linked unchanged (no GPU code is instrumented,
and the CPU tests run no kernel).  The library goes to dbie/libdbie_<kind>.so and the CPU tests run
with DBIE_LIB pointing at it and the sanitizer runtime preloaded.  Exit status: that of pytest.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import build as B  # noqa: E402

FLAGS = {
    "address": ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"],
    "thread": ["-fsanitize=thread", "-fno-omit-frame-pointer"],
}
RUNTIME = {"address": "libasan.so", "thread": "libtsan.so"}
# tests that exercise the host code paths on the CPU
TESTS = ["tests/test_host.py", "tests/test_targets.py", "tests/test_multibox.py", "tests/test_boundary.py",
         "tests/test_lattice.py"]


def main(kind: str) -> int:
    if kind not in FLAGS:
        print(f"usage: sanitize.py {{{'|'.join(FLAGS)}}}", file=sys.stderr)
        return 2
    B.build(force=False, jobs=8)  # the device objects of the normal build
    out_dir = os.path.join(HERE, f"_build_{kind}")
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src in B._sources():
        if src.endswith(".cpp"):
            obj = os.path.join(out_dir, src + ".o")
            cmd = ["g++", *B.HOST_FLAGS, "-O1", "-g", *FLAGS[kind], "-c", os.path.join(B.SRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                print(r.stderr, file=sys.stderr)
                return 1
            objs.append(obj)
        else:
            objs.append(os.path.join(B.OUT_DIR, src + ".o"))
    lib = os.path.join(HERE, "dbie", f"libdbie_{kind}.so")
    link = ["g++", "-shared", "-o", lib, *objs, *FLAGS[kind], f"-L{B.ROCM}/lib", "-lrocsolver", "-lrocblas",
            "-lamdhip64", "-pthread", f"-Wl,-rpath,{B.ROCM}/lib"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stderr, file=sys.stderr)
        return 1
    def libpath(name):
        return subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    # libstdc++ is preloaded right after the runtime: the python executable does not link it, and the
    # runtime's __cxa_throw interceptor resolves the real one only if it is loaded at start-up
    preload = f"{libpath(RUNTIME[kind])}:{libpath('libstdc++.so')}"
    env = dict(os.environ, DBIE_LIB=lib, LD_PRELOAD=preload,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1:report_signal_unsafe=0")
    if kind == "thread":
        # the oracle the tests check against runs OpenMP loops, and libgomp's barriers are not
        # instrumented: TSan would flag the oracle's own thread hand-offs.  One OpenMP thread keeps the
        # report to libdbie's threads (std::thread pools, the request server).
        env["OMP_NUM_THREADS"] = "1"
    root = os.path.dirname(HERE)
    # -s: a sanitizer report goes to fd 2 just before the abort, which pytest's capture would swallow
    cmd = [sys.executable, "-m", "pytest", "-q", "-s", "-x", "-m", "not gpu", "-p", "no:cacheprovider", *TESTS]
    print(f"[sanitize] {kind}: {lib}; running {' '.join(TESTS)}", flush=True)
    return subprocess.run(cmd, cwd=root, env=env).returncode


if __name__ == "__main__":
    sys.exit(main(sys.argv[1] if len(sys.argv) > 1 else "address"))
