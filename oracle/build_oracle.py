"""TEST INFRASTRUCTURE: builds the CPU oracle (liboracle_mr.so) from oracle/*.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
FP32 expressions are compiled without contraction (-ffp-contract=off) on
x86-64 SSE2 scalar arithmetic, so they round exactly as written.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liboracle_mr.so")


def build(verbose=True):
    tmp = LIBRARY + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wno-unused-function", os.path.join(HERE, "oracle_world.cpp"), os.path.join(HERE, "exact_cone.cpp"),
           os.path.join(HERE, "oracle_task.cpp"), "-o", tmp, "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
