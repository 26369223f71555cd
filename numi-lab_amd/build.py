"""Builds libmetalrobo_amd.so (HIP kernels for gfx950 + C++ host + C ABI) in-tree.

Usage: python numi-lab_amd/build.py [--jobs N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libmetalrobo_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MR_OFFLOAD_ARCH", "gfx950")

# Device code: no contraction into FMA, correctly rounded division / square
# root, IEEE denormals, no packed-FP32 instruction selection.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]
SOURCES = [
    ("mr_kernels.hip", "hip"),
    ("mr_mlp.hip", "hip"),
    ("mr_task.hip", "hip"),
    ("mr_task_host.cpp", "cxx"),
    ("mr_packs.cpp", "cxx"),
    ("mr_policy.cpp", "cxx"),
    ("mr_builder.cpp", "cxx"),
    ("mr_world.cpp", "cxx"),
    ("mr_c_api.cpp", "cxx"),
]


def _compile(src, kind):
    out = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if kind == "hip":
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *DEVICE_FLAGS, "-c", path, "-o", out]
    else:
        cmd = ["g++", *HOST_FLAGS, "-c", path, "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and "warning: feature flag" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    return out


def build(jobs=4, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda sk: _compile(*sk), SOURCES))
    tmp = LIBRARY + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
           "-o", tmp, f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    try:
        build(args.jobs)
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
