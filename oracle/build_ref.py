"""TEST INFRASTRUCTURE: builds oracle/_ref/libmr_ref.so, the reference's own FP64
CPU oracle compiled from its sources where they lie, plus this directory's C
harness (ref_harness.cpp) over it.

Compiled reference files (under /root/reference, read in place, never copied):
  src/core/ArticulatedDynamics.cpp   FP64 CRBA + RNEA forward dynamics
  src/core/Collision.cpp             FP64 collision oracle (collideCpuReference)
  src/core/ReferenceConicSolver.cpp  exact-cone contact solve
  src/core/EngineModel.cpp           EngineModel::valid, used by the dynamics
  src/core/ConstraintIR.cpp          validateConstraintIR, used by EngineModel::valid
  src/core/ConstraintSolver.cpp      projected Gauss-Seidel over free bodies
                                     (solveContactConstraints)
They need only the C++ standard library and the reference's own headers
(include/metalrobo), so g++ builds them directly; nothing else of the
reference's build is run. Only tests/ load the result.

It also copies one data file the reference ships, a trained G1 locomotion
policy pack (artifacts/g1-legs-locomotion-60min-20260805/retained-policy/
candidate.policypack), to oracle/_ref/g1_legs_locomotion.policypack, the
golden input of the policy-pack tests.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"
OUT_DIR = os.path.join(HERE, "_ref")
LIBRARY = os.path.join(OUT_DIR, "libmr_ref.so")
SOURCES = ["ArticulatedDynamics.cpp", "Collision.cpp", "ReferenceConicSolver.cpp", "EngineModel.cpp",
           "ConstraintIR.cpp", "ConstraintSolver.cpp"]


POLICY_PACK = os.path.join(REFERENCE, "artifacts", "g1-legs-locomotion-60min-20260805", "retained-policy",
                           "candidate.policypack")
POLICY_PACK_COPY = os.path.join(OUT_DIR, "g1_legs_locomotion.policypack")


def copy_policy_pack(verbose=True):
    """Copies the reference's trained policy pack into oracle/_ref when it is present."""
    if not os.path.isfile(POLICY_PACK):
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    shutil.copyfile(POLICY_PACK, POLICY_PACK_COPY)
    if verbose:
        print(f"copied {POLICY_PACK_COPY}")
    return POLICY_PACK_COPY


def available() -> bool:
    return all(os.path.isfile(os.path.join(REFERENCE, "src", "core", s)) for s in SOURCES)


def build(verbose=True):
    """Builds the library when the reference sources are present; returns its path, or None."""
    if not available():
        if verbose:
            print("reference sources absent: oracle/_ref not built")
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [os.path.join(REFERENCE, "src", "core", s) for s in SOURCES] + [os.path.join(HERE, "ref_harness.cpp")]
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(OUT_DIR, os.path.splitext(os.path.basename(s))[0] + ".o")
        objs.append(o)
        cmd = ["g++", "-std=c++23", "-O2", "-fPIC", "-c", "-I" + os.path.join(REFERENCE, "include"), s, "-o", o]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for cmd, p in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"reference oracle build failed: {' '.join(cmd)}\n{err}")
    tmp = LIBRARY + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"reference oracle link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {LIBRARY}")
    copy_policy_pack(verbose)
    return LIBRARY


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
