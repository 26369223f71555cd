"""The C-ABI library loads, exports every entry point declared in
include/*.h, agrees with the C compiler on record layouts, and reports
errors through the thread-local mr_last_error (c_api.cpp:183, 321-332
conventions) — all without launching GPU work."""

import ctypes as ct
import os
import re
import subprocess
import tempfile
import threading

import numpy as np
import pytest

import metalrobo as mr
from metalrobo import native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = []
    for header in HEADERS:
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(mr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point():
    lib = mr.load().lib
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(nat.exported_symbols()), set(names) ^ set(nat.exported_symbols())


def test_version_and_fingerprint():
    assert "gfx950" in mr.version()
    assert mr.load().lib.mr_runtime_abi_fingerprint() != 0


def test_record_layouts_match_the_c_compiler():
    src = r'''
    #include <stdio.h>
    #include <stddef.h>
    #include "mr_c_api.h"
    int main(void) {
        printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n",
            sizeof(MRBodyC), sizeof(MRDofC), sizeof(MRShapeC), sizeof(MRMaterialC), sizeof(MRCompiledPairC),
            sizeof(MRModelHeaderC), sizeof(MRBodyStateC), sizeof(MRManifoldHeaderC), sizeof(MRManifoldPointC),
            sizeof(MRWorldStatusC), sizeof(MRContactStatusC), sizeof(MRPairRecordC), sizeof(MRWorldStepConfigC),
            sizeof(MRWorldModelDescC), offsetof(MRWorldModelDescC, heightfield), offsetof(MRBodyC, damping_scale));
        return 0;
    }'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sizes.c")
        exe = os.path.join(d, "sizes")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    expect = [nat.BODY_DTYPE.itemsize, nat.DOF_DTYPE.itemsize, nat.SHAPE_DTYPE.itemsize,
              nat.MATERIAL_DTYPE.itemsize, nat.PAIR_DTYPE.itemsize, nat.MODEL_HEADER_DTYPE.itemsize,
              nat.BODY_STATE_DTYPE.itemsize, nat.MANIFOLD_HEADER_DTYPE.itemsize, nat.MANIFOLD_POINT_DTYPE.itemsize,
              nat.WORLD_STATUS_DTYPE.itemsize, nat.CONTACT_STATUS_DTYPE.itemsize, nat.PAIR_RECORD_DTYPE.itemsize,
              ct.sizeof(nat.WorldStepConfigC), ct.sizeof(nat.WorldModelDescC),
              nat.WorldModelDescC.heightfield.offset, nat.BODY_DTYPE.fields["damping_scale"][1]]
    assert got == expect


def test_errors_are_reported_through_last_error():
    b = mr.load()
    cfg = nat.SyntheticWorldConfigC(99, 4, 0)
    assert not b.lib.mr_world_model_create_synthetic(ct.byref(cfg))
    assert "unknown synthetic world kind" in b.last_error()
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 2, 1)
    assert b.lib.mr_world_model_commands(model.handle, 0, 1, None, 0) == -1
    assert "null argument" in b.last_error()
    bad = model.default_config()
    bad.physics_substeps = 0
    with pytest.raises(mr.MetalRoboError, match="physics_substeps"):
        mr.World(model, bad)
    bad = model.default_config()
    bad.pair_capacity = 0
    with pytest.raises(mr.MetalRoboError, match="pair_capacity"):
        mr.World(model, bad)


def test_last_error_is_thread_local():
    b = mr.load()
    cfg = nat.SyntheticWorldConfigC(99, 4, 0)
    assert not b.lib.mr_world_model_create_synthetic(ct.byref(cfg))
    seen = []

    def other():
        seen.append(b.lib.mr_last_error())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen == [b""]
    assert b.last_error() != ""


def test_model_validation_rejects_malformed_trees():
    import model_util as mu

    links = [mu.body(1.0, [0.01] * 3), mu.body(1.0, [0.01] * 3, parent=2, joint_type=0, q_index=0, v_index=0),
             mu.body(1.0, [0.01] * 3, parent=0, joint_type=0, q_index=1, v_index=1)]
    with pytest.raises(mr.MetalRoboError, match="parent before child"):
        mu.CustomWorld(links, [mu.dof(1, 0), mu.dof(2, 1)], [], link_count=3, root_type=0, nq=2, nv=2)


def test_task_rollout_rejects_unknown_source():
    with pytest.raises(mr.MetalRoboError, match="unsupported run source"):
        mr.TaskRollout(7, 4)


def test_synthetic_models_have_the_documented_shape():
    chain = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 4, 1)
    h = chain.models()
    assert np.all(h["nv"] == 7) and np.all(h["link_count"] == 8) and np.all(h["shape_count"] == 28)
    # 15 capsule-capsule + 140 capsule-box + 190 box-box + 7 capsule-plane + 20 box-plane
    assert np.all(h["pair_count"] == 372)
    hum = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 2, 1)
    h = hum.models()
    assert h["nq"][0] == 36 and h["nv"][0] == 35 and h["link_count"][0] == 30 and h["shape_count"][0] == 31
    assert hum.heights().shape == (256, 256)
    assert abs(float(hum.heights().max()) - 0.05) < 1e-6 and abs(float(hum.heights().min()) + 0.05) < 1e-6
