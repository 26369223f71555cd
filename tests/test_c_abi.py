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
    with pytest.raises(mr.MetalRoboError, match="robot-factory"):
        mr.TaskRollout(0, 4)  # MR_RUN_SOURCE_UNITREE_G1: not built here


def _golden_layout():
    import json

    return json.load(open(os.path.join(ROOT, "tests", "golden", "task_rollout_layout.json")))["records"]


def test_task_rollout_records_have_the_reference_layout():
    """Every field of the task-rollout records sits at the reference header's
    offset (fixture written from include/metalrobo/c_api.h by
    tests/golden/task_rollout_layout.py), in C and in the ctypes mirrors."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("trl", os.path.join(ROOT, "tests", "golden", "task_rollout_layout.py"))
    trl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(trl)
    golden = _golden_layout()
    fields = trl.record_fields(os.path.join(ROOT, "include", "mr_c_api.h"))
    assert set(golden) == set(fields)
    for name, fs in fields.items():
        assert set(golden[name]) == set(fs) | {"size"}, name
    ours = trl.layout(os.path.join(ROOT, "include"), "mr_c_api.h", fields)
    assert ours == golden
    mirrors = {"MRTaskRolloutDynamicSphereC": nat.TaskRolloutDynamicSphereC,
               "MRTaskRolloutConfigC": nat.TaskRolloutConfigC, "MRTaskRolloutLayoutC": nat.TaskRolloutLayoutC,
               "MRRunManifestC": nat.RunManifestC, "MRTaskRolloutStageHighWaterC": nat.TaskRolloutStageHighWaterC,
               "MRTaskRolloutAdvanceC": nat.TaskRolloutAdvanceC}
    for name, cls in mirrors.items():
        assert ct.sizeof(cls) == golden[name]["size"], name
        for f in fields[name]:
            assert getattr(cls, f).offset == golden[name][f], (name, f)


def test_policy_pack_records_have_the_reference_layout():
    """MRPolicyDenseLayerC, MRPolicyPackC and MRPolicyRolloutBatchC sit at the
    reference header's offsets (fixture tests/golden/policy_pack_layout.json,
    written from c_api.h:334-394), in C and in the ctypes mirrors; where the
    reference is present its header declares the same fields in order."""
    import importlib.util
    import json

    spec = importlib.util.spec_from_file_location("trl", os.path.join(ROOT, "tests", "golden", "task_rollout_layout.py"))
    trl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(trl)
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "policy_pack_layout.json")))["records"]
    fields = trl.record_fields(os.path.join(ROOT, "include", "mr_packs.h"), trl.PACK_RECORDS)
    assert trl.layout(os.path.join(ROOT, "include"), "mr_packs.h", fields) == golden
    for name, cls in {"MRPolicyDenseLayerC": nat.PolicyDenseLayerC, "MRPolicyPackC": nat.PolicyPackC,
                      "MRPolicyRolloutBatchC": nat.PolicyRolloutBatchC}.items():
        assert ct.sizeof(cls) == golden[name]["size"], name
        assert [f for f, _ in cls._fields_] == fields[name], name
        for f in fields[name]:
            assert getattr(cls, f).offset == golden[name][f], (name, f)
    if os.path.isfile("/root/reference/include/metalrobo/c_api.h"):
        assert trl.record_fields("/root/reference/include/metalrobo/c_api.h", trl.PACK_RECORDS) == fields


@pytest.mark.skipif(not os.path.isfile("/root/reference/include/metalrobo/c_api.h"),
                    reason="the reference header is not present here")
def test_task_rollout_layout_fixture_matches_the_reference_header():
    """Where the reference is present: its header declares the same fields in
    the same order, and the compiler lays them out as the fixture says."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("trl", os.path.join(ROOT, "tests", "golden", "task_rollout_layout.py"))
    trl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(trl)
    ref_fields = trl.record_fields("/root/reference/include/metalrobo/c_api.h")
    assert ref_fields == trl.record_fields(os.path.join(ROOT, "include", "mr_c_api.h"))
    assert trl.layout("/root/reference/include", "metalrobo/c_api.h", ref_fields) == _golden_layout()


def test_task_rollout_manifest_fields_without_meaning_are_rejected():
    """Fields of the reference manifest that name task programs, packs or
    visual sensors must be zero here; a non-zero value is an error, not
    silently ignored."""
    b = mr.load()

    def manifest(**changes):
        m = nat.RunManifestC()
        m.profile.environment_count = 2
        m.source = nat.MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES
        for k, v in changes.items():
            if k.startswith("profile_"):
                setattr(m.profile, k[len("profile_"):], v)
            else:
                setattr(m, k, v)
        return m

    for changes, msg in [({"task": 2}, "task program"), ({"urdf_path": b"robot.urdf"}, "pack paths"),
                         ({"surface": 1}, "surface"), ({"profile_dynamic_sphere_count": 1}, "dynamic_spheres"),
                         ({"profile_interaction_reference_mode": 1}, "interaction"),
                         ({"profile_environment_count": 0}, "environment_count")]:
        m = manifest(**changes)
        assert not b.lib.mr_create_task_rollout(ct.byref(m)), changes
        assert msg in b.last_error(), (changes, b.last_error())
    out = nat.TaskRolloutAdvanceC()
    assert b.lib.mr_task_rollout_advance(None, None, 0, None, 0, 1, 0, 0, None) == -1
    assert "result is null" in b.last_error()
    assert b.lib.mr_task_rollout_advance(None, None, 0, None, 0, 1, 0, 0, ct.byref(out)) == -1
    assert out.first_failing_environment == 0xFFFFFFFF and out.first_failing_control_step == 0xFFFFFFFF


def test_malformed_descriptors_are_rejected_before_use():
    """mr_world_model_create validates the descriptor before reading through
    it: out-of-range model indices, under-stated strides, short heightfields
    and pair classes that contradict their shapes are errors."""
    import model_util as mu

    def world(**kw):
        bodies = [mu.body(1.0, [0.01] * 3), mu.body(1.0, [0.01] * 3), mu.body(0.0, [1, 1, 1], motion=0)]
        shapes = [mu.shape(1, 0, [0.05]), mu.shape(2, 3, [0, 0, 0], rot=mu.Y_TO_Z)]
        return mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, **kw)

    w = world()
    b = mr.load()

    def try_create(mutate):
        desc = nat.WorldModelDescC()
        ct.memmove(ct.byref(desc), ct.byref(w._desc), ct.sizeof(desc))
        keep = mutate(desc)
        h = b.lib.mr_world_model_create(ct.byref(desc), 0.005)
        del keep
        return h

    env_model = np.array([5], np.uint32)
    assert not try_create(lambda d: (setattr(d, "environment_model", env_model.ctypes.data), env_model))
    assert "environment model index" in b.last_error()
    pairs = w.pairs.copy()
    pairs["pair_class"] = 12  # heightfield class on a sphere-plane pair
    assert not try_create(lambda d: (setattr(d, "pairs", pairs.ctypes.data), pairs))
    assert "pair class" in b.last_error()
    hdr = w.header.copy()
    hdr["body_count"] = 40
    hdr["scene_body_count"] = 39
    assert not try_create(lambda d: (setattr(d, "models", hdr.ctypes.data), hdr))
    assert "offsets outside" in b.last_error()
    # an under-stated stride is raised to the model's need, never read past
    h = try_create(lambda d: (setattr(d, "max_scene_bodies", 0), None))
    assert h
    b.lib.mr_world_model_destroy(h)
    heights = np.zeros((4, 4), np.float32)
    hf = mu.CustomWorld([mu.body(1.0, [0.01] * 3), mu.body(1.0, [0.01] * 3), mu.body(0.0, [1, 1, 1], motion=0)], [],
                        [mu.shape(1, 0, [0.05]), mu.shape(2, 7, [0, 0, 0])], link_count=1, root_type=0, nq=0, nv=0,
                        heights=heights, hf_origin=(0, 0), hf_spacing=0.1)
    desc = nat.WorldModelDescC()
    ct.memmove(ct.byref(desc), ct.byref(hf._desc), ct.sizeof(desc))
    desc.heightfield_vertex_count = 8  # fewer heights than nx * ny
    assert not b.lib.mr_world_model_create(ct.byref(desc), 0.005)
    assert "heightfield" in b.last_error()


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
