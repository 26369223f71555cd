"""The oracle's broadphase and narrowphase against the reference's own
known-answer collision scenes (apps/collision_probe.cpp), transcribed as data
in tests/golden/collision_probe_kat.json (CPU only).

Checked per scene, as the reference probe checks them:
  * the candidate pair set (canonical A < B) and, per pair, the number of raw
    contacts, including the just-outside-the-contact-offset capsule pair that
    the padded broadphase keeps but the narrowphase rejects;
  * feature ids ((shape type << 28) | local feature) of every single-contact
    pair, signed separations and normals where the probe states them;
  * the witness convention on every contact: unit normal (2e-6) and
    (p_B - p_A) . n = separation (3e-6);
  * input-order reciprocity (swapping A and B flips the normal, keeps the
    separation);
  * group/mask and static-static filtering;
  * the eight-vertex box patch reduced to a four-point manifold.
"""

import json
import os

import numpy as np
import pytest

import model_util as mu
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "collision_probe_kat.json")))
SCENES = {s["name"]: s for s in KAT["scenes"]}
TOL = KAT["witness_tolerances"]


def feature(shape_type, local):
    return ((shape_type & 0xF) << 28) | (local & 0x0FFFFFFF)


def resolve(scene):
    """Applies a derived scene's edits to its base scene."""
    if "base" not in scene:
        return scene["bodies"], scene["shapes"]
    base = SCENES[scene["base"]]
    bodies = [dict(b) for b in base["bodies"]]
    shapes = [dict(s) for s in base["shapes"]]
    for k, edit in scene.get("edit_bodies", {}).items():
        bodies[int(k)].update(edit)
    for k, edit in scene.get("edit_shapes", {}).items():
        shapes[int(k)].update(edit)
    return bodies, shapes


def build(scene, substeps=1):
    """A world with a fixed, shapeless root link followed by the scene bodies."""
    bodies_spec, shapes_spec = resolve(scene)
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01])]
    states = []
    for b in bodies_spec:
        motion = 0 if b["motion"] == "static" else 2
        bodies.append(mu.body(1.0 if motion == 2 else 0.0, [0.01, 0.01, 0.01] if motion == 2 else [1, 1, 1],
                              motion=motion))
        states.append(mu.scene_state(b["pos"], b["rot"], inv_mass=1.0 if motion == 2 else 0.0, motion=motion))
    shapes = []
    for s in shapes_spec:
        sh = mu.shape(1 + s["body"], s["type"], s["dims"], rot=tuple(s.get("local_rot", (0, 0, 0, 1))),
                      contact_offset=s["offset"])
        sh["collision_group"] = s["group"]
        sh["collision_mask"] = s["mask"]
        shapes.append(sh)
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0)
    scene_state = np.array(states, w.model.initial_state()[2].dtype)[None, :]
    return w, scene_state


def collide(scene):
    w, state = build(scene)
    ov, cnt, reals, feats = ol.collide(w.model, w.config(), 0, np.zeros(1, np.float32), np.zeros(1, np.float32),
                                       state[0])
    out = {}
    for k, p in enumerate(w.pairs):
        if ov[k]:
            key = f"{int(p['collider_a'])},{int(p['collider_b'])}"
            n = int(cnt[k])
            out[key] = {"normal": reals[k, :n, :3].astype(np.float64), "sep": reals[k, :n, 3].astype(np.float64),
                        "pa": reals[k, :n, 4:7].astype(np.float64), "pb": reals[k, :n, 7:10].astype(np.float64),
                        "feat": feats[k, :n].copy()}
    return out


def check_witnesses(frame):
    for key, c in frame.items():
        for i in range(len(c["sep"])):
            n = c["normal"][i]
            assert abs(np.linalg.norm(n) - 1.0) <= TOL["unit_normal"], key
            proj = float(np.dot(c["pb"][i] - c["pa"][i], n))
            assert abs(proj - c["sep"][i]) <= TOL["separation_convention"], (key, proj, c["sep"][i])


@pytest.mark.parametrize("name", ["primary", "four_point_reduction", "capsule_pairs", "sphere_box",
                                  "capsule_box_box"])
def test_known_answer_scene(name):
    scene = SCENES[name]
    exp = scene["expect"]
    frame = collide(scene)
    assert sorted(frame) == sorted(f"{a},{b}" for a, b in exp["pairs"])
    check_witnesses(frame)
    if "contacts" in exp:
        for key, n in exp["contacts"].items():
            assert len(frame[key]["sep"]) == n, (key, len(frame[key]["sep"]), n)
        assert sum(len(c["sep"]) for c in frame.values()) == exp["raw_contacts"]
    if "min_contacts" in exp:
        for key, n in exp["min_contacts"].items():
            assert len(frame[key]["sep"]) >= n, key
        assert sum(len(c["sep"]) for c in frame.values()) >= exp["min_raw_contacts"]
    for key, ((ta, la), (tb, lb)) in exp.get("features", {}).items():
        assert tuple(frame[key]["feat"][0]) == (feature(ta, la), feature(tb, lb)), key
    for key, (value, tol) in exp.get("separations", {}).items():
        assert abs(frame[key]["sep"][0] - value) <= tol, (key, frame[key]["sep"][0], value)
    for key, (components, tol) in exp.get("normals", {}).items():
        for axis, value in enumerate(components):
            if value is not None:
                assert abs(frame[key]["normal"][0][axis] - value) <= tol, (key, axis)
    for ka, kb in exp.get("reciprocal", []):
        assert abs(frame[ka]["normal"][0][0] + frame[kb]["normal"][0][0]) <= 2.0e-6, (ka, kb)
        assert abs(frame[ka]["sep"][0] - frame[kb]["sep"][0]) <= 2.0e-6, (ka, kb)


@pytest.mark.parametrize("name", ["primary_masked", "primary_static"])
def test_filters(name):
    scene = SCENES[name]
    frame = collide(scene)
    banned = set(scene["expect"]["no_pairs_with"])
    for key in frame:
        a, b = (int(x) for x in key.split(","))
        assert a not in banned and b not in banned, key
    assert frame, "the filter removed every pair"


def test_eight_point_patch_reduces_to_four_manifold_points():
    scene = SCENES["four_point_reduction"]
    w, state = build(scene)
    cfg = w.config()
    cfg.physics_substeps = 1
    st = ol.OracleState(w.model, cfg)
    st.scene[:] = state
    status, cs, *_ = st.step(np.zeros((1, 0), np.float32))
    assert int(cs["raw_contacts"][0]) == 8
    assert int(st.mcount[0]) == 1
    assert int(st.mh[0, 0]["pair_and_count"][3]) == scene["expect"]["manifold_points"]["0,1"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["primary", "four_point_reduction", "capsule_pairs", "sphere_box",
                                  "capsule_box_box", "primary_masked"])
def test_gpu_known_answer_scene_matches_oracle(gpu, name):
    """One control step of each known-answer scene on the GPU: broadphase
    evidence, manifold features and counts, statuses and scene states equal
    the oracle's bit for bit."""
    import metalrobo as mr

    w, state = build(SCENES[name])
    cfg = w.config()
    cfg.capture_evidence = 1
    cfg.capture_substep = 0
    world = mr.World(w.model, cfg)
    q, v, _ = w.model.initial_state()
    world.set_state(q, v, state)
    st = ol.OracleState(w.model, cfg)
    st.scene[:] = state
    o_status, o_cs, _, o_pairs, o_counts = st.step(np.zeros((1, 0), np.float32))
    world.advance(np.zeros((1, 1, 0), np.float32))
    g_status = world.statuses()[0]
    g_pairs, g_counts = world.evidence()
    np.testing.assert_array_equal(g_status["code"], o_status["code"])
    np.testing.assert_array_equal(g_counts, o_counts)
    n = int(o_counts[0])
    np.testing.assert_array_equal(g_pairs[0, :n], o_pairs[0, :n])
    mh, mp, mc = world.manifolds()
    np.testing.assert_array_equal(mc, st.mcount)
    m = int(mc[0])
    np.testing.assert_array_equal(mh[0, :m]["pair_and_count"], st.mh[0, :m]["pair_and_count"])
    np.testing.assert_array_equal(mp[0, :m]["feature_and_life"], st.mp[0, :m]["feature_and_life"])
    _, _, s = world.state()
    assert np.array_equal(s.view(np.uint8), st.scene.view(np.uint8))
