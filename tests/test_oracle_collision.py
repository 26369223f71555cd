"""Pins the oracle's collider projection, broadphase and narrowphase (CPU only).

Closed forms for the analytic primitive pairs (the reference covers the same
classes in apps/collision_probe.cpp:674-1675 with analytic expectations),
stable feature keys (Collision.metal:1415 featureKey), the contact-normal
convention (A -> B), and a brute-force FP64 AABB broadphase on random scenes:
the FP32 pair list must contain every exact overlap (no false negatives,
NUMERICS.md:144-151) and may add only pairs within the conservative pad.
"""

import numpy as np
import pytest

import model_util as mu
import oracle_lib as ol

SPHERE, CAPSULE, BOX, PLANE, HEIGHTFIELD = 0, 1, 2, 3, 7


def feature(shape_type, local):
    return ((shape_type & 0xF) << 28) | (local & 0x0FFFFFFF)


def make_scene(shapes_spec, with_plane=True):
    """shapes_spec: list of (type, dims, pos, rot). One free body per shape, then a plane."""
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01])]  # unused fixed root (no shapes)
    shapes = []
    states = []
    for i, (t, dims, pos, rot) in enumerate(shapes_spec):
        bodies.append(mu.body(1.0, [0.01, 0.01, 0.01]))
        shapes.append(mu.shape(1 + i, t, dims, contact_offset=0.01))
        states.append(mu.scene_state(pos, rot, inv_mass=1.0))
    if with_plane:
        bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
        shapes.append(mu.shape(len(bodies) - 1, PLANE, [0, 0, 0], rot=mu.Y_TO_Z, contact_offset=0.01))
        states.append(mu.scene_state((0, 0, 0), motion=0))
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0)
    scene = np.array(states, w.model.initial_state()[2].dtype)[None, :]
    return w, scene


def collide(w, scene):
    cfg = w.config()
    return ol.collide(w.model, cfg, 0, np.zeros(1, np.float32), np.zeros(1, np.float32), scene[0])


def pair_index(w, a, b):
    for k, p in enumerate(w.pairs):
        if p["collider_a"] == a and p["collider_b"] == b:
            return k
    raise KeyError((a, b))


def test_sphere_plane_closed_form():
    w, scene = make_scene([(SPHERE, [0.1], (0.3, -0.2, 0.05), (0, 0, 0, 1))])
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert ov[k] == 1 and cnt[k] == 1
    n, sep, pa, pb = reals[k, 0, :3], reals[k, 0, 3], reals[k, 0, 4:7], reals[k, 0, 7:10]
    np.testing.assert_allclose(n, [0, 0, -1], atol=1e-6)  # sphere (A) toward plane (B)
    assert abs(sep - (-0.05)) < 1e-6
    np.testing.assert_allclose(pa, [0.3, -0.2, -0.05], atol=1e-6)
    np.testing.assert_allclose(pb, [0.3, -0.2, 0.0], atol=1e-6)
    assert tuple(feats[k, 0]) == (feature(SPHERE, 0), feature(PLANE, 0))


def test_sphere_above_contact_distance_has_no_witness():
    w, scene = make_scene([(SPHERE, [0.1], (0, 0, 0.13), (0, 0, 0, 1))])
    ov, cnt, _, _ = collide(w, scene)
    assert cnt[pair_index(w, 0, 1)] == 0  # separation 0.03 > contact offsets 0.02


def test_capsule_plane_two_endpoints():
    rot = mu.Y_TO_Z * 0
    rot = np.array([0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)], np.float32)  # local +Y -> world -X (horizontal)
    w, scene = make_scene([(CAPSULE, [0.05, 0.2], (0, 0, 0.04), rot)])
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 2
    np.testing.assert_allclose(reals[k, :2, 3], [-0.01, -0.01], atol=1e-6)
    xs = sorted(reals[k, :2, 4])
    np.testing.assert_allclose(xs, [-0.2, 0.2], atol=1e-6)
    assert sorted(feats[k, :2, 0].tolist()) == [feature(CAPSULE, 0), feature(CAPSULE, 1)]


def test_box_resting_on_plane_gives_bottom_vertices():
    w, scene = make_scene([(BOX, [0.1, 0.1, 0.1], (0, 0, 0.1), (0, 0, 0, 1))])
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 4
    np.testing.assert_allclose(reals[k, :4, 3], 0.0, atol=1e-6)
    assert sorted(feats[k, :4, 0].tolist()) == [feature(BOX, v) for v in range(4)]


def test_sphere_sphere_normal_and_depth():
    w, scene = make_scene([(SPHERE, [0.1], (0, 0, 1), (0, 0, 0, 1)),
                           (SPHERE, [0.1], (0.09, 0.12, 1), (0, 0, 0, 1))], with_plane=False)
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 1
    np.testing.assert_allclose(reals[k, 0, :3], [0.6, 0.8, 0.0], atol=1e-6)
    assert abs(reals[k, 0, 3] - (0.15 - 0.2)) < 1e-6


def test_crossing_capsules_closest_interior_points():
    rx = np.array([0, 0, np.sin(-np.pi / 4), np.cos(-np.pi / 4)], np.float32)  # +Y -> +X
    w, scene = make_scene([(CAPSULE, [0.05, 0.3], (0, 0, 1), rx),
                           (CAPSULE, [0.05, 0.3], (0, 0, 1.08), (0, 0, 0, 1))], with_plane=False)
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 1
    np.testing.assert_allclose(reals[k, 0, :3], [0, 0, 1], atol=1e-6)
    assert abs(reals[k, 0, 3] - (0.08 - 0.1)) < 1e-6
    assert tuple(feats[k, 0]) == (feature(CAPSULE, 2), feature(CAPSULE, 2))


def test_sphere_over_box_face():
    w, scene = make_scene([(SPHERE, [0.05], (0.02, -0.03, 0.14), (0, 0, 0, 1)),
                           (BOX, [0.1, 0.2, 0.1], (0, 0, 0), (0, 0, 0, 1))], with_plane=False)
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 1
    np.testing.assert_allclose(reals[k, 0, :3], [0, 0, -1], atol=1e-6)
    assert abs(reals[k, 0, 3] - (0.04 - 0.05)) < 1e-6
    # region code: x inside (1), y inside (1), z above (2) -> 1 + 3 + 18 = 22
    assert tuple(feats[k, 0]) == (feature(SPHERE, 0), feature(BOX, 22))


def test_stacked_boxes_face_contact():
    w, scene = make_scene([(BOX, [0.1, 0.1, 0.1], (0, 0, 0.1), (0, 0, 0, 1)),
                           (BOX, [0.05, 0.05, 0.05], (0.01, 0.02, 0.249), (0, 0, 0, 1))], with_plane=False)
    ov, cnt, reals, feats = collide(w, scene)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 4
    np.testing.assert_allclose(reals[k, :4, :3], np.tile([0, 0, 1.0], (4, 1)), atol=1e-6)
    np.testing.assert_allclose(reals[k, :4, 3], -0.001, atol=2e-6)
    # B's bottom vertices (ids 0..3) against A's +z face axis (SAT axis 2 -> 128 + 2)
    assert sorted(feats[k, :4, 1].tolist()) == [feature(BOX, v) for v in range(4)]
    assert set(feats[k, :4, 0].tolist()) == {feature(BOX, 130)}


def exact_aabb(t, dims, pos, rot, co):
    pos = np.asarray(pos, np.float64)
    R = np.stack([ol_qrot(rot, e) for e in np.eye(3)], axis=1)
    if t == SPHERE:
        e = np.full(3, dims[0] + co)
        return pos - e, pos + e
    if t == CAPSULE:
        axis = R @ np.array([0, dims[1], 0.0])
        lo = np.minimum(pos - axis, pos + axis) - (dims[0] + co)
        hi = np.maximum(pos - axis, pos + axis) + (dims[0] + co)
        return lo, hi
    ext = np.abs(R) @ np.asarray(dims[:3], np.float64) + co
    return pos - ext, pos + ext


def ol_qrot(q, v):
    q = np.asarray(q, np.float64) / np.linalg.norm(q)
    u = q[:3]
    t = 2 * np.cross(u, v)
    return v + q[3] * t + np.cross(u, t)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_broadphase_superset_of_exact_overlaps(seed):
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(24):
        t = [SPHERE, CAPSULE, BOX][i % 3]
        dims = [rng.uniform(0.03, 0.15), rng.uniform(0.03, 0.2), rng.uniform(0.03, 0.15)]
        pos = rng.uniform(-0.5, 0.5, 3)
        quat = rng.normal(size=4)
        spec.append((t, dims, tuple(pos), tuple((quat / np.linalg.norm(quat)).astype(np.float32))))
    w, scene = make_scene(spec, with_plane=False)
    ov, cnt, _, _ = collide(w, scene)
    exact = []
    for k, p in enumerate(w.pairs):
        a, b = int(p["collider_a"]), int(p["collider_b"])
        la, ha = exact_aabb(spec[a][0], spec[a][1], spec[a][2], spec[a][3], 0.01)
        lb, hb = exact_aabb(spec[b][0], spec[b][1], spec[b][2], spec[b][3], 0.01)
        gap = np.max(np.maximum(la - hb, lb - ha))
        exact.append(gap)
    exact = np.array(exact)
    assert np.all(ov[exact <= 0.0] == 1), "FP32 broadphase missed an exact overlap"
    extra = (ov == 1) & (exact > 0.0)
    assert np.all(exact[extra] < 1e-4), "FP32 broadphase added a pair outside the conservative pad"
    # canonical order: strictly increasing (A, B)
    keys = w.pairs["collider_a"].astype(np.int64) * 4096 + w.pairs["collider_b"]
    assert np.all(np.diff(keys) > 0)


def test_heightfield_flat_patch_matches_plane():
    """On a flat heightfield the support-point test reduces to the plane case."""
    heights = np.zeros((8, 8), np.float32)
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, BOX, [0.1, 0.1, 0.1]), mu.shape(2, HEIGHTFIELD, [0, 0, 0])]
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, heights=heights,
                       hf_origin=(-0.35, -0.35), hf_spacing=0.1)
    states = [mu.scene_state((0.01, 0.02, 0.095), inv_mass=1.0), mu.scene_state((0, 0, 0), motion=0)]
    scene = np.array(states, w.model.initial_state()[2].dtype)
    ov, cnt, reals, feats = ol.collide(w.model, w.config(), 0, np.zeros(1, np.float32), np.zeros(1, np.float32),
                                       scene)
    assert ov[0] == 1 and cnt[0] == 4
    np.testing.assert_allclose(reals[0, :4, 3], -0.005, atol=2e-6)
    np.testing.assert_allclose(reals[0, :4, :3], np.tile([0, 0, -1.0], (4, 1)), atol=1e-6)
    assert all((int(f) >> 28) == HEIGHTFIELD for f in feats[0, :4, 1])


def test_heightfield_slope_normal():
    """A uniformly sloped heightfield z = 0.2 x gives the analytic plane normal."""
    xs = np.arange(8) * 0.1 - 0.35
    heights = np.tile((0.2 * xs).astype(np.float32), (8, 1))
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, SPHERE, [0.05]), mu.shape(2, HEIGHTFIELD, [0, 0, 0])]
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, heights=heights,
                       hf_origin=(-0.35, -0.35), hf_spacing=0.1)
    center = np.array([0.03, 0.01, 0.2 * 0.03 + 0.05 / np.cos(np.arctan(0.2)) - 0.004])
    states = [mu.scene_state(tuple(center), inv_mass=1.0), mu.scene_state((0, 0, 0), motion=0)]
    scene = np.array(states, w.model.initial_state()[2].dtype)
    ov, cnt, reals, feats = ol.collide(w.model, w.config(), 0, np.zeros(1, np.float32), np.zeros(1, np.float32),
                                       scene)
    assert cnt[0] == 1
    n = np.array([-0.2, 0.0, 1.0]) / np.sqrt(1.04)
    np.testing.assert_allclose(reals[0, 0, :3], -n, atol=2e-6)
    assert abs(reals[0, 0, 3] - (-0.004 * np.cos(np.arctan(0.2)))) < 2e-6
