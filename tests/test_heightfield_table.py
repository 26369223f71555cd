"""Heightfield query table at its limit.

The kernel evaluates the heightfield query points of all active pairs one per
lane into a table of 128 slots; a pair whose queries do not fit evaluates
them in its own lane. This world holds 15 boxes (8 queries each) and 3
capsules (3 each) on a non-planar heightfield: 129 queries, so the last
capsule's three queries straddle the table end (slots 126-128). The GPU must
read only the slots the fitting pairs wrote and match the oracle bit for bit.
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol

SPHERE, CAPSULE, BOX, PLANE, HEIGHTFIELD = 0, 1, 2, 3, 7


def straddling_world(envs=2):
    n = 12
    xs = np.arange(n) * 0.1 - 0.55
    X, Y = np.meshgrid(xs, xs)
    heights = (0.02 * np.sin(3.0 * X) * np.cos(2.0 * Y)).astype(np.float32)
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01])]  # fixed root link, no shape
    shapes = []
    spots = [(-0.44 + 0.22 * (i % 5), -0.33 + 0.22 * (i // 5)) for i in range(18)]
    kinds = [BOX] * 15 + [CAPSULE] * 3
    for i, kind in enumerate(kinds):
        bodies.append(mu.body(0.5, [0.001] * 3))
    bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
    hf_body = len(bodies) - 1
    shapes.append(mu.shape(hf_body, HEIGHTFIELD, [0, 0, 0]))  # collider 0: its pairs come first
    for i, kind in enumerate(kinds):
        dims = [0.04, 0.04, 0.04] if kind == BOX else [0.03, 0.05]
        shapes.append(mu.shape(1 + i, kind, dims))
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, envs=envs, heights=heights,
                       hf_origin=(-0.55, -0.55), hf_spacing=0.1)
    cfg = w.config(pair_capacity=64, manifold_capacity=32, constraint_capacity=128)
    q, v, scene = w.model.initial_state()
    rng = np.random.default_rng(5)
    for e in range(envs):
        for i, kind in enumerate(kinds):
            x, y = spots[i]
            z = 0.055 if kind == BOX else 0.04
            quat = np.array([0, 0, 0, 1.0]) + 0.05 * rng.normal(size=4)
            quat /= np.linalg.norm(quat)
            scene[e, i] = mu.scene_state((x, y, z + 0.002 * e), rot=tuple(quat.astype(np.float32)), inv_mass=2.0)
        scene[e, len(kinds)] = mu.scene_state((0, 0, 0), motion=0)
    return w, cfg, q, v, scene


def test_world_puts_129_queries_across_the_table_end():
    """Pins the scene: every heightfield pair is active and their queries in
    canonical pair order are 15 x 8 + 3 x 3 = 129, the last pair starting at 126."""
    w, cfg, q, v, scene = straddling_world(envs=1)
    ov, cnt, _, _ = ol.collide(w.model, cfg, 0, q[0], v[0], scene[0])
    hf_pairs = [k for k, p in enumerate(w.pairs) if int(p["pair_class"]) == 12]
    assert len(hf_pairs) == 18 and hf_pairs == list(range(18))
    assert np.all(ov[hf_pairs] == 1)
    queries = [8 if int(w.shapes[int(w.pairs[k]["collider_b"])]["shape_type"]) == BOX else 3 for k in hf_pairs]
    first = np.cumsum([0] + queries[:-1])
    assert sum(queries) == 129 and first[-1] == 126 and first[-1] + queries[-1] > 128
    # the terrain is not planar under the bodies: contact normals differ
    assert np.sum(cnt[hf_pairs]) > 18


@pytest.mark.gpu
def test_gpu_straddling_query_table_matches_oracle(gpu):
    w, cfg, q, v, scene = straddling_world(envs=2)
    world = mr.World(w.model, cfg)
    world.set_state(q, v, scene)
    st = ol.OracleState(w.model, cfg)
    st.scene[:] = scene
    steps = 3
    cmds = np.zeros((steps, 2, 0), np.float32)
    for k in range(steps):
        print(f"heightfield table: control step {k}", flush=True)
        world.advance(cmds[k:k + 1])
        o_status, o_cs, *_ = st.step(cmds[k], control_step=k)
        g = world.statuses()[0]
        gcs = world.contact_statuses()[0]
        np.testing.assert_array_equal(g["code"], o_status["code"])
        np.testing.assert_array_equal(gcs["active_pairs"], o_cs["active_pairs"])
        np.testing.assert_array_equal(gcs["raw_contacts"], o_cs["raw_contacts"])
        _, _, s = world.state()
        assert np.array_equal(s.view(np.uint8), st.scene.view(np.uint8)), f"step {k}: scene bodies differ"
        _, _, mcount = world.manifolds()
        np.testing.assert_array_equal(mcount, st.mcount)
    assert np.all(world.statuses()[0]["code"] == 0)
