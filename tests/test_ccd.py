"""Hybrid continuous collision (MR_CCD_HYBRID): microsteps split at the
earliest time of impact of CCD-enabled pairs.

CPU (the oracle against closed forms):
  * a 60 m/s sphere crosses a 1 cm wall within one microstep: the speculative
    mode lets it through, the hybrid mode stops it at the wall;
  * two spheres closing head-on at 80 m/s: the first segment ends at the
    closed-form time of impact (d - R) / v, so they end touching, with zero
    total momentum and no relative velocity (restitution 0);
  * two impacts in one microstep need two segments: max_ccd_events = 1 fails
    the step with MR_STEP_CCD_EVENT_BUDGET_EXHAUSTED, 2 passes;
  * a fast CCD sphere over a heightfield has no time-of-impact query: the
    step fails with MR_STEP_DID_NOT_CONVERGE, unless its motion stays inside
    the speculative contact offset;
  * CCD flags change nothing in the speculative mode.
GPU: the same scenes and a randomised shooting gallery match the oracle bit
for bit (state, statuses and the impact-segment count).
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol

SPHERE, CAPSULE, BOX, PLANE, HEIGHTFIELD = 0, 1, 2, 3, 7
CCD = 2  # MR_SHAPE_FLAG_ENABLE_CCD
HYBRID = 1
BUDGET, DID_NOT_CONVERGE = 9, 11


def ccd_config(w, **kw):
    args = dict(ccd_mode=HYBRID, max_ccd_events=4, max_ccd_iterations=16, ccd_distance_tolerance=1e-3,
                ccd_time_tolerance=1e-5, ccd_minimum_advance=1e-6)
    args.update(kw)
    return w.config(**args)


def flagged(s, on=True):
    if on:
        s["flags"] |= CCD
    return s


def scene_world(movers, statics, envs=1, ccd=True, heights=None, hf_origin=None, hf_spacing=None):
    """A fixed dummy link, dynamic movers (shape_type, dims, mass) and static
    shapes (shape_type, dims, body position, body rotation), no gravity; the
    positions go into the scene states (`static`), each shape sits at its
    body's origin."""
    bodies = [mu.body(1.0, [0.01] * 3)]
    shapes = []
    for k, (stype, dims, mass) in enumerate(movers):
        r = max(dims)
        bodies.append(mu.body(mass, [0.4 * mass * r * r] * 3))
        shapes.append(flagged(mu.shape(1 + k, stype, dims, contact_offset=0.01), ccd))
    for stype, dims, pos, rot in statics:
        bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
        shapes.append(mu.shape(len(bodies) - 1, stype, dims, rot=rot, contact_offset=0.01))
    return mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, envs=envs,
                          gravity=(0, 0, 0), heights=heights, hf_origin=hf_origin, hf_spacing=hf_spacing)


def place(st, env, states):
    for k, s in enumerate(states):
        st.scene[env, k] = s


def mover(pos, lin, mass, ang=(0, 0, 0)):
    return mu.scene_state(pos, lin=lin, ang=ang, inv_mass=1.0 / mass)


def static(pos=(0, 0, 0)):
    return mu.scene_state(pos, motion=0)


def bullet_and_wall(ccd=True):
    w = scene_world([(SPHERE, [0.02], 0.1)], [(BOX, [0.005, 0.5, 0.5], (0.5, 0, 0), (0, 0, 0, 1))], ccd=ccd)
    return w, [mover((0.0, 0, 0), (60.0, 0, 0), 0.1), static((0.5, 0, 0))]


def run_oracle(w, cfg, states, steps=1):
    st = ol.OracleState(w.model, cfg)
    for e in range(w.envs):
        place(st, e, states[e] if isinstance(states[0], list) else states)
    out = []
    for k in range(steps):
        out.append(st.step(np.zeros((w.envs, 0), np.float32), control_step=k)[:2])
    return st, out


def test_speculative_mode_lets_the_bullet_through_the_wall():
    w, states = bullet_and_wall()
    st, out = run_oracle(w, w.config(), states)
    assert out[0][0]["code"][0] == 0
    x = float(st.scene[0, 0]["position"][0])
    assert x > 0.5 + 0.005 + 0.02, x  # entirely behind the wall
    assert out[0][1]["diagnostics"][0, 3] == 0.0


def test_hybrid_mode_stops_the_bullet_at_the_wall():
    w, states = bullet_and_wall()
    st, out = run_oracle(w, ccd_config(w), states)
    status, cs = out[0]
    assert status["code"][0] == 0
    x = float(st.scene[0, 0]["position"][0])
    vx = float(st.scene[0, 0]["linear_velocity_inv_mass"][0])
    face = 0.5 - 0.005 - 0.02  # sphere centre at contact with the near face
    assert face - 0.011 < x <= face + 0.002, x
    assert abs(vx) < 0.5, vx  # the contact stopped it (restitution 0)
    assert cs["diagnostics"][0, 3] == 1.0  # one segment ended at a time of impact


def test_head_on_spheres_end_touching_at_the_closed_form_time_of_impact():
    """Gap 0.56 m between surfaces, closing at 80 m/s: impact 7 ms into the
    10 ms microstep. The first segment ends there with the spheres touching;
    the second resolves the contact."""
    w = scene_world([(SPHERE, [0.02], 0.1), (SPHERE, [0.02], 0.1)], [])
    states = [mover((-0.3, 0, 0), (40.0, 0, 0), 0.1), mover((0.3, 0, 0), (-40.0, 0, 0), 0.1)]
    cfg = ccd_config(w, physics_substeps=1, control_timestep=0.01)
    st, out = run_oracle(w, cfg, states)
    assert out[0][0]["code"][0] == 0 and out[0][1]["diagnostics"][0, 3] == 1.0
    xa, xb = (float(st.scene[0, k]["position"][0]) for k in range(2))
    va, vb = (float(st.scene[0, k]["linear_velocity_inv_mass"][0]) for k in range(2))
    assert abs(va + vb) < 1e-4  # momentum conserved
    assert abs(vb - va) < 1e-3  # no relative normal velocity
    # at the impact the centres are 2R apart; after it the pair stays there
    assert abs(xb - xa - 0.04) < 2e-4, (xa, xb)
    assert abs(xa + 0.02) < 2e-4 and abs(xb - 0.02) < 2e-4
    # without the split (speculative mode) they pass through each other
    st2, _ = run_oracle(w, w.config(physics_substeps=1, control_timestep=0.01), states)
    assert float(st2.scene[0, 0]["position"][0]) > float(st2.scene[0, 1]["position"][0])


def two_walls():
    w = scene_world([(SPHERE, [0.02], 0.1), (SPHERE, [0.02], 0.1)],
                    [(BOX, [0.005, 0.5, 0.5], (0.5, 0, 0), (0, 0, 0, 1)),
                     (BOX, [0.005, 0.5, 0.5], (-0.6, 0, 0), (0, 0, 0, 1))])
    states = [mover((0.0, 0, 0), (60.0, 0, 0), 0.1), mover((-0.1, 0, 0), (-60.0, 0, 0), 0.1),
              static((0.5, 0, 0)), static((-0.6, 0, 0))]
    return w, states


def test_event_budget():
    w, states = two_walls()
    # the walls are 0.475 and 0.475 m away at t=0 ... and 7.9 ms vs 7.9 ms: stagger the second
    states[1] = mover((-0.1, 0, 0), (-50.0, 0, 0), 0.1)
    _, out = run_oracle(w, ccd_config(w, max_ccd_events=1), states)
    assert out[0][0]["code"][0] == BUDGET
    assert out[0][0]["failing_substep"][0] == 1
    st, out = run_oracle(w, ccd_config(w, max_ccd_events=2), states)
    assert out[0][0]["code"][0] == 0
    assert out[0][1]["diagnostics"][0, 3] == 2.0
    assert float(st.scene[0, 0]["position"][0]) < 0.5 and float(st.scene[0, 1]["position"][0]) > -0.6


def spinning_bullet():
    """A 4 cm box spinning at 300 rad/s while flying at 60 m/s at the wall: its
    rotation makes the closing-speed bound conservative, so a time-of-impact
    query advances in several steps."""
    w = scene_world([(BOX, [0.02, 0.02, 0.02], 0.1)], [(BOX, [0.005, 0.5, 0.5], (0.5, 0, 0), (0, 0, 0, 1))])
    return w, [mover((0.0, 0, 0), (60.0, 0, 0), 0.1, ang=(0.0, 0.0, 300.0)), static((0.5, 0, 0))]


def test_an_unresolved_pair_splits_the_microstep_instead_of_failing():
    """With one conservative-advancement iteration per query the spinning
    bullet's pair is not resolved within its query (which used to fail the step
    with MR_STEP_DID_NOT_CONVERGE). The query still proves the pair separated up
    to the time it advanced, so the microstep ends there; from that state the
    pair is within its contact distance and the segment's speculative contacts
    stop the bullet at the wall."""
    w, states = spinning_bullet()
    for iterations in (16, 1):
        st, out = run_oracle(w, ccd_config(w, max_ccd_iterations=iterations, max_ccd_events=16), states)
        assert out[0][0]["code"][0] == 0, iterations
        assert out[0][1]["diagnostics"][0, 3] >= 1.0  # the microstep was split
        x = float(st.scene[0, 0]["position"][0])
        assert 0.5 - 0.005 - 0.04 < x < 0.5, (iterations, x)  # at the wall's near face, not through it


def test_hybrid_impulses_record_their_segment_duration():
    """A manifold's impulses come from one segment's solve; in the hybrid mode
    the point records that segment's duration (impulses.w), and the next
    segment's warm start scales them by its own duration over it. The
    speculative mode keeps w = 0 and unscaled warm starts."""
    w, states = bullet_and_wall()
    cfg = ccd_config(w)
    st, out = run_oracle(w, cfg, states)
    mh = st.mh[0, :int(st.mcount[0])]
    pts = st.mp[0, :int(st.mcount[0])]
    live = np.arange(4)[None, :] < mh["pair_and_count"][:, 3:4]
    assert live.any()
    h = cfg.control_timestep / cfg.physics_substeps
    wv = pts["impulses"][..., 3][live]
    assert (wv > 0.0).all() and (wv <= h * (1.0 + 1e-6)).all(), wv
    # speculative mode: a slow sphere that comes to rest against the wall
    slow = [mover((0.5 - 0.005 - 0.02 - 0.002, 0, 0), (2.0, 0, 0), 0.1), static((0.5, 0, 0))]
    st, out = run_oracle(w, w.config(), slow)
    mh = st.mh[0, :int(st.mcount[0])]
    live = np.arange(4)[None, :] < mh["pair_and_count"][:, 3:4]
    assert live.any()
    assert (st.mp[0, :int(st.mcount[0])]["impulses"][..., 3][live] == 0.0).all()


def heightfield_world(ccd=True):
    heights = np.zeros((8, 8), np.float32)
    w = scene_world([(SPHERE, [0.05], 1.0)], [(HEIGHTFIELD, [0, 0, 0], (0, 0, 0), (0, 0, 0, 1))], ccd=ccd,
                    heights=heights, hf_origin=(-1.0, -1.0), hf_spacing=2.0 / 7.0)
    return w


def test_a_pair_without_a_time_of_impact_query_fails_unless_speculative_contacts_cover_it():
    w = heightfield_world()
    fast = [mover((0, 0, 0.5), (0, 0, -40.0), 1.0), static()]
    _, out = run_oracle(w, ccd_config(w), fast)
    assert out[0][0]["code"][0] == DID_NOT_CONVERGE
    slow = [mover((0, 0, 0.5), (0, 0, -1.0), 1.0), static()]  # 5 mm per microstep, inside the 1 cm offset
    _, out = run_oracle(w, ccd_config(w), slow)
    assert out[0][0]["code"][0] == 0


def test_flags_change_nothing_in_the_speculative_mode():
    w1, states = bullet_and_wall(ccd=True)
    w0, _ = bullet_and_wall(ccd=False)
    st1, _ = run_oracle(w1, w1.config(), states, steps=3)
    st0, _ = run_oracle(w0, w0.config(), states, steps=3)
    assert np.array_equal(st1.scene.view(np.uint8), st0.scene.view(np.uint8))


S = float(np.sqrt(0.5))
# inward-facing planes of the room |x|, |y| <= 0.6, 0 <= z <= 0.9: (position, rotation taking +y to the normal)
ROOM = [((0.6, 0, 0), (0, 0, S, S)), ((-0.6, 0, 0), (0, 0, -S, S)), ((0, 0.6, 0), (0, 0, 1, 0)),
        ((0, -0.6, 0), (0, 0, 0, 1)), ((0, 0, 0), (S, 0, 0, S)), ((0, 0, 0.9), (-S, 0, 0, S))]


def gallery(envs=16, seed=3, ccd=True):
    """Randomised shooting gallery in a closed room of six planes: per
    environment a sphere, a box and a capsule fired at 20-80 m/s, spinning,
    at the walls and each other."""
    rng = np.random.default_rng(seed)
    w = scene_world([(SPHERE, [0.03], 0.2), (BOX, [0.03, 0.02, 0.04], 0.3), (CAPSULE, [0.02, 0.05], 0.2)],
                    [(PLANE, [0, 0, 0], pos, rot) for pos, rot in ROOM], envs=envs, ccd=ccd)
    states = []
    for e in range(envs):
        row = []
        for k in range(3):
            pos = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.2, 0.7))
            speed = rng.uniform(20.0, 80.0)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            row.append(mover(pos, tuple(speed * d), [0.2, 0.3, 0.2][k], ang=tuple(rng.normal(size=3) * 5.0)))
        row += [static(pos) for pos, _ in ROOM]
        states.append(row)
    return w, states


def inside_room(st, margin=0.02):
    p = st.scene[:, :3]["position"][:, :, :3]
    return (np.abs(p[..., 0]) < 0.6 + margin) & (np.abs(p[..., 1]) < 0.6 + margin) & (p[..., 2] > -margin) & \
        (p[..., 2] < 0.9 + margin)


def test_gallery_keeps_every_mover_in_the_room():
    w, states = gallery(envs=8)
    st, out = run_oracle(w, ccd_config(w, max_ccd_events=16), states, steps=4)
    for status, cs in out:
        assert np.all(status["code"] == 0), status["code"]
    assert sum(float(cs["diagnostics"][:, 3].sum()) for _, cs in out) > 0
    assert inside_room(st).all(), st.scene[:, :3]["position"]
    # the speculative mode loses movers through the walls
    w0, states0 = gallery(envs=8, ccd=False)
    st0, _ = run_oracle(w0, w0.config(), states0, steps=4)
    assert not inside_room(st0).all()


# ------------------------------------------------------------------ GPU


def gpu_matches(w, cfg, states, steps):
    st, out = run_oracle(w, cfg, states, steps=0)
    world = mr.World(w.model, cfg)
    world.set_state(st.q, st.v, st.scene)
    for k in range(steps):
        world.advance(np.zeros((1, w.envs, 0), np.float32))
        o_status, o_cs, *_ = st.step(np.zeros((w.envs, 0), np.float32), control_step=k)
        g = world.statuses()[0]
        gc = world.contact_statuses()[0]
        np.testing.assert_array_equal(g["code"], o_status["code"])
        np.testing.assert_array_equal(g["failing_index"], o_status["failing_index"])
        np.testing.assert_array_equal(gc["diagnostics"][:, 3], o_cs["diagnostics"][:, 3])
        _, _, scene = world.state()
        assert np.array_equal(scene.view(np.uint8), st.scene.view(np.uint8)), f"step {k}: scene"
    return st


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bullet", "head_on", "budget", "heightfield", "split"])
def test_gpu_ccd_scenes_match_oracle(gpu, name):
    print(f"ccd scene {name}", flush=True)
    if name == "split":
        # one advancement iteration: the query ends the segment at the time it proved safe
        w, states = spinning_bullet()
        cfg = ccd_config(w, max_ccd_iterations=1, max_ccd_events=16)
    elif name == "bullet":
        w, states = bullet_and_wall()
        cfg = ccd_config(w)
    elif name == "head_on":
        w = scene_world([(SPHERE, [0.02], 0.1), (SPHERE, [0.02], 0.1)], [])
        states = [mover((-0.3, 0, 0), (40.0, 0, 0), 0.1), mover((0.3, 0, 0), (-40.0, 0, 0), 0.1)]
        cfg = ccd_config(w, physics_substeps=1, control_timestep=0.01)
    elif name == "budget":
        w, states = two_walls()
        states[1] = mover((-0.1, 0, 0), (-50.0, 0, 0), 0.1)
        cfg = ccd_config(w, max_ccd_events=1)
    else:
        w = heightfield_world()
        states = [mover((0, 0, 0.5), (0, 0, -40.0), 1.0), static()]
        cfg = ccd_config(w)
    gpu_matches(w, cfg, states, steps=2)


@pytest.mark.gpu
def test_gpu_gallery_matches_oracle(gpu):
    w, states = gallery(envs=64)
    print("ccd gallery: 64 environments, 4 control steps", flush=True)
    st = gpu_matches(w, ccd_config(w, max_ccd_events=16), states, steps=4)
    assert inside_room(st).all()
