"""Closed-form pins of oracle stages that no reference fixture covers.

Each test derives its expectation from the mathematics of the reference
stage, in FP64, independently of the oracle's code, and is chosen so that a
plausible slip (a dropped term, a flipped sign, a wrong index) fails it:

- implicit PD drive and effective armature (ArticulatedABA.metal:301-315,
  MetalWorld.metal:280-366): one revolute DoF integrated microstep by
  microstep, (I + a + h kd + h^2 kp) qdd = kp (target - q - h v) - kd v;
- dry friction, stick and slip branches (MetalWorld.metal:330-344);
- actuator torque-speed envelope (MetalWorld.metal:252-269, 345-365) at
  speed fractions 0, 1/2 and 1, held over a control step in effort mode and
  refreshed every microstep for implicit drives;
- floating-root SO(3) integration (MetalWorldContact.metal:8876-9043): a
  torque-free isotropic body spinning at constant world-frame omega must
  follow exp(omega t) q0;
- heightfield contacts on a non-planar patch: brute-force point-in-triangle
  over every triangle of the patch (normal, signed distance and feature id
  (HEIGHTFIELD, 2 cell + tri));
- IR evaluation (MetalWorldContact.metal:3883-3960): the stabilisation
  target -h err / (tau^2 (1 + 2 zeta h / tau + (h / tau)^2)) with slop and
  the max-depenetration clamp, restitution on a new impact, and the
  regularisation compliance / h^2 + dissipation / h.
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol
from metalrobo import native as nat

ACTUATED, POSITION_LIMIT, EFFORT_LIMIT, DRIVE, ROOT = 2, 4, 16, 32, 1
SPHERE, CAPSULE, BOX, PLANE, HEIGHTFIELD = 0, 1, 2, 3, 7


# ---------------------------------------------------------------- drives


def one_joint(kp=0.0, kd=0.0, arm=0.0, dry=0.0, flags=ACTUATED | EFFORT_LIMIT | DRIVE, effort_limit=1000.0,
              actuators=None, ix=0.05):
    """A fixed root and one revolute link whose joint axis (x) passes through
    the link's COM: the joint-space inertia is exactly ix."""
    root = mu.body(1.0, [0.01] * 3)
    link = mu.body(2.0, [ix, 0.08, 0.08], parent=0, joint_type=0, q_index=0, v_index=0, axis=(1, 0, 0),
                   parent_anchor=(0, 0, 0.5), child_anchor=(0, 0, 0))
    d = mu.dof(1, 0, flags=flags, limits=(0, 0, 0, effort_limit if flags & EFFORT_LIMIT else 0.0),
               drive=(kp, kd, arm, dry))
    return mu.CustomWorld([root, link], [d], [], link_count=2, root_type=0, nq=1, nv=1, gravity=(0, 0, 0),
                          actuators=actuators)


def run_joint(w, q0, v0, command, implicit, substeps=4, dt=0.02):
    cfg = w.config(actuation_mode=1 if implicit else 0, physics_substeps=substeps, control_timestep=dt,
                   apply_body_damping=0)
    st = ol.OracleState(w.model, cfg)
    st.q[0, 0] = q0
    st.v[0, 0] = v0
    status, *_ = st.step(np.array([[command]], np.float32))
    assert status["code"][0] == 0
    return float(st.q[0, 0]), float(st.v[0, 0])


def integrate(q, v, h, substeps, torque, inertia):
    """Semi-implicit Euler of one DoF: v += h tau(q, v) / I(q, v), then q += h v."""
    for _ in range(substeps):
        v = v + h * torque(q, v) / inertia
        q = q + h * v
    return q, v


@pytest.mark.parametrize("q0,v0,target", [(0.1, 1.5, 0.3), (-0.4, -2.0, 0.25), (0.7, 0.0, -0.2)])
def test_implicit_pd_drive_matches_closed_form(q0, v0, target):
    kp, kd, arm, ix, h = 400.0, 8.0, 0.01, 0.05, 0.005
    w = one_joint(kp=kp, kd=kd, arm=arm, ix=ix)
    q, v = run_joint(w, q0, v0, target, implicit=True)
    tau = lambda q_, v_: kp * (target - q_ - h * v_) - kd * v_
    eq, ev = integrate(q0, v0, h, 4, tau, ix + arm + h * kd + h * h * kp)
    assert abs(v - ev) <= 2e-6 * max(1.0, abs(ev)) and abs(q - eq) <= 2e-6 * max(1.0, abs(eq)), (q, v, eq, ev)
    # the effective armature's h^2 kp term alone moves the answer far outside that bound
    wq, wv = integrate(q0, v0, h, 4, tau, ix + arm + h * kd)
    assert abs(wv - ev) > 1e-3


def test_dry_friction_slip_and_stick():
    kp, kd, arm, ix, h, dry = 400.0, 8.0, 0.01, 0.05, 0.005, 2.0
    inertia = ix + arm + h * kd + h * h * kp
    w = one_joint(kp=kp, kd=kd, arm=arm, dry=dry, ix=ix)
    # slip: |v| > 1e-4 subtracts dry * sign(v) for every microstep
    q0, v0, target = 0.1, 1.5, 0.3
    q, v = run_joint(w, q0, v0, target, implicit=True)
    eq, ev = integrate(q0, v0, h, 4, lambda q_, v_: kp * (target - q_ - h * v_) - kd * v_ - dry * np.sign(v_),
                       inertia)
    assert abs(v - ev) <= 2e-6 * max(1.0, abs(ev)) and abs(q - eq) <= 2e-6
    # stick: at rest with a PD effort inside the friction band nothing moves
    q, v = run_joint(w, 0.0, 0.0, 0.004, implicit=True)  # kp * 0.004 = 1.6 < dry
    assert q == 0.0 and v == 0.0
    # at rest beyond the band the effort is reduced by dry, keeping its sign
    q, v = run_joint(w, 0.0, 0.0, -0.02, implicit=True, substeps=1, dt=h)  # kp * (-0.02) = -8
    assert abs(v - h * (-8.0 + dry) / inertia) <= 1e-6


@pytest.mark.parametrize("v0,expect", [(0.0, 6.0), (2.0, 4.0), (-2.0, 4.0), (4.0, 0.0), (5.0, 0.0)])
def test_actuator_envelope_in_effort_mode(v0, expect):
    """Torque constant 2 N m/A x 5 A = stall 10 N m, efficiency 0.8, no-load
    4 rad/s, effort limit 6 N m; a command of 100 N m becomes
    min(6, 8 (1 - |v| / 4)) at the control step's start velocity and is held."""
    ix, h = 0.05, 0.005
    prof = mu.actuator(0, torque_constant=2.0, current_limit=5.0, no_load_speed=4.0, efficiency=0.8)
    w = one_joint(flags=ACTUATED | EFFORT_LIMIT, effort_limit=6.0, actuators=[prof], ix=ix)
    q, v = run_joint(w, 0.0, v0, 100.0, implicit=False)
    eq, ev = integrate(0.0, v0, h, 4, lambda q_, v_: expect, ix)
    assert abs(v - ev) <= 2e-6 * max(1.0, abs(ev)) and abs(q - eq) <= 2e-6
    # inactive profile: the effort limit alone; a DoF without one receives no effort
    w = one_joint(flags=ACTUATED | EFFORT_LIMIT, effort_limit=6.0, ix=ix)
    _, v = run_joint(w, 0.0, v0, 100.0, implicit=False)
    assert abs(v - (v0 + 4 * h * 6.0 / ix)) <= 2e-6 * max(1.0, abs(v))
    w = one_joint(flags=ACTUATED, ix=ix)
    _, v = run_joint(w, 0.0, v0, 100.0, implicit=False)
    assert v == np.float32(v0)


def test_actuator_envelope_refreshes_for_implicit_drives():
    """A stiff drive saturates the envelope; for an implicit drive it is
    re-evaluated from the current velocity every microstep."""
    kp, kd, ix, h = 5000.0, 0.0, 0.05, 0.005
    prof = mu.actuator(0, torque_constant=2.0, current_limit=5.0, no_load_speed=4.0, efficiency=0.8)
    w = one_joint(kp=kp, kd=kd, effort_limit=6.0, actuators=[prof], ix=ix)
    inertia = ix + h * kd + h * h * kp

    def tau(q_, v_):
        pd = min(max(kp * (1.0 - q_ - h * v_) - kd * v_, -6.0), 6.0)
        env = min(6.0, 8.0 * (1.0 - min(abs(v_) / 4.0, 1.0)))
        return min(max(pd, -env), env)

    q, v = run_joint(w, 0.0, 1.0, 1.0, implicit=True)
    eq, ev = integrate(0.0, 1.0, h, 4, tau, inertia)
    assert abs(v - ev) <= 2e-6 * max(1.0, abs(ev)) and abs(q - eq) <= 2e-6
    # a held envelope (the start velocity's) gives a different answer
    held = lambda q_, v_: min(max(min(max(kp * (1.0 - q_ - h * v_), -6.0), 6.0), -6.0), 8.0 * (1 - 0.25))
    _, hv = integrate(0.0, 1.0, h, 4, held, inertia)
    assert abs(hv - ev) > 1e-3


def test_actuator_profiles_are_validated():
    bad = mu.actuator(0, torque_constant=2.0, current_limit=5.0, no_load_speed=4.0, efficiency=1.5)
    with pytest.raises(mr.MetalRoboError, match="efficiency"):
        one_joint(flags=ACTUATED | EFFORT_LIMIT, effort_limit=6.0, actuators=[bad])
    stall = mu.actuator(0, torque_constant=2.0, current_limit=5.0, no_load_speed=4.0, efficiency=0.5)
    stall["transmission_and_envelope"][2] = 12.0
    with pytest.raises(mr.MetalRoboError, match="stall torque"):
        one_joint(flags=ACTUATED | EFFORT_LIMIT, effort_limit=6.0, actuators=[stall])
    no_limit = mu.actuator(0, torque_constant=2.0, current_limit=5.0, no_load_speed=4.0, efficiency=0.5)
    with pytest.raises(mr.MetalRoboError, match="effort-limited"):
        one_joint(flags=ACTUATED, actuators=[no_limit])


# ------------------------------------------------------- floating root


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def test_floating_root_follows_the_so3_exponential():
    """A torque-free isotropic body: world-frame omega is constant, so the
    orientation is exp(omega t) q0 (left multiplication) and the COM moves
    at constant velocity, whatever the step size."""
    body = mu.body(1.5, [0.1, 0.1, 0.1])
    dofs = [mu.dof(0, i if i < 3 else mu.INVALID, flags=ROOT) for i in range(6)]
    w = mu.CustomWorld([body], dofs, [], link_count=1, root_type=1, nq=7, nv=6, gravity=(0, 0, 0))
    cfg = w.config(apply_body_damping=0)
    st = ol.OracleState(w.model, cfg)
    rng = np.random.default_rng(3)
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    p0 = np.array([0.2, -0.1, 1.0])
    lin = np.array([0.3, -0.2, 0.1])
    omega = np.array([1.2, -0.7, 2.0])
    st.q[0, :3] = p0
    st.q[0, 3:7] = q0
    st.v[0, :3] = lin
    st.v[0, 3:6] = omega
    steps = 25
    for k in range(steps):
        status, *_ = st.step(np.zeros((1, 6), np.float32), control_step=k)
        assert status["code"][0] == 0
    t = steps * cfg.control_timestep
    theta = omega * t
    angle = np.linalg.norm(theta)
    expq = np.concatenate([np.sin(angle / 2) * theta / angle, [np.cos(angle / 2)]])
    expect = qmul(expq, q0)
    got = st.q[0, 3:7].astype(np.float64)
    if np.dot(got, expect) < 0:
        got = -got
    assert np.max(np.abs(got - expect)) < 2e-5, (got, expect)
    np.testing.assert_allclose(st.q[0, :3], p0 + lin * t, atol=1e-5)  # 100 float32 additions near 1 m
    np.testing.assert_allclose(st.v[0, 3:6], omega, atol=1e-5)
    # body-frame (right) multiplication would be far off for this q0 and omega
    assert np.max(np.abs(qmul(q0, expq) - expect)) > 0.1


# --------------------------------------------------------- heightfield


def patch_world(queries):
    """A 3 x 3-cell non-planar heightfield (4 x 4 vertices) and free bodies."""
    rng = np.random.default_rng(7)
    heights = rng.uniform(-0.03, 0.03, (4, 4)).astype(np.float32)
    bodies = [mu.body(1.0, [0.01] * 3)]
    shapes = [mu.shape(len(queries) + 1, HEIGHTFIELD, [0, 0, 0])]
    for i, (kind, dims) in enumerate(queries):
        bodies.append(mu.body(1.0, [0.001] * 3))
        shapes.append(mu.shape(1 + i, kind, dims))
    bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, heights=heights,
                       hf_origin=(-0.15, -0.15), hf_spacing=0.1)
    return w, heights.astype(np.float64)


def brute_force_triangle(heights, x, y, origin=(-0.15, -0.15), s=0.1):
    """The triangle under (x, y), found by a barycentric test over every
    triangle of the patch: cell (ix, iy) splits into (v00, v10, v11) and
    (v00, v11, v01). Returns the upward unit normal, the vertex v00 and the
    feature index 2 cell + tri."""
    ny, nx = heights.shape
    hits = []
    for iy in range(ny - 1):
        for ix in range(nx - 1):
            def v(dx, dy):
                return np.array([origin[0] + (ix + dx) * s, origin[1] + (iy + dy) * s, heights[iy + dy, ix + dx]])
            for tri, (a, b, c) in enumerate([(v(0, 0), v(1, 0), v(1, 1)), (v(0, 0), v(1, 1), v(0, 1))]):
                m = np.array([[b[0] - a[0], c[0] - a[0]], [b[1] - a[1], c[1] - a[1]]])
                l1, l2 = np.linalg.solve(m, [x - a[0], y - a[1]])
                if l1 >= -1e-12 and l2 >= -1e-12 and l1 + l2 <= 1 + 1e-12:
                    n = np.cross(b - a, c - a)
                    hits.append((n / np.linalg.norm(n), a, 2 * (iy * (nx - 1) + ix) + tri))
    assert len(hits) == 1, "query on a triangle edge"
    return hits[0]


def test_heightfield_contacts_match_brute_force_triangles():
    rng = np.random.default_rng(11)
    xy = list(rng.uniform(-0.14, 0.14, (10, 2)))
    xy += [(0.02, 0.031), (0.031, 0.02), (-0.049, 0.06), (0.051, -0.098)]  # beside diagonals and cell edges
    radius = 0.05
    w, heights = patch_world([(SPHERE, [radius])] * len(xy))
    cfg = w.config()
    q, v, scene = w.model.initial_state()
    expected = []
    for i, (x, y) in enumerate(xy):
        n, a, feat = brute_force_triangle(heights, x, y)
        z = a[2] + (-(n[0] * (x - a[0]) + n[1] * (y - a[1]))) / n[2] + radius / n[2] - 0.004
        scene[0, i] = mu.scene_state((x, y, z), inv_mass=1.0)
        x32 = scene[0, i]["position"][:3].astype(np.float64)
        expected.append((n, float(np.dot(n, x32 - a) - radius), feat))
    scene[0, len(xy)] = mu.scene_state((0, 0, 0), motion=0)
    ov, cnt, reals, feats = ol.collide(w.model, cfg, 0, q[0], v[0], scene[0])
    for i, (n, d, feat) in enumerate(expected):
        k = i  # pair (heightfield 0, sphere i + 1)
        assert cnt[k] == 1
        # A is the heightfield (collider 0): the witness normal points from the surface to the sphere
        np.testing.assert_allclose(reals[k, 0, :3], n, atol=2e-6)
        assert abs(reals[k, 0, 3] - d) < 2e-6
        assert feats[k, 0, 0] == (HEIGHTFIELD << 28) | feat, (i, hex(int(feats[k, 0, 0])), feat)
        assert feats[k, 0, 1] == (SPHERE << 28)


def test_heightfield_box_vertices_match_brute_force_triangles():
    """A tilted box: each of its eight vertices queries the triangle beneath it;
    a witness exists exactly where the vertex is within the accepted distance."""
    w, heights = patch_world([(BOX, [0.04, 0.05, 0.03])])
    cfg = w.config()
    q, v, scene = w.model.initial_state()
    quat = np.array([0.1, -0.05, 0.3, 1.0])
    quat /= np.linalg.norm(quat)
    scene[0, 0] = mu.scene_state((0.013, -0.027, 0.02), rot=tuple(quat.astype(np.float32)), inv_mass=1.0)
    scene[0, 1] = mu.scene_state((0, 0, 0), motion=0)
    ov, cnt, reals, feats = ol.collide(w.model, cfg, 0, q[0], v[0], scene[0])
    qq = scene[0, 0]["orientation"].astype(np.float64)
    R = np.array([[1 - 2 * (qq[1] ** 2 + qq[2] ** 2), 2 * (qq[0] * qq[1] - qq[2] * qq[3]), 2 * (qq[0] * qq[2] + qq[1] * qq[3])],
                  [2 * (qq[0] * qq[1] + qq[2] * qq[3]), 1 - 2 * (qq[0] ** 2 + qq[2] ** 2), 2 * (qq[1] * qq[2] - qq[0] * qq[3])],
                  [2 * (qq[0] * qq[2] - qq[1] * qq[3]), 2 * (qq[1] * qq[2] + qq[0] * qq[3]), 1 - 2 * (qq[0] ** 2 + qq[1] ** 2)]])
    c = scene[0, 0]["position"][:3].astype(np.float64)
    half = np.array([0.04, 0.05, 0.03])
    got = {int(feats[0, j, 1]) & 0xFFFFFFF: j for j in range(cnt[0])}
    accepted = 0.02  # both shapes' contact offsets
    emitted = 0
    for vi in range(8):
        local = half * np.array([1 if vi & 1 else -1, 1 if vi & 2 else -1, 1 if vi & 4 else -1])
        p = c + R @ local
        n, a, feat = brute_force_triangle(heights, p[0], p[1])
        d = float(np.dot(n, p - a))
        if d <= accepted - 1e-5:
            j = got[vi]
            emitted += 1
            assert abs(reals[0, j, 3] - d) < 3e-6
            np.testing.assert_allclose(reals[0, j, :3], n, atol=2e-6)
            assert feats[0, j, 0] == (HEIGHTFIELD << 28) | feat
        elif d > accepted + 1e-5:
            assert vi not in got
    assert emitted >= 3


# -------------------------------------------------------- IR evaluation


def box_on_plane(material=((0.8, 0.8, 0.8, 0.8), (0.0, 0.5, 0.0, 0.0))):
    bodies = [mu.body(1.0, [0.01] * 3), mu.body(1.0, [0.00667] * 3), mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, BOX, [0.1, 0.1, 0.1], contact_offset=0.01),
              mu.shape(2, PLANE, [0, 0, 0], rot=mu.Y_TO_Z, contact_offset=0.01)]
    return mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, materials=[material])


def one_microstep(w, z, vz, h=0.005, slop=0.001, maxdep=1.0):
    cfg = w.config(physics_substeps=1, control_timestep=h, contact_slop=slop, max_depenetration_velocity=maxdep,
                   residual_tolerance=1e-10, max_iterations=3000)
    st = ol.OracleState(w.model, cfg)
    st.scene[0, 0] = mu.scene_state((0.0, 0.0, z), lin=(0.0, 0.0, vz), inv_mass=1.0)
    st.scene[0, 1] = mu.scene_state((0, 0, 0), motion=0)
    status, cs, *_ = st.step(np.zeros((1, 0), np.float32))
    assert status["code"][0] == 0 and cs["constraints"][0] == 4
    return float(st.scene[0, 0]["linear_velocity_inv_mass"][2]), st


def stabilisation(sep, h, slop, maxdep, tau_factor=2.0, zeta=1.0):
    tau = tau_factor * h
    err = min(sep + slop, 0.0)
    r = h / tau
    return min(max(-h * err / (tau * tau * (1.0 + 2.0 * zeta * r + r * r)), 0.0), maxdep)


@pytest.mark.parametrize("depth", [0.004, 0.012, 0.2])
def test_stabilisation_target_of_a_penetrating_box(depth):
    """A box resting `depth` into a plane: the converged normal velocity of
    its four contacts is the stabilisation target (clamped at 1 m/s for the
    deep case), from a free velocity of -g h."""
    h, slop = 0.005, 0.001
    w = box_on_plane()
    vz, _ = one_microstep(w, 0.1 - depth, 0.0, h=h, slop=slop)
    expect = stabilisation(-depth, h, slop, 1.0)
    assert abs(vz - expect) <= 2e-5 * max(1.0, expect), (vz, expect)
    if depth < 0.1:  # without the (h / tau)^2 term of the denominator the answer moves by 12 %
        wrong = min(0.5 * h * (depth - slop) / ((2 * h) ** 2), 1.0)
        assert abs(vz - wrong) > 1e-3


def test_restitution_on_a_new_impact():
    """A box arriving at 2 m/s onto the plane, touching: restitution 0.5 above
    a 0.1 m/s threshold sets the target to 0.5 x the incoming speed."""
    h = 0.005
    w = box_on_plane(((0.8, 0.8, 0.8, 0.8), (0.5, 0.1, 0.0, 0.0)))
    vz, _ = one_microstep(w, 0.1, -2.0, h=h)
    incoming = -2.0 - 9.81 * h
    assert abs(vz - 0.5 * -incoming) <= 2e-5, (vz, 0.5 * -incoming)
    # below the threshold there is no bounce: the contact just stops the box
    vz, _ = one_microstep(w, 0.1, -0.05, h=h)
    assert abs(vz) < 2e-5


def test_regularisation_of_a_compliant_contact():
    """Compliance c and dissipation d per material (summed over the pair) give
    every contact row R = c / h^2 + d / h. Four symmetric contacts under a
    resting box then carry Lambda = g h / (1/m + R/4) in total, and the box
    keeps -g h + Lambda / m."""
    h, c, d = 0.005, 0.5e-4, 0.5e-3
    w = box_on_plane(((0.8, 0.8, 0.8, 0.8), (0.0, 0.5, c, d)))
    vz, _ = one_microstep(w, 0.1, 0.0, h=h)
    R = 2 * c / (h * h) + 2 * d / h
    lam = 9.81 * h / (1.0 + R / 4.0)
    expect = -9.81 * h + lam
    assert abs(vz - expect) <= 2e-6, (vz, expect)
