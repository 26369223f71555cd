"""Pins the oracle's temporalCone cone solve and integration by physical
invariants with closed-form answers (CPU only):

  * a box resting on a plane carries its weight: the normal impulses of the
    final sweep sum to m g h per microstep;
  * a box sliding on a plane decelerates at mu g (Coulomb), per tangent axis
    of an elliptic cone (mu_u along the manifold tangent, mu_v along n x t);
  * a free box in flight follows the ballistic closed form;
  * the block projection keeps every impulse inside its cone (violation 0).
The reference's own solver checks are tolerance/invariant based too
(apps/contact_gpu_probe.mm:26-28, reference_conic_probe.cpp:9-100).
"""

import numpy as np
import pytest

import model_util as mu
import oracle_lib as ol

BOX, PLANE = 2, 3


def box_on_plane(mu_u=0.8, mu_v=0.8, half=0.1, mass=1.0, z=None, lin=(0, 0, 0), envs=1):
    he = np.broadcast_to(np.asarray(half, np.float64), (3,))
    a, b, c = 2 * he
    inertia = [mass * (b * b + c * c) / 12.0, mass * (a * a + c * c) / 12.0, mass * (a * a + b * b) / 12.0]
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(mass, inertia), mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, BOX, list(he)), mu.shape(2, PLANE, [0, 0, 0], rot=mu.Y_TO_Z)]
    mats = [((mu_u, mu_v, mu_u, mu_v), (0.0, 0.5, 0.0, 0.0))]
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, materials=mats, envs=envs)
    cfg = w.config()
    st = ol.OracleState(w.model, cfg)
    zz = float(he[2]) if z is None else z
    for e in range(envs):
        st.scene[e, 0] = mu.scene_state((0, 0, zz), lin=lin, inv_mass=1.0 / mass)
        st.scene[e, 1] = mu.scene_state((0, 0, 0), motion=0)
    return w, cfg, st


def run(st, steps, nv=0):
    cmd = np.zeros((st.q.shape[0], max(nv, 0)), np.float32)
    out = []
    for k in range(steps):
        status, cs, acc, pairs, counts = st.step(cmd, control_step=k)
        assert np.all(status["code"] == 0), status
        out.append(cs)
    return out


def test_resting_box_normal_impulse_equals_weight():
    w, cfg, st = box_on_plane()
    run(st, 50)
    h = cfg.control_timestep / cfg.physics_substeps
    n = int(st.mcount[0])
    assert n == 1
    lam = st.mp[0, 0]["impulses"][:4, 0].sum()
    assert abs(lam - 1.0 * 9.81 * h) / (9.81 * h) < 0.02, lam
    z = st.scene[0, 0]["position"][2]
    assert 0.095 < z < 0.1005, z
    assert np.linalg.norm(st.scene[0, 0]["linear_velocity_inv_mass"][:3]) < 1e-3


def sliding_sphere(mu_u, mu_v, lin):
    """A sphere with a very large rotational inertia on a plane: one contact
    whose block is isotropic (normal along the lever arm, rotation frozen),
    so one block step with the radial projection is the exact Coulomb step."""
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(1.0, [1.0e4] * 3), mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, 0, [0.1]), mu.shape(2, PLANE, [0, 0, 0], rot=mu.Y_TO_Z)]
    mats = [((mu_u, mu_v, mu_u, mu_v), (0.0, 0.5, 0.0, 0.0))]
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, materials=mats)
    cfg = w.config()
    st = ol.OracleState(w.model, cfg)
    st.scene[0, 0] = mu.scene_state((0, 0, 0.1), lin=lin, inv_mass=1.0)
    st.scene[0, 1] = mu.scene_state((0, 0, 0), motion=0)
    return w, cfg, st


@pytest.mark.parametrize("direction,mu_axis", [((1, 0, 0), 0.7), ((0, 1, 0), 1.0)])
def test_sliding_decelerates_at_elliptic_coefficient(direction, mu_axis):
    """Coulomb sliding: |dv/dt| = mu g along each cone axis. The contact
    frame is u = stable tangent of n = (0, 0, -1) = +y (mu_u) and v = n x u
    = +x (mu_v)."""
    v0 = 2.0
    w, cfg, st = sliding_sphere(1.0, 0.7, tuple(v0 * np.array(direction, float)))
    steps = 5
    run(st, steps)
    t = steps * cfg.control_timestep
    v = st.scene[0, 0]["linear_velocity_inv_mass"][:3] @ np.array(direction, np.float64)
    expect = v0 - mu_axis * 9.81 * t
    assert abs(v - expect) < 0.01 * v0, (v, expect)
    u = st.mh[0, 0]["tangent_and_metric"][:3]
    np.testing.assert_allclose(np.abs(u), [0, 1, 0], atol=1e-5)


def test_ballistic_flight_closed_form():
    w, cfg, st = box_on_plane(z=2.0, lin=(0.5, -0.25, 1.0))
    steps = 10
    run(st, steps)
    h = cfg.control_timestep / cfg.physics_substeps
    n = steps * cfg.physics_substeps
    vz = 1.0 - 9.81 * h * n
    # symplectic Euler: z_n = z0 + h * sum_{k=1..n} (vz0 - g h k)
    z = 2.0 + h * (n * 1.0 - 9.81 * h * n * (n + 1) / 2)
    s = st.scene[0, 0]
    np.testing.assert_allclose(s["linear_velocity_inv_mass"][:3], [0.5, -0.25, vz], atol=1e-5)
    np.testing.assert_allclose(s["position"][:3], [0.5 * h * n, -0.25 * h * n, z], atol=1e-5)


def test_impulses_stay_inside_the_cone():
    w, cfg, st = box_on_plane(mu_u=1.0, mu_v=0.7, half=(0.2, 0.2, 0.005), lin=(1.5, 1.0, 0.0), z=0.005)
    for cs in run(st, 10):
        assert np.all(cs["residuals"][:, 2] <= 1e-6)
    imp = st.mp[0, 0]["impulses"][:4]
    lam_n = imp[:, 0]
    assert np.all(lam_n >= 0.0)
    ell = np.sqrt((imp[:, 1] / np.maximum(1.0 * lam_n, 1e-30)) ** 2 + (imp[:, 2] / np.maximum(0.7 * lam_n, 1e-30)) ** 2)
    assert np.all(ell[lam_n > 0] <= 1.0 + 1e-5)


def test_adaptive_mode_reaches_the_tolerance_or_the_cap():
    """With a residual tolerance the sweeps stop at the first sweep whose max
    |dlambda| is at or below it (MetalWorldContact.metal:8040-8218 iteration
    control), never later than max_iterations."""
    w, cfg, st = box_on_plane()
    cfg.residual_tolerance = 1.0e-7
    cfg.max_iterations = 64
    for cs in run(st, 20):
        it = int(cs["solver_iterations"][0])
        res = float(cs["residuals"][0, 0])
        assert 1 <= it <= 64
        assert res <= 1.0e-7 or it == 64, (it, res)


def test_maximum_normal_impulse_caps_every_contact():
    """The cone's maximum normal impulse (MetalWorldContact.metal:488-493): with the cap below
    the share of the weight each of the four corner contacts carries, every stored normal
    impulse equals the cap and the box sinks at g - 4 cap / (m h) per microstep; with the cap
    above it, the resting box is unchanged bit for bit."""
    w, cfg, st = box_on_plane()
    run(st, 40)
    h = cfg.control_timestep / cfg.physics_substeps
    share = 9.81 * h / 4.0
    snap = [a.copy() for a in (st.q, st.v, st.scene, st.mh, st.mp, st.mcount)]
    free = ol.OracleState(w.model, cfg)
    free.q, free.v, free.scene, free.mh, free.mp, free.mcount = [a.copy() for a in snap]
    run(free, 1)
    cfg_hi = w.config()
    cfg_hi.maximum_normal_impulse = 10.0 * share
    hi = ol.OracleState(w.model, cfg_hi)
    hi.q, hi.v, hi.scene, hi.mh, hi.mp, hi.mcount = [a.copy() for a in snap]
    run(hi, 1)
    assert np.array_equal(hi.scene.view(np.uint8), free.scene.view(np.uint8))
    cap = 0.25 * share
    cfg_lo = w.config()
    cfg_lo.maximum_normal_impulse = cap
    lo = ol.OracleState(w.model, cfg_lo)
    lo.q, lo.v, lo.scene, lo.mh, lo.mp, lo.mcount = [a.copy() for a in snap]
    vz0 = float(lo.scene[0, 0]["linear_velocity_inv_mass"][2])
    cs = run(lo, 1)[-1]
    lam = lo.mp[0, 0]["impulses"][:4, 0]
    np.testing.assert_array_equal(lam, np.float32(cap))
    assert cs["residuals"][0, 1] > 0.0  # the capped normal rows are no longer satisfied
    vz = float(lo.scene[0, 0]["linear_velocity_inv_mass"][2])
    steps = cfg.physics_substeps
    expect = vz0 + steps * (-9.81 * h + 4.0 * cap)  # unit mass
    assert abs(vz - expect) < 1e-4, (vz, expect)
