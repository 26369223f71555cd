"""The oracle against the reference's own FP64 CPU oracle, compiled from the
reference sources (oracle/build_ref.py -> oracle/_ref/libmr_ref.so) and
called through oracle/ref_harness.cpp. CPU only; skipped where the library
was not built (no reference sources at build time).

  * Articulated dynamics: the oracle's FP32 ABA against
    computeArticulatedForwardDynamics (ArticulatedDynamics.cpp:2241, FP64
    CRBA + RNEA) on the chain, humanoid and random-tree models at random
    states, within the reference ABA probe's 5e-5 scaled-acceleration bound
    (articulated_aba_gpu_probe.mm:811-813).
  * Collision: for every environment of the three workloads at states along
    an oracle trajectory, the oracle's per-pair overlap flags, witness counts,
    feature ids and witnesses against collideCpuReference
    (Collision.cpp:2959-3228) given the same colliders, body poses and the
    complement of the compiled pair table as exclusions. Heightfield pairs
    are left out (the reference CPU oracle has no heightfield). Two FP32
    effects are allowed, each checked for what it must preserve:
      - at an exact tie between two box features (a capsule closest to a
        corner shared by two edges) FP32 rounding may pick the other label;
        the witness geometry must still agree;
      - a box-box edge-edge witness whose support edges tie (parallel edges
        of one box equally far along the normal) may sit on the other edge;
        normal and separation must agree and the points must satisfy the
        witness relation (p_B - p_A) . n = separation.
  * Exact-cone solve: oracle/exact_cone.cpp against solveReferenceConicProblem
    (ReferenceConicSolver.cpp:182) on random circular-cone problems, and
    against the reference's projected Gauss-Seidel solveContactConstraints
    (ConstraintSolver.cpp:1047) on free-body scenes whose contacts stick.
"""

import os

import numpy as np
import pytest

import metalrobo as mr
import oracle_lib as ol
import ref_lib as rl
from fp64_dynamics import TreeModel
from metalrobo.native import BODY_STATE_DTYPE
from test_oracle_dynamics import random_state

pytestmark = pytest.mark.skipif(not rl.available(), reason="oracle/_ref not built (reference sources absent)")

HEIGHTFIELD = 7
BOX = 2
MOTION_DYNAMIC = 2
KINDS = [mr.MR_SYNTHETIC_CHAIN_BOXES, mr.MR_SYNTHETIC_HUMANOID_TERRAIN, mr.MR_SYNTHETIC_RANDOM_TREES]


@pytest.mark.parametrize("kind", KINDS)
def test_aba_matches_reference_fp64_dynamics(kind):
    model = mr.WorldModel.synthetic(kind, 6 if kind != mr.MR_SYNTHETIC_HUMANOID_TERRAIN else 2, 17)
    cfg = model.default_config()
    cfg.actuation_mode = mr.native.MR_ACTUATION_EFFORT
    headers, env_model = model.models(), model.environment_model()
    bodies, dofs = model.bodies(), model.dofs()
    q0, _, _ = model.initial_state()
    h_step = cfg.control_timestep / cfg.physics_substeps
    rng = np.random.default_rng(1)
    worst = 0.0
    for env in range(model.environment_count):
        h = headers[env_model[env]]
        tm = TreeModel(h, bodies, dofs, [0, 0, -9.81])
        for _ in range(4):
            q, v = random_state(tm, rng, q0[env][:tm.nq])
            tau = rng.normal(0, 5.0, tm.nv).astype(np.float32)
            code, acc = ol.forward_dynamics(model, cfg, env, q, v, tau)
            assert code == 0
            rc, ref, err = rl.forward_dynamics(h, bodies, dofs, q, v, tau, timestep=h_step)
            assert rc == 0, err
            scaled = np.max(np.abs(acc[:tm.nv] - ref)) / max(1.0, np.max(np.abs(ref)))
            worst = max(worst, scaled)
            assert scaled < 5.0e-5, (env, scaled)
    print(f"worst scaled acceleration error against the reference {worst:.2e}")


def env_scene(model, cfg, env, q, v, scene):
    """Body states (links from the oracle's kinematics, then scene bodies), colliders and
    compiled pairs of one environment."""
    h = model.models()[model.environment_model()[env]]
    nl, ns = int(h["link_count"]), int(h["scene_body_count"])
    code, poses = ol.link_poses(model, cfg, env, q, v)
    assert code == 0
    bs = np.zeros(nl + ns, BODY_STATE_DTYPE)
    for i in range(nl):
        bs[i]["position"] = [*poses[i, :3], 1.0]
        bs[i]["orientation"] = poses[i, 3:]
        bs[i]["flags_indices"] = [MOTION_DYNAMIC, 0, i, 0]
    bs[nl:] = scene[:ns]
    shapes = model.shapes()[int(h["shape_offset"]):int(h["shape_offset"]) + int(h["shape_count"])]
    pairs = model.pairs()[int(h["pair_offset"]):int(h["pair_offset"]) + int(h["pair_count"])]
    return bs, shapes, pairs


def compare_environment(model, cfg, env, q, v, scene, tally):
    bs, shapes, pairs = env_scene(model, cfg, env, q, v, scene)
    keep = np.nonzero(shapes["shape_type"] != HEIGHTFIELD)[0]
    local = {int(s): i for i, s in enumerate(keep)}
    compiled = {(int(p["collider_a"]), int(p["collider_b"])): k for k, p in enumerate(pairs)}
    exclusions = [(local[int(a)], local[int(b)]) for i, a in enumerate(keep) for b in keep[i + 1:]
                  if (int(a), int(b)) not in compiled]
    rc, rpairs, raw, raw_u, err = rl.collide(shapes[keep], bs, exclusions)
    assert rc == 0, err
    overlap, counts, reals, feats = ol.collide(model, cfg, env, q, v, scene)

    ref_pairs = {(int(keep[a]), int(keep[b])) for a, b in rpairs}
    ref_contacts = {}
    for c in range(len(raw)):
        a, b = rpairs[raw_u[c, 2]]
        ref_contacts.setdefault((int(keep[a]), int(keep[b])), []).append(c)
    for (a, b), k in compiled.items():
        if HEIGHTFIELD in (int(shapes[a]["shape_type"]), int(shapes[b]["shape_type"])):
            continue
        tally["pairs"] += 1
        # broadphase: the padded-AABB candidate set is the reference's
        assert bool(overlap[k]) == ((a, b) in ref_pairs), (env, a, b)
        cs = ref_contacts.get((a, b), [])
        # narrowphase: the same number of witnesses per pair
        assert int(counts[k]) == len(cs), (env, a, b, int(counts[k]), len(cs))
        box_box = int(shapes[a]["shape_type"]) == BOX and int(shapes[b]["shape_type"]) == BOX
        for c in cs:
            tally["contacts"] += 1
            fa, fb = int(raw_u[c, 0]), int(raw_u[c, 1])
            geo = np.abs(reals[k, :counts[k]] - raw[c][None, :]).max(axis=1)
            same = [j for j in range(counts[k]) if feats[k, j, 0] == fa and feats[k, j, 1] == fb]
            if not same:
                # a tie between two features: same witness geometry under the other label
                tally["relabelled"] += 1
                assert geo.min() < 5.0e-4, (env, a, b, hex(fa), hex(fb), geo.min())
                continue
            j = min(same, key=lambda x: geo[x])
            # normal (unit, A->B) and separation
            assert np.abs(reals[k, j, :4] - raw[c, :4]).max() < 1.0e-5, (env, a, b, reals[k, j], raw[c])
            if geo[j] < 1.0e-4:
                continue
            # only a box-box edge-edge witness may sit on another support edge
            assert box_box and fa == fb and fa & 0x0FFFFFFF >= 256 + 6, (env, a, b, hex(fa), geo[j])
            tally["edge_ties"] += 1
            n, sep = reals[k, j, :3], reals[k, j, 3]
            gap = float(np.dot(reals[k, j, 7:10] - reals[k, j, 4:7], n))
            assert abs(gap - sep) < 1.0e-5, (env, a, b, gap, sep)


@pytest.mark.parametrize("kind,envs,steps", [(mr.MR_SYNTHETIC_CHAIN_BOXES, 8, 60),
                                             (mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 4, 60),
                                             (mr.MR_SYNTHETIC_RANDOM_TREES, 6, 40)])
def test_collision_matches_reference_fp64_collision(kind, envs, steps):
    model = mr.WorldModel.synthetic(kind, envs, 5)
    cfg = model.default_config()
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, steps)
    tally = {"pairs": 0, "contacts": 0, "relabelled": 0, "edge_ties": 0}
    for t in range(steps):
        if t % 5 == 0:
            for env in range(envs):
                compare_environment(model, cfg, env, st.q[env], st.v[env], st.scene[env], tally)
        st.step(cmds[t], control_step=t)
    assert tally["pairs"] > 1000
    if kind != mr.MR_SYNTHETIC_HUMANOID_TERRAIN:
        assert tally["contacts"] > 500
    # the tie cases stay rare
    assert tally["relabelled"] + tally["edge_ties"] <= 0.03 * max(1, tally["contacts"]), tally
    print(tally)


def random_conic_problem(rng, nv, k):
    A = rng.normal(size=(nv, nv))
    M = A @ A.T + nv * np.eye(nv)
    Minv = np.linalg.inv(M)
    J = rng.normal(size=(k, 3, nv))
    vf = rng.normal(size=nv)
    target = np.zeros((k, 3))
    target[:, 0] = rng.uniform(0.0, 0.5, k)
    reg = np.full((k, 3), 1.0e-4)
    mu = rng.uniform(0.3, 1.2, k)
    return Minv, vf, J, target, reg, mu


@pytest.mark.parametrize("seed", range(6))
def test_exact_cone_matches_reference_conic_solver(seed):
    rng = np.random.default_rng(100 + seed)
    nv, k = int(rng.integers(4, 12)), int(rng.integers(1, 5))
    Minv, vf, J, target, reg, mu = random_conic_problem(rng, nv, k)
    rc, vel_ref, lam_ref, info, err = rl.conic_solve(Minv, vf, J, target, reg, mu)
    assert rc == 0, err
    # the same optimality tolerance as the reference solve; ill-conditioned draws (Delassus
    # condition ~1e4 with 1e-4 regularization) take ~1e5 projected-gradient iterations in both
    code, lam, vel, stats = ol.exact_cone(Minv, vf, J.reshape(k * 3, nv), target.reshape(-1), reg.reshape(-1),
                                         np.stack([mu, mu], 1), max_iterations=2000000, tolerance=1.0e-10)
    assert code == 0
    # the optimum of the strictly convex program is unique: equal objectives, and impulses
    # equal to what the conditioning allows at that tolerance
    assert abs(stats["objective"] - info[1]) <= 1.0e-9 * max(1.0, abs(info[1]))
    scale = max(1.0, np.abs(lam_ref).max())
    np.testing.assert_allclose(lam.reshape(k, 3), lam_ref, atol=1.0e-5 * scale)
    np.testing.assert_allclose(vel, vel_ref, atol=1.0e-5 * max(1.0, np.abs(vel_ref).max()))


# ---- the reference's projected Gauss-Seidel (ConstraintSolver.cpp:1047) ----

def _free_body(position, mass, inertia_diag, rotation, velocity, spin):
    b = np.zeros((), rl.PGS_BODY_DTYPE)
    b["position"][:3] = position
    b["orientation"][:] = (0.0, 0.0, 0.0, 1.0)
    b["linear_velocity_inverse_mass"][:] = (*velocity, 1.0 / mass if mass > 0 else 0.0)
    b["angular_velocity"][:3] = spin
    if mass > 0:
        inv = rotation @ np.diag(1.0 / np.asarray(inertia_diag)) @ rotation.T
        b["inverse_inertia_world"][:, :3] = inv
    b["flags_and_indices"][0] = MOTION_DYNAMIC if mass > 0 else 0
    return b


def _contact(a, b, feature, point, separation, normal, tangent, mu, compliance):
    c = np.zeros((), rl.PGS_CONTACT_DTYPE)
    c["body_a"], c["body_b"] = a, b
    c["pair_key"] = (a << 32) | b
    c["feature_key"] = feature
    c["point_and_separation"][:] = (*point, separation)
    c["normal"][:3] = normal
    c["tangent"][:3] = tangent
    c["friction"][:2] = (mu, mu)
    c["response"][2] = compliance
    return c


def _pgs_scene(name, h):
    """Free bodies resting or sticking on the static ground (body 0) and on
    each other, with contacts sorted by (pair, feature) as the deterministic
    solve requires. Velocities already include one step of gravity."""
    g = -9.81 * h
    up = np.array([0.0, 0.0, 1.0])
    ground = _free_body((0, 0, 0), 0.0, (1, 1, 1), np.eye(3), (0, 0, 0), (0, 0, 0))
    bodies, contacts = [ground], []
    if name == "sphere_sliding_to_stick":
        r = 0.1
        bodies.append(_free_body((0, 0, r - 0.002), 1.5, (0.006,) * 3, np.eye(3),
                                 (0.02, -0.015, g), (0.3, 0.1, 0.05)))
        contacts.append(_contact(0, 1, 0, (0, 0, -0.002), -0.002, up, (1, 0, 0), 1.0, 1.0e-7))
    elif name == "stacked_spheres":
        r = 0.1
        bodies.append(_free_body((0, 0, r - 0.001), 1.0, (0.004,) * 3, np.eye(3), (0.01, 0, g), (0, 0.2, 0)))
        off = np.array([0.01, -0.004, 0.0])
        d = np.array([off[0], off[1], np.sqrt((2 * r) ** 2 - off[0] ** 2 - off[1] ** 2)])
        top = bodies[1]["position"][:3] + d - 0.0015 * d / np.linalg.norm(d)
        bodies.append(_free_body(top, 0.7, (0.0028,) * 3, np.eye(3), (0, 0.01, g), (0, 0, 0.1)))
        n = d / np.linalg.norm(d)
        t = np.cross(n, [0.0, 1.0, 0.0])
        contacts.append(_contact(0, 1, 0, (0, 0, -0.001), -0.001, up, (1, 0, 0), 1.2, 0.0))
        contacts.append(_contact(1, 2, 0, bodies[1]["position"][:3] + r * n, -0.0015, n, t / np.linalg.norm(t),
                                 1.2, 1.0e-8))
    elif name == "box_on_an_edge":
        hx = 0.15
        # two corners of one bottom edge touch the ground, the box leans 20 degrees about x
        lean = np.deg2rad(20.0)
        Rx = np.array([[1, 0, 0], [0, np.cos(lean), -np.sin(lean)], [0, np.sin(lean), np.cos(lean)]])
        com = np.array([0.0, 0.1 * np.sin(lean) * 0.5, 0.08])
        bodies.append(_free_body(com, 2.0, (0.02, 0.015, 0.01), Rx, (0.005, -0.01, g), (0.05, 0.0, 0.02)))
        for k, x in enumerate((-hx, hx)):
            contacts.append(_contact(0, 1, k, (x, 0.0, -0.0005), -0.0005, up, (1, 0, 0), 1.0, 1.0e-8))
    else:  # sphere pyramid: two spheres on the ground, a third resting on both
        r = 0.1
        gap = 0.19
        for k, x in enumerate((-gap / 2 - 0.005, gap / 2 + 0.005)):
            bodies.append(_free_body((x, 0, r - 0.001), 1.0, (0.004,) * 3, np.eye(3), (0, 0, g), (0, 0, 0)))
        x1, x2 = bodies[1]["position"][:3], bodies[2]["position"][:3]
        dx = (x2[0] - x1[0]) / 2
        top = np.array([0.0, 0.0, x1[2] + np.sqrt((2 * r) ** 2 - dx ** 2) - 0.001])
        bodies.append(_free_body(top, 0.8, (0.0032,) * 3, np.eye(3), (0, 0, g), (0, 0, 0)))
        contacts.append(_contact(0, 1, 0, (x1[0], 0, -0.001), -0.001, up, (1, 0, 0), 1.5, 0.0))
        contacts.append(_contact(0, 2, 0, (x2[0], 0, -0.001), -0.001, up, (1, 0, 0), 1.5, 0.0))
        for a, xa in ((1, x1), (2, x2)):
            n = top - xa
            n /= np.linalg.norm(n)
            contacts.append(_contact(a, 3, 0, xa + r * n, -0.001, n, (0, 1, 0), 1.5, 1.0e-8))
    return np.array(bodies, rl.PGS_BODY_DTYPE), np.array(contacts, rl.PGS_CONTACT_DTYPE)


def _cone_problem(bodies, contacts, h, erp, slop, max_depenetration):
    """The same contacts as one exact-cone program over the dynamic bodies'
    world velocities [v, w]: rows (normal, u, v) per contact with the
    reference's basis (ConstraintSolver.cpp:242-251), the normal target of
    normalTargetVelocity (:380-405) and the normal softness compliance / h^2
    (:373-378); tangents unregularised."""
    dyn = [i for i in range(len(bodies)) if bodies[i]["flags_and_indices"][0] == MOTION_DYNAMIC]
    col = {b: 6 * k for k, b in enumerate(dyn)}
    nv = 6 * len(dyn)
    Minv = np.zeros((nv, nv))
    vf = np.zeros(nv)
    for b in dyn:
        o = col[b]
        Minv[o:o + 3, o:o + 3] = np.eye(3) * float(bodies[b]["linear_velocity_inverse_mass"][3])
        Minv[o + 3:o + 6, o + 3:o + 6] = bodies[b]["inverse_inertia_world"][:, :3].astype(np.float64)
        vf[o:o + 3] = bodies[b]["linear_velocity_inverse_mass"][:3]
        vf[o + 3:o + 6] = bodies[b]["angular_velocity"][:3]
    k = len(contacts)
    J = np.zeros((k, 3, nv))
    target = np.zeros((k, 3))
    reg = np.zeros((k, 3))
    mu = np.zeros(k)
    for i, c in enumerate(contacts):
        p = c["point_and_separation"][:3].astype(np.float64)
        n = c["normal"][:3].astype(np.float64)
        n /= np.linalg.norm(n)
        t = c["tangent"][:3].astype(np.float64)
        u = t - n * (n @ t)
        u /= np.linalg.norm(u)
        v = np.cross(n, u)
        for sign, b in ((-1.0, int(c["body_a"])), (1.0, int(c["body_b"]))):
            if b not in col:
                continue
            r = p - bodies[b]["position"][:3].astype(np.float64)
            for row, d in enumerate((n, u, v)):
                J[i, row, col[b]:col[b] + 3] += sign * d
                J[i, row, col[b] + 3:col[b] + 6] += sign * np.cross(r, d)
        penetration = min(float(c["point_and_separation"][3]) + slop, 0.0)
        target[i, 0] = max(min(max_depenetration, -erp * penetration / h), 0.0)
        reg[i, 0] = float(c["response"][2]) / (h * h)
        mu[i] = float(c["friction"][0])
    return Minv, vf, J, target, reg, mu, col


@pytest.mark.parametrize("name", ["sphere_sliding_to_stick", "stacked_spheres", "box_on_an_edge", "sphere_pyramid"])
def test_exact_cone_matches_reference_pgs_on_sticking_contacts(name):
    """The reference's PGS (normal then tangent block per contact, radial
    friction projection) and the FP64 exact cone share their fixed point when
    every contact sticks strictly inside its cone: relative velocity at each
    contact equal to its target, so the post-solve body velocities agree.
    Sliding contacts are not compared: PGS keeps the normal velocity at its
    target while the exact cone's dual adds mu*|v_t| to it."""
    h, erp, slop, vmax = 1.0 / 240.0, 0.2, 1.0e-4, 2.0
    bodies, contacts = _pgs_scene(name, h)
    rc, b_pgs, c_pgs, info, err = rl.pgs_solve(bodies, contacts, h, erp, slop, vmax, 6000)
    assert rc == 0, err
    Minv, vf, J, target, reg, mu, col = _cone_problem(bodies, contacts, h, erp, slop, vmax)
    k, nv = J.shape[0], J.shape[2]
    code, lam, vel, stats = ol.exact_cone(Minv, vf, J.reshape(3 * k, nv), target.reshape(-1), reg.reshape(-1),
                                         np.stack([mu, mu], 1), max_iterations=2000000, tolerance=1.0e-13)
    assert code == 0
    lam = lam.reshape(k, 3)
    # every contact is loaded and sticks strictly inside its cone
    assert (lam[:, 0] > 1.0e-4).all()
    assert (np.hypot(lam[:, 1], lam[:, 2]) < 0.9 * mu * lam[:, 0]).all()
    # the PGS result also sticks: relative velocities at their softened targets
    v_pgs = np.zeros(nv)
    for b, o in col.items():
        v_pgs[o:o + 3] = b_pgs[b]["linear_velocity_inverse_mass"][:3]
        v_pgs[o + 3:o + 6] = b_pgs[b]["angular_velocity"][:3]
    rel = np.einsum("krn,n->kr", J, v_pgs)
    np.testing.assert_allclose(rel, target - reg * c_pgs["impulses"][:, :3], atol=2.0e-5)
    # equal post-solve velocities; a sign or index mistake in the program's rows fails here
    np.testing.assert_allclose(vel, v_pgs, atol=2.0e-5 * max(1.0, np.abs(vf).max()))
    # equal impulses where the rows are independent (the solution is unique)
    if np.linalg.matrix_rank(J.reshape(3 * k, nv)) == 3 * k:
        np.testing.assert_allclose(c_pgs["impulses"][:, :3], lam, atol=2.0e-5 * max(1.0, lam.max()))
