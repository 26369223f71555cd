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
    (ReferenceConicSolver.cpp:182) on random circular-cone problems.
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
