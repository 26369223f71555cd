"""GPU parity: the HIP control-step kernel against the CPU oracle, through the C ABI.

Teacher-forced: before every control step the oracle receives the GPU's
accepted state (q, v, scene bodies, persistent manifolds), both step once
with the same command, and the outputs are compared:
  * broadphase pair lists (environment, collider A, collider B, class) and
    manifold feature ids: bit-exact;
  * status / rollback codes and failing indices: bit-exact;
  * q, v, scene bodies, impulses: within 1e-4 relative (the kernels follow
    the oracle's FP32 evaluation order, so they are expected to be equal).
"""

import numpy as np
import pytest

import metalrobo as mr
import oracle_lib as ol

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / (1.0 + np.abs(b))))


def copy_gpu_state_to_oracle(world, ost):
    q, v, s = world.state()
    mh, mp, mc = world.manifolds()
    ost.q[:] = q
    ost.v[:] = v
    ost.scene[:] = s
    ost.mh[:] = mh
    ost.mp[:] = mp
    ost.mcount[:] = mc


def scene_floats(s):
    return np.concatenate([s["position"][..., :3].reshape(len(s), -1), s["orientation"].reshape(len(s), -1),
                           s["linear_velocity_inv_mass"][..., :3].reshape(len(s), -1),
                           s["angular_velocity"][..., :3].reshape(len(s), -1)], axis=1)


def compare_step(world, ost, k, o_status, o_cs, o_pairs, o_counts, log):
    g_status = world.statuses()[0]
    g_cs = world.contact_statuses()[0]
    g_pairs, g_counts = world.evidence()
    # discrete outputs: bit-exact
    np.testing.assert_array_equal(g_status["code"], o_status["code"], err_msg=f"step {k} status codes")
    np.testing.assert_array_equal(g_status["failing_index"], o_status["failing_index"])
    np.testing.assert_array_equal(g_status["successful_substeps"], o_status["successful_substeps"])
    np.testing.assert_array_equal(g_counts, o_counts, err_msg=f"step {k} pair counts")
    for e in range(len(g_counts)):
        n = min(int(g_counts[e]), g_pairs.shape[1])
        np.testing.assert_array_equal(g_pairs[e, :n], o_pairs[e, :n], err_msg=f"step {k} env {e} pair list")
    for f in ("active_pairs", "raw_contacts", "manifolds", "constraints", "islands", "retained_points",
              "new_points", "solver_iterations", "max_constraints"):
        np.testing.assert_array_equal(g_cs[f], o_cs[f], err_msg=f"step {k} contact status {f}")
    q, v, s = world.state()
    mh, mp, mc = world.manifolds()
    np.testing.assert_array_equal(mc, ost.mcount, err_msg=f"step {k} manifold counts")
    for e in range(len(mc)):
        n = int(mc[e])
        np.testing.assert_array_equal(mh[e, :n]["pair_and_count"], ost.mh[e, :n]["pair_and_count"])
        np.testing.assert_array_equal(mp[e, :n]["feature_and_life"], ost.mp[e, :n]["feature_and_life"])
    errs = {
        "q": rel_err(q, ost.q), "v": rel_err(v, ost.v),
        "scene": rel_err(scene_floats(s.reshape(-1)), scene_floats(ost.scene.reshape(-1))),
        "impulse": max([rel_err(mp[e, :int(mc[e])]["impulses"], ost.mp[e, :int(mc[e])]["impulses"])
                        for e in range(len(mc))] + [0.0]),
        "residual": rel_err(g_cs["residuals"], o_cs["residuals"]),
    }
    for name, err in errs.items():
        assert err <= REL_TOL, f"step {k}: {name} relative error {err:.3e}"
    log.append(errs)
    return errs


SOLVERS = [mr.native.MR_SOLVER_ORDERED_SWEEPS, mr.native.MR_SOLVER_RELAXED_JACOBI]


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("kind,envs,steps,seed", [
    (mr.MR_SYNTHETIC_CHAIN_BOXES, 16, 48, 3),
    (mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 8, 24, 5),
    (mr.MR_SYNTHETIC_RANDOM_TREES, 12, 30, 9),
])
def test_teacher_forced_parity(gpu, kind, envs, steps, seed, solver):
    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = model.default_config()
    cfg.solver = solver
    cfg.capture_evidence = 1
    cfg.capture_substep = 0
    world = mr.World(model, cfg)
    ost = ol.OracleState(model, cfg)
    cmds = model.commands(0, steps)
    log = []
    exact = True
    total_pairs = 0
    total_cons = 0
    for k in range(steps):
        copy_gpu_state_to_oracle(world, ost)
        o_status, o_cs, _, o_pairs, o_counts = ost.step(cmds[k], control_step=k)
        world.advance(cmds[k:k + 1])
        errs = compare_step(world, ost, k, o_status, o_cs, o_pairs, o_counts, log)
        exact = exact and all(v == 0.0 for v in errs.values())
        total_pairs += int(o_counts.sum())
        total_cons += int(o_cs["constraints"].sum())
        assert np.all(o_status["code"] == 0), f"step {k}: oracle failures {o_status['code']}"
    worst = {n: max(d[n] for d in log) for n in log[0]}
    print(f"kind={kind} solver={solver} envs={envs} steps={steps} pairs={total_pairs} constraints={total_cons} "
          f"bit_exact={exact} worst={worst}", flush=True)
    assert total_pairs > 0
    assert total_cons > 0, "the run never reached a contact"


@pytest.mark.parametrize("solver", SOLVERS)
def test_free_running_horizon_matches_oracle(gpu, solver):
    """Whole-horizon run in one advance: the GPU trajectory equals the oracle's."""
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 8, 11)
    cfg = model.default_config()
    cfg.solver = solver
    world = mr.World(model, cfg)
    ost = ol.OracleState(model, cfg)
    steps = 40
    cmds = model.commands(0, steps)
    world.advance(cmds)
    for k in range(steps):
        ost.step(cmds[k], control_step=k)
    q, v, s = world.state()
    st = world.statuses()
    assert np.all(st["code"] == 0)
    print(f"free-running q err={rel_err(q, ost.q):.3e} v err={rel_err(v, ost.v):.3e}", flush=True)
    assert rel_err(q, ost.q) <= REL_TOL
    assert rel_err(v, ost.v) <= REL_TOL


@pytest.mark.parametrize("solver", SOLVERS)
def test_adaptive_residual_mode_matches_oracle(gpu, solver):
    """Adaptive cone solve (iterations until an iteration's max |dlambda| reaches
    the tolerance, capped): same iteration counts, residuals and trajectory."""
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 8, 13)
    cfg = model.default_config()
    cfg.solver = solver
    cfg.residual_tolerance = 1.0e-4
    cfg.max_iterations = 40
    world = mr.World(model, cfg)
    ost = ol.OracleState(model, cfg)
    steps = 20
    cmds = model.commands(0, steps)
    world.advance(cmds)
    o_cs = None
    for k in range(steps):
        _, o_cs, *_ = ost.step(cmds[k], control_step=k)
    q, v, s = world.state()
    g_cs = world.contact_statuses()[-1]
    np.testing.assert_array_equal(g_cs["solver_iterations"], o_cs["solver_iterations"])
    np.testing.assert_array_equal(g_cs["residuals"], o_cs["residuals"])
    assert np.any(o_cs["solver_iterations"] > 6), "the tolerance never asked for more sweeps than the default"
    assert rel_err(q, ost.q) <= REL_TOL and rel_err(v, ost.v) <= REL_TOL


@pytest.mark.parametrize("solver", SOLVERS)
def test_full_size_humanoid_batch_matches_oracle_on_sampled_worlds(gpu, solver):
    """The benchmark's shape (4096 humanoid worlds, two rounds of resident workgroups):
    16 worlds from both ends and the middle of the grid, stepped by the oracle alone,
    end where the GPU's batch puts them, bit for bit (state, manifolds, statuses)."""
    E = 4096
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, E, 2024)
    cfg = model.default_config()
    cfg.solver = solver
    world = mr.World(model, cfg)
    steps = 10
    cmds = model.commands(0, steps)
    world.advance(cmds)
    q, v, s = world.state()
    mh, mp, mc = world.manifolds()
    g_status = world.statuses()[-1]
    ost = ol.OracleState(model, cfg)
    sample = list(range(0, 6)) + list(range(2045, 2050)) + list(range(E - 5, E))
    for k in range(steps):
        for e in sample:
            o_status, *_ = ost.step(cmds[k], control_step=k, env_begin=e, env_end=e + 1)
    for e in sample:
        assert np.array_equal(q[e].view(np.uint32), ost.q[e].view(np.uint32)), e
        assert np.array_equal(v[e].view(np.uint32), ost.v[e].view(np.uint32)), e
        assert np.array_equal(s[e].view(np.uint8), ost.scene[e].view(np.uint8)), e
        n = int(mc[e])
        assert n == int(ost.mcount[e])
        np.testing.assert_array_equal(mh[e, :n]["pair_and_count"], ost.mh[e, :n]["pair_and_count"])
        for f in ("feature_and_life", "impulses", "local_anchor_a", "local_anchor_b"):
            np.testing.assert_array_equal(mp[e, :n][f], ost.mp[e, :n][f], err_msg=f"env {e} {f}")
        assert int(g_status["code"][e]) == 0
    assert sum(int(mc[e]) for e in sample) > 0


@pytest.mark.parametrize("solver", SOLVERS)
def test_maximum_normal_impulse_matches_oracle(gpu, solver):
    """A normal-impulse cap that binds (humanoid feet and chain-box contacts) in both solver
    modes: the GPU trajectory and stored impulses equal the oracle's bit for bit, and the
    cap holds on every stored impulse."""
    for kind, envs, steps, cap in ((mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 6, 12, 0.05),
                                   (mr.MR_SYNTHETIC_CHAIN_BOXES, 6, 60, 0.02)):
        model = mr.WorldModel.synthetic(kind, envs, 4)
        cfg = model.default_config()
        cfg.solver = solver
        cfg.maximum_normal_impulse = cap
        world = mr.World(model, cfg)
        ost = ol.OracleState(model, cfg)
        cmds = model.commands(0, steps)
        world.advance(cmds)
        for k in range(steps):
            ost.step(cmds[k], control_step=k)
        q, v, s = world.state()
        mh, mp, mc = world.manifolds()
        assert np.array_equal(q.view(np.uint32), ost.q.view(np.uint32)), kind
        assert np.array_equal(v.view(np.uint32), ost.v.view(np.uint32)), kind
        assert np.array_equal(mc, ost.mcount)
        bound = 0
        for e in range(envs):
            n = int(mc[e])
            lam = mp[e, :n]["impulses"][..., 0]
            np.testing.assert_array_equal(lam.view(np.uint32), ost.mp[e, :n]["impulses"][..., 0].view(np.uint32))
            assert np.all(lam <= np.float32(cap))
            bound += int(np.sum(lam == np.float32(cap)))
        assert bound > 0, f"the cap never bound ({kind})"
