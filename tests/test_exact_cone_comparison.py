"""temporalCone against the FP64 exact-cone oracle over a horizon (CPU only).

For one humanoid world, every control step's final-microstep contact
problem is captured from the FP32 oracle (which the GPU matches bit for
bit): Jacobian rows J, response rows M^-1 J^T, free velocity, targets,
regularisation and elliptic cone coefficients. The same program is solved
exactly in FP64 (H = J M^-1 J^T + R, l = J v_free - target) and the
post-impulse velocities are compared. temporalCone runs a fixed 4+2 block
sweeps with a radial cone projection, so it is not the exact Coulomb
solution; the test states the gap and checks that the solve moves the
velocity towards the exact one. The reference reports the same comparison
against its exact solver (apps/reference_conic_probe.cpp, ReferenceConicSolver).
"""

import numpy as np

import metalrobo as mr
import oracle_lib as ol


def horizon_errors(kind, envs, steps, seed, env):
    model = mr.WorldModel.synthetic(kind, envs, seed)
    cfg = model.default_config()
    st = ol.OracleState(model, cfg)
    cmds = model.commands(0, steps)
    rows = []
    for k in range(steps):
        _, prob = ol.contact_problem_capture(st, cmds[k], env, control_step=k)
        if prob is None or prob["contacts"] == 0 or prob["dynamic_scene"]:
            continue
        J, R = prob["J"], prob["R"]
        H = J @ R.T + np.diag(prob["reg"])
        H = 0.5 * (H + H.T)
        lin = J @ prob["v_free"] - prob["target"]
        code, lam, stats = ol.exact_cone_program(H, lin, prob["mu"])
        assert code == 0, stats
        v_exact = prob["v_free"] + R.T @ lam
        scale = 1.0 + np.linalg.norm(v_exact)
        rows.append((np.linalg.norm(prob["v_final"] - v_exact) / scale,
                     np.linalg.norm(prob["v_free"] - v_exact) / scale,
                     stats["iterations"], prob["contacts"]))
    return np.array(rows)


def test_humanoid_temporal_cone_vs_exact_cone_over_horizon():
    rows = horizon_errors(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 2, 60, 2024, 0)
    assert len(rows) >= 20, "too few contact steps in the horizon"
    tc, free = rows[:, 0], rows[:, 1]
    print(f"humanoid exact-cone gap over {len(rows)} contact steps: velocity rel err mean {tc.mean():.3e} "
          f"max {tc.max():.3e} (free-velocity gap mean {free.mean():.3e}); exact iterations mean "
          f"{rows[:, 2].mean():.0f}; contacts mean {rows[:, 3].mean():.1f}", flush=True)
    assert np.all(np.isfinite(tc))
    assert np.mean(tc < free) >= 0.9
