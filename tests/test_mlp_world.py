"""The learned warm start inside the control-step kernel (the MLP instance).

Bit-exact plumbing checks against the plain kernel on the chain-and-boxes
workload:
  * sample capture with the plain warm start changes nothing (states,
    manifolds, statuses), and its targets are the impulses the step stored;
  * an MLP whose correction is zero reproduces the plain path, and a
    constant correction changes only the sweeps the solve needs;
each from a settled state with live contacts (40 plain control steps), and
comparing manifolds up to each environment's count.
  * when every MLP-warm-started solve misses the tolerance, the fallback
    re-solves from the plain warm start and reproduces the plain path bit for
    bit, counting one fallback per microstep and twice the sweeps.
Training on captured samples and the iterations saved at equal residual are
measured by bench.py's warm-start leg.
"""

import numpy as np
import pytest


def zero_mlp():
    from metalrobo import warmstart as ws

    m = ws.WarmStartMLP()
    state = {k: np.zeros(shape, np.float32) for k, shape in ws.SHAPES.items()}
    state.update(in_shift=np.zeros(8, np.float32), in_scale=np.ones(8, np.float32),
                 out_scale=np.ones(3, np.float32))
    m.load_state(state)
    m.min_correction = 0.0
    return m


def chain_world(envs=8, **overrides):
    import metalrobo as mr

    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, envs, 7)
    cfg = model.default_config()
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return model, cfg


_SETTLED = {}


def settled(model):
    """State and manifolds after 40 plain control steps (boxes landed, contacts live)."""
    import metalrobo as mr

    key = id(model)
    if key not in _SETTLED:
        w = mr.World(model, model.default_config())
        w.advance(model.commands(0, 40))
        _SETTLED[key] = (*w.state(), *w.manifolds())
    return _SETTLED[key]


def run(model, cfg, steps, setup=None, before_step=None):
    import metalrobo as mr

    q0, v0, s0, h0, p0, c0 = settled(model)
    assert c0.sum() > 0
    w = mr.World(model, cfg)
    w.set_state(q0, v0, s0)
    w.write_manifolds(h0, p0, c0)
    if setup:
        setup(w)
    cmds = model.commands(40, steps)
    for t in range(steps):
        if before_step:
            before_step(w)
        w.advance(cmds[t:t + 1])
    q, v, s = w.state()
    h, p, c = w.manifolds()
    # entries past each environment's manifold count are not state: clear them
    for e in range(len(c)):
        h[e, c[e]:] = np.zeros(1, h.dtype)
        p[e, c[e]:] = np.zeros(1, p.dtype)
    return w, (q, v, s, h, p, c, w.statuses(), w.contact_statuses())


NAMES = ("q", "v", "scene", "manifold headers", "manifold points", "manifold counts", "statuses",
         "contact statuses")


def assert_same(a, b):
    for i, (x, y) in enumerate(zip(a, b)):
        same = np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
        assert same, (NAMES[i], a[5], b[5], a[6]["code"], b[6]["code"])


@pytest.mark.gpu
def test_capture_changes_nothing_and_records_the_stored_impulses(gpu):
    model, cfg = chain_world(capture_substep=3)
    _, plain = run(model, cfg, 6)
    w, cap = run(model, cfg, 6, setup=lambda w: w.set_mlp_capture(True))
    assert cap[7][-1]["constraints"].sum() > 0
    assert_same(plain, cap)
    samples = w.mlp_samples()
    cs = cap[7][-1]
    _, p, c = cap[3], cap[4], cap[5]
    for env in range(model.environment_count):
        rows = samples[env]
        valid = rows[:, 11] == 1.0
        assert valid.sum() == cs["constraints"][env]
        assert np.all(rows[valid.sum():, 11] == 0.0)
        hdr = cap[3][env]
        stored = []
        for m in range(int(c[env])):
            n = int(hdr[m]["pair_and_count"][3])
            stored += [tuple(p[env, m, k]["impulses"][:3]) for k in range(n)]
        got = sorted(tuple(r) for r in rows[valid, 8:11])
        assert got == sorted(stored)
        assert np.all(np.isfinite(rows[valid, :8]))
        assert np.all(rows[valid, 2] >= 0.0) and np.all(rows[valid, 1] >= -0.05)


@pytest.mark.gpu
def test_zero_correction_reproduces_the_plain_path(gpu):
    """A network whose correction is zero leaves every warm start untouched: the
    MLP instance then reproduces the plain path bit for bit."""
    model, cfg = chain_world(residual_tolerance=1.0e-5, max_iterations=64)
    mlp = zero_mlp()
    _, plain = run(model, cfg, 6)
    _, learned = run(model, cfg, 6, setup=lambda w: w.set_warm_start_mlp(mlp))
    assert learned[7]["constraints"].sum() > 0
    assert_same(plain, learned)


@pytest.mark.gpu
def test_learned_correction_changes_the_warm_start_only(gpu):
    """A constant correction moves the warm start; every microstep still
    reaches the tolerance (or falls back) and the contact sets stay those the
    plain path finds on its first step."""
    model, cfg = chain_world(residual_tolerance=1.0e-5, max_iterations=128)
    mlp = zero_mlp()
    b3 = mlp.master["b3"]
    b3.fill_(0.25)
    _, plain = run(model, cfg, 1)
    _, learned = run(model, cfg, 1, setup=lambda w: w.set_warm_start_mlp(mlp))
    pc, lc = plain[7][-1], learned[7][-1]
    np.testing.assert_array_equal(pc["constraints"], lc["constraints"])
    assert np.all((lc["residuals"][:, 0] <= 1.0e-5) | (lc["diagnostics"][:, 2] > 0))
    assert not np.array_equal(pc["residuals"][:, 3], lc["residuals"][:, 3])


@pytest.mark.gpu
def test_forced_fallback_reproduces_the_plain_path(gpu):
    model, cfg = chain_world(residual_tolerance=1.0e-30, max_iterations=3)
    mlp = zero_mlp()
    mlp.master["b3"].fill_(0.25)  # a correction that replaces every warm start
    _, plain = run(model, cfg, 6)
    _, fb = run(model, cfg, 6, setup=lambda w: w.set_warm_start_mlp(mlp))
    assert fb[7]["constraints"].sum() > 0
    assert_same(plain[:7], fb[:7])
    pc, fc = plain[7], fb[7]
    for k in ("code", "constraints", "solver_iterations", "manifolds"):
        assert np.array_equal(pc[k], fc[k]), k
    np.testing.assert_array_equal(pc["residuals"][..., :3], fc["residuals"][..., :3])
    np.testing.assert_array_equal(fc["residuals"][..., 3], 2.0 * pc["residuals"][..., 3])
    # every microstep with constraints fell back once
    assert np.all(fc["diagnostics"][..., 2] == pc["residuals"][..., 3] / 3.0)


def _project_cone(lam, mu_u, mu_v):
    """The cone projection the kernel applies to a learned warm start (normal clamp, radial
    elliptic scaling of the tangents), in float32."""
    lam = lam.astype(np.float32).copy()
    f32 = np.float32
    lam[:, 0] = np.maximum(lam[:, 0], f32(0.0))
    lu, lv = np.float32(mu_u) * lam[:, 0], np.float32(mu_v) * lam[:, 0]
    ok = (lu > 0) & (lv > 0)
    lam[~ok, 1:] = 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        ns = (lam[:, 1] * lam[:, 1]) / (lu * lu) + (lam[:, 2] * lam[:, 2]) / (lv * lv)
        s = (f32(1.0) / np.sqrt(ns)).astype(np.float32)
    big = ok & (ns > 1.0)
    lam[big, 1] *= s[big]
    lam[big, 2] *= s[big]
    return lam


@pytest.mark.gpu
def test_in_solver_inference_matches_the_mlp_oracle(gpu):
    """Trained weights inside the solver, against the CPU restatement oracle/mlp_oracle.py on the
    features the kernel captured at the start of each contact's cone solve:
      * the network's scaled correction (normalize, forward with BF16 operands, scale) agrees
        with the oracle within 2e-3 of each output's scale, the forward-pass tolerance (BF16
        operands, FP32 against float64 accumulation), and the in-bounds flag agrees exactly;
      * given the kernel's own correction, the decision and the warm start are exact: a contact
        is replaced iff its features are in bounds and some component exceeds min_correction,
        and then its warm start is the previous impulse plus the correction projected onto the
        mu = 0.8 cone (normal clamp, radial tangent scaling), bit for bit; otherwise it keeps
        the previous impulse bit for bit."""
    import os
    import sys

    import metalrobo as mr
    from metalrobo import warmstart as ws

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import mlp_oracle as mo

    model, cfg = chain_world(envs=32, residual_tolerance=1.0e-5, max_iterations=128, capture_substep=1)
    mats = model.materials()["friction"]
    assert np.all(mats[:, :2] == np.float32(0.8))
    f, t = ws.collect_samples(model, cfg, control_steps=8, settle_steps=40)
    mlp = ws.WarmStartMLP(seed=5)
    mlp.fit(f, t, steps=150, batch=4096)
    gate = np.float32(0.05)
    mlp.min_correction = float(gate)
    mlp.max_normal_residual = 0.0
    q0, v0, s0, h0, p0, c0 = settled(model)
    w = mr.World(model, cfg)
    w.set_state(q0, v0, s0)
    w.write_manifolds(h0, p0, c0)
    w.set_warm_start_mlp(mlp)
    w.set_mlp_capture(True)
    w.advance(model.commands(40, 1))
    rows = w.mlp_samples().reshape(-1, 20)
    rows = rows[rows[:, 11] == 1.0]
    assert len(rows) > 200
    st = mlp.numpy_state()
    x = rows[:, :8]
    xn = ((x - st["in_shift"]) * st["in_scale"]).astype(np.float32)
    in_range = np.all(np.abs(xn) <= 8.0, axis=1)
    np.testing.assert_array_equal(rows[:, 19] == 1.0, in_range)
    d_ref = mo.forward(st, x)["y_hat"] * st["out_scale"].astype(np.float64)
    d_gpu = rows[:, 16:19]
    for k in range(3):
        scale = max(float(np.abs(d_ref[in_range, k]).max()), 1e-12)
        err = float(np.abs(d_gpu[in_range, k] - d_ref[in_range, k]).max())
        assert err <= 2e-3 * scale, (k, err, scale)
    changed = in_range & (np.abs(d_gpu).max(axis=1) > gate)
    replaced = rows[:, 15] == 1.0
    print(f"in-solver inference: {len(rows)} contacts, {int(in_range.sum())} in bounds, "
          f"{int(replaced.sum())} warm starts replaced", flush=True)
    assert replaced.sum() > 0 and (~replaced).sum() > 0
    np.testing.assert_array_equal(replaced, changed)
    keep = ~replaced
    assert np.array_equal(rows[keep, 12:15].view(np.uint32), rows[keep, 5:8].view(np.uint32))
    mu = np.sqrt(np.float32(0.8) * np.float32(0.8))  # the kernel's material mixing, sqrt(mu_a mu_b)
    expect = _project_cone((rows[:, 5:8] + d_gpu).astype(np.float32), mu, mu)
    assert np.array_equal(rows[replaced, 12:15].view(np.uint32), expect[replaced].view(np.uint32))
    w.close()


@pytest.mark.gpu
def test_residual_guard_falls_back_to_the_plain_path(gpu):
    """A learned solve that reaches the tolerance but leaves a normal residual above
    max_normal_residual is solved again from the plain warm start. With the bound at 1e-30
    every microstep the network touched falls back, so the MLP instance reproduces the plain
    path bit for bit and counts the residual fallbacks (contact status diagnostics.w) among
    all fallbacks (diagnostics.z)."""
    model, cfg = chain_world(residual_tolerance=1.0e-5, max_iterations=128)
    mlp = zero_mlp()
    mlp.master["b3"].fill_(0.25)  # a correction that replaces every warm start
    mlp.max_normal_residual = 1.0e-30
    _, plain = run(model, cfg, 3)
    _, guarded = run(model, cfg, 3, setup=lambda w: w.set_warm_start_mlp(mlp))
    assert_same(plain[:7], guarded[:7])
    gc = guarded[7]
    assert gc["diagnostics"][..., 3].sum() > 0
    assert np.all(gc["diagnostics"][..., 3] <= gc["diagnostics"][..., 2])


@pytest.mark.gpu
def test_trained_warm_start_against_the_plain_one(gpu):
    """Trains the MLP on samples the plain solve produced (chain-and-boxes
    worlds), then runs held-out worlds at residual 1e-5 with and without it,
    and asserts the sign of what the trained network gains. Measured: it
    saves no sweeps on this workload (about 2.5 % more, with fallbacks on a
    few percent of microsteps), so the saving is negative and the learned
    warm start stays off by default. With the normal-residual guard, no
    world that stays within the healthy bound under the plain warm start
    leaves it under the learned one."""
    import metalrobo as mr
    from metalrobo import warmstart as ws

    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 256, 31)
    cfg = model.default_config()
    cfg.residual_tolerance = 1.0e-5
    cfg.max_iterations = 128
    f, t = ws.collect_samples(model, cfg, control_steps=20, settle_steps=40, max_samples=400_000)
    assert len(f) > 100_000
    mlp = ws.WarmStartMLP(seed=3)
    rep = mlp.fit(f, t, steps=300, batch=4096)
    assert rep.final_loss < rep.first_loss
    ev_model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_CHAIN_BOXES, 256, 32)
    ev = ws.evaluate(ev_model, cfg, mlp, control_steps=5, settle_steps=40)
    plain, learned = ev["plain"]["sweeps_per_step"], ev["mlp"]["sweeps_per_step"]
    print(f"sweeps per control step: plain {plain:.1f}, learned {learned:.1f} "
          f"({100.0 * (1.0 - learned / plain):+.1f} % saved), fallbacks {ev['mlp']['fallbacks_per_step']:.3f}, "
          f"healthy worlds leaving the bound {ev['mlp']['healthy_worlds_leaving_bound']}, "
          f"ms per step plain {ev['plain']['ms_per_step']:.2f} learned {ev['mlp']['ms_per_step']:.2f}", flush=True)
    assert ev["healthy_worlds"] > 0
    saving = 1.0 - learned / plain
    assert saving < 0.0, f"the trained network now saves {100.0 * saving:.1f} % of the sweeps: update the record"
    assert ev["mlp"]["healthy_worlds_leaving_bound"] == 0


@pytest.mark.gpu
def test_learned_warm_start_needs_the_ordered_sweeps(gpu):
    """The learned warm start and its fallback are built around the ordered sweeps; a world
    configured for the relaxed Jacobi iteration refuses it, and an unknown solver mode is
    refused when the world is created."""
    import metalrobo as mr

    model, cfg = chain_world(envs=2, residual_tolerance=1.0e-5, max_iterations=16)
    cfg.solver = mr.MR_SOLVER_RELAXED_JACOBI
    w = mr.World(model, cfg)
    with pytest.raises(mr.MetalRoboError, match="ordered sweeps"):
        w.set_warm_start_mlp(zero_mlp())
    w.close()
    cfg.solver = 7
    with pytest.raises(mr.MetalRoboError, match="unknown solver mode"):
        mr.World(model, cfg)
