"""Learning packs (include/mr_packs.h) and compiled policies.

CPU:
  * the content hash against the published FNV-1a 64 test vectors;
  * the reference's trained G1 policy pack (copied into oracle/_ref where the
    reference is present): header, content hash and every table read, checked
    against an independent Python parse of the payload, and written back byte
    for byte identical;
  * round trips, the wire-version-3 form without a contract, and every
    rejection readPack and validatePolicyArtifact make;
  * rollout packs written through the C ABI and parsed independently, and
    their validation.
GPU:
  * the reference's G1 actor and critic on the device against the oracle's
    dense-layer and sample restatement;
  * a policy driving the locomotion task (step without actions) against the
    oracle task driven by the oracle policy, bit for bit;
  * a rollout pack published from that task.
"""

import os
import struct

import numpy as np
import pytest

import metalrobo as mr
import oracle_lib as ol
from metalrobo import packs as pk
from metalrobo import task as tk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G1_PACK = os.path.join(ROOT, "oracle", "_ref", "g1_legs_locomotion.policypack")
needs_g1 = pytest.mark.skipif(not os.path.isfile(G1_PACK),
                              reason="the reference's policy pack was not copied (reference absent at build time)")
CONTRACT = (1, 0x1111, 0x2222, 0x3333, 0x4444)


def fnv1a64(data: bytes) -> int:
    h = 14695981039346656037
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h or 1


def parse_policy_payload(payload: bytes, version: int):
    """An independent reader of serializePolicy (LearningPacks.cpp:1709-1816)."""
    at = 0

    def take(fmt):
        nonlocal at
        v = struct.unpack_from(fmt, payload, at)
        at += struct.calcsize(fmt)
        return v[0]

    def floats():
        nonlocal at
        n = take("<Q")
        a = np.frombuffer(payload, "<f4", n, at)
        at += 4 * n
        return a

    def layers():
        out = []
        for _ in range(take("<Q")):
            i, o, act = take("<I"), take("<I"), take("<I")
            out.append((i, o, act, floats(), floats()))
        return out

    n = take("<Q")
    pid = payload[at:at + n].decode()
    at += n
    d = {"id": pid, "revision": take("<Q"), "mean": floats(), "inv": floats(), "layers": layers(),
         "cmean": floats(), "cinv": floats(), "critic": layers(), "log_std": floats(), "bias": floats(),
         "scale": floats(), "obs_clip": take("<f"), "act_clip": take("<f")}
    d["contract"] = tuple(take("<Q") for _ in range(5)) if version == 4 else (0,) * 5
    assert at == len(payload)
    return d


def small_pack(rng, inputs=6, hidden=5, actions=3, critic_inputs=4, stochastic=True, activation=pk.ACTIVATION_ELU):
    def layer(i, o, act):
        return pk.DenseLayer((rng.standard_normal((o, i)) / np.sqrt(i)).astype(np.float32),
                             (0.1 * rng.standard_normal(o)).astype(np.float32), act)

    return pk.PolicyPack(
        id="unit_policy", revision=3, contract=CONTRACT,
        layers=[layer(inputs, hidden, activation), layer(hidden, actions, pk.ACTIVATION_IDENTITY)],
        observation_mean=(0.1 * rng.standard_normal(inputs)).astype(np.float32),
        observation_inverse_std=(1.0 + rng.random(inputs)).astype(np.float32),
        critic_layers=[layer(critic_inputs, hidden, activation), layer(hidden, 1, pk.ACTIVATION_IDENTITY)],
        action_log_std=np.full(actions, -1.0, np.float32) if stochastic else np.zeros(0, np.float32),
        action_bias=(0.05 * rng.standard_normal(actions)).astype(np.float32),
        action_scale=np.full(actions, 0.5, np.float32), observation_clip=5.0, action_clip=1.0)


def test_content_hash_is_fnv1a64_with_the_published_vectors():
    # FNV-1a 64-bit reference values (offset basis 0xcbf29ce484222325, prime 0x100000001b3)
    assert mr.learning_pack_content_hash(b"") == 0xCBF29CE484222325
    assert mr.learning_pack_content_hash(b"a") == 0xAF63DC4C8601EC8C
    assert mr.learning_pack_content_hash(b"foobar") == 0x85944171F73967E8
    blob = bytes(np.random.default_rng(1).integers(0, 256, 4099, dtype=np.uint8))
    assert mr.learning_pack_content_hash(blob) == fnv1a64(blob)


@needs_g1
def test_reference_policy_pack_reads_as_its_payload_says_and_rewrites_identically(tmp_path):
    raw = open(G1_PACK, "rb").read()
    magic, version, kind, nbytes, chash = struct.unpack_from("<8sIIQQ", raw, 0)
    assert (magic, version, kind, nbytes) == (b"MRLEARN\0", 4, 2, len(raw) - 32)
    assert fnv1a64(raw[32:]) == chash
    ref = parse_policy_payload(raw[32:], version)
    p = mr.PolicyPack.read(G1_PACK)
    assert (p.id, p.revision, p.format_version, p.content_hash) == (ref["id"], ref["revision"], 4, chash)
    assert p.contract == ref["contract"] and p.contract[0] == 1
    assert [(l.input_count, l.output_count, l.activation) for l in p.layers] == \
        [(i, o, a) for i, o, a, _, _ in ref["layers"]] == [(240, 256, 3), (256, 256, 3), (256, 128, 3), (128, 29, 0)]
    assert [(l.input_count, l.output_count) for l in p.critic_layers] == [(98, 256), (256, 256), (256, 128), (128, 1)]
    for mine, (_, _, _, w, b) in zip(p.layers + p.critic_layers, ref["layers"] + ref["critic"]):
        assert np.array_equal(mine.weights.reshape(-1), w) and np.array_equal(mine.bias, b)
    for mine, theirs in ((p.observation_mean, "mean"), (p.observation_inverse_std, "inv"),
                         (p.critic_observation_mean, "cmean"), (p.critic_observation_inverse_std, "cinv"),
                         (p.action_log_std, "log_std"), (p.action_bias, "bias"), (p.action_scale, "scale")):
        assert np.array_equal(mine, ref[theirs]), theirs
    assert p.observation_clip == ref["obs_clip"] == 5.0
    out = tmp_path / "rewritten.policypack"
    p.write(str(out))
    assert out.read_bytes() == raw


def test_policy_pack_round_trip_and_version_three(tmp_path):
    p = small_pack(np.random.default_rng(3))
    path = tmp_path / "p.policypack"
    p.write(str(path))
    raw = path.read_bytes()
    assert struct.unpack_from("<IIQ", raw, 8) == (4, 2, len(raw) - 32)
    q = mr.PolicyPack.read(str(path))
    assert q.content_hash == fnv1a64(raw[32:]) == mr.learning_pack_content_hash(raw[32:])
    assert (q.id, q.revision, q.contract, q.observation_clip, q.action_clip) == \
        (p.id, p.revision, CONTRACT, 5.0, 1.0)
    for a, b in zip(p.layers + p.critic_layers, q.layers + q.critic_layers):
        assert np.array_equal(a.weights, b.weights) and np.array_equal(a.bias, b.bias)
        assert a.activation == b.activation
    assert np.array_equal(p.action_log_std, q.action_log_std)
    # version 3: the same payload without the trailing contract (LearningPacks.cpp:1801-1804)
    v3 = raw[32:-40]
    (tmp_path / "v3.policypack").write_bytes(struct.pack("<8sIIQQ", b"MRLEARN\0", 3, 2, len(v3), fnv1a64(v3)) + v3)
    r = mr.PolicyPack.read(str(tmp_path / "v3.policypack"))
    assert r.format_version == 3 and r.contract == (0, 0, 0, 0, 0) and r.revision == 3
    # a version-3 pack carries no exact contract, so it cannot be written as version 4
    with pytest.raises(mr.MetalRoboError, match="contract"):
        r.write(str(tmp_path / "again.policypack"))


def _expect_read_error(path, match):
    with pytest.raises(mr.MetalRoboError, match=match):
        mr.PolicyPack.read(str(path))


def test_policy_pack_reader_rejections(tmp_path):
    p = small_pack(np.random.default_rng(4))
    good = tmp_path / "good.policypack"
    p.write(str(good))
    raw = good.read_bytes()
    payload = raw[32:]
    header = lambda magic=b"MRLEARN\0", version=4, kind=2, n=len(payload), h=fnv1a64(payload): \
        struct.pack("<8sIIQQ", magic, version, kind, n, h)  # noqa: E731
    cases = {
        "flipped": (header() + payload[:-1] + bytes([payload[-1] ^ 1]), "fingerprint"),
        "magic": (header(magic=b"MRLEARX\0") + payload, "header or kind"),
        "kind": (header(kind=3) + payload, "header or kind"),
        "new_version": (header(version=5) + payload, "version"),
        "old_version": (header(version=2) + payload, "version"),
        "length": (header(n=len(payload) + 4) + payload, "length"),
        "truncated": (raw[:20], "length"),
    }
    extra = payload + b"\0\0\0\0"
    cases["trailing"] = (struct.pack("<8sIIQQ", b"MRLEARN\0", 4, 2, len(extra), fnv1a64(extra)) + extra, "malformed")
    # a valid hash over a layer count no payload could hold (2^31 layers): refused before allocating
    at = 8 + struct.unpack_from("<Q", payload, 0)[0] + 8  # id, revision
    for _ in range(2):  # observation mean and inverse std
        at += 8 + 4 * struct.unpack_from("<Q", payload, at)[0]
    huge = payload[:at] + struct.pack("<Q", 1 << 31) + payload[at + 8:]
    cases["layer_count"] = (struct.pack("<8sIIQQ", b"MRLEARN\0", 4, 2, len(huge), fnv1a64(huge)) + huge, "malformed")
    for name, (data, match) in cases.items():
        path = tmp_path / f"{name}.policypack"
        path.write_bytes(data)
        _expect_read_error(path, match)
    _expect_read_error(tmp_path / "missing.policypack", "missing")


def test_policy_pack_writer_rejections(tmp_path):
    rng = np.random.default_rng(5)
    out = str(tmp_path / "x.policypack")
    bad = []
    p = small_pack(rng); p.revision = 0; bad.append(("revision", p))
    p = small_pack(rng); p.contract = (2,) + CONTRACT[1:]; bad.append(("contract", p))
    p = small_pack(rng); p.contract = CONTRACT[:4] + (0,); bad.append(("contract", p))
    p = small_pack(rng); p.observation_clip = 0.0; bad.append(("clipping", p))
    p = small_pack(rng); p.layers[0].bias = p.layers[0].bias[:-1]; bad.append(("shape", p))
    p = small_pack(rng); p.layers[1].weights[0, 0] = np.nan; bad.append(("shape", p))
    p = small_pack(rng); p.layers[0].activation = 5; bad.append(("shape", p))
    p = small_pack(rng); p.critic_layers = []; bad.append(("stochastic", p))
    p = small_pack(rng); p.action_bias[0] = np.inf; bad.append(("stochastic", p))
    for match, pack in bad:
        with pytest.raises(mr.MetalRoboError, match=match):
            pack.write(out)
        assert not os.path.exists(out)
    p = small_pack(rng)
    p.layers = []
    with pytest.raises(mr.MetalRoboError, match="required"):
        p.write(out)


def _rollout(rng, E=3, T=2, O=5, C=4, A=2, revision=7):
    S = E * T
    tr = np.zeros(S, mr.LEARNING_TRANSITION_DTYPE)
    tr["reward_and_bootstrap"][:, 0] = rng.standard_normal(S)
    tr["termination"][:, 0] = rng.integers(0, 2, S)
    tr["context"][:, 2] = np.arange(S)
    tr["policy_revision"] = revision
    return dict(id="batch-0", task_fingerprint=11, policy_fingerprint=22, policy_revision=revision,
                environment_count=E, control_step_count=T, actor_observation_count=O, critic_observation_count=C,
                action_count=A, actor_observations=rng.standard_normal(S * O), critic_observations=rng.standard_normal(S * C),
                latents=rng.standard_normal(S * A), log_probabilities=rng.standard_normal(S), values=rng.standard_normal(S),
                bootstrap_values=rng.standard_normal(E), transitions=tr,
                outcomes=[("distance", "m", 1)], outcome_values=rng.standard_normal(S))


def test_rollout_pack_written_and_read_back(tmp_path):
    args = _rollout(np.random.default_rng(6))
    path = str(tmp_path / "r.rolloutpack")
    mr.write_policy_rollout_pack(path, **args)
    got = mr.read_policy_rollout_pack(path)
    for k in ("id", "task_fingerprint", "policy_fingerprint", "policy_revision", "environment_count",
              "control_step_count", "actor_observation_count", "critic_observation_count", "action_count"):
        assert got[k] == args[k], k
    for k in ("actor_observations", "critic_observations", "latents", "log_probabilities", "values",
              "bootstrap_values", "outcome_values"):
        assert np.array_equal(got[k], np.asarray(args[k], np.float32)), k
    assert got["motion_features"].size == 0 and got["teacher_actions"].size == 0
    assert got["outcomes"] == [("distance", "m", 1)]
    assert got["transitions"].tobytes() == args["transitions"].tobytes()


def test_rollout_pack_validation(tmp_path):
    path = str(tmp_path / "r.rolloutpack")
    rng = np.random.default_rng(7)
    mutations = {
        "dimensions": lambda a: a.update(latents=a["latents"][:-1]),
        "bootstrap": lambda a: a.update(bootstrap_values=a["bootstrap_values"][:-1]),
        "fingerprints": lambda a: a.update(policy_fingerprint=0),
        "inconsistent": lambda a: a["transitions"].__setitem__("policy_revision", 8),
        "non-finite": lambda a: a["values"].__setitem__(0, np.nan),
        "termination": lambda a: a["transitions"]["termination"].__setitem__((0, 1), 2),
        "outcome": lambda a: a.update(outcomes=[("distance", "m", 3)]),
    }
    for name, mutate in mutations.items():
        args = _rollout(rng)
        mutate(args)
        with pytest.raises(mr.MetalRoboError):
            mr.write_policy_rollout_pack(path, **args)
        assert not os.path.exists(path), name


# ------------------------------------------------------------------ GPU


def oracle_policy(pack, obs, states, seed, control_step=0, critic_obs=None):
    """The oracle restatement of inference: dense layers in order (the first
    normalises and clips), then sample and score."""
    def network(layers, x, mean, inv):
        for i, l in enumerate(layers):
            m = (mean if mean.size else np.zeros(l.input_count, np.float32)) if i == 0 else None
            s = (inv if inv.size else np.ones(l.input_count, np.float32)) if i == 0 else None
            x = ol.policy_dense(l.weights, l.bias, x, l.output_count, l.activation, m, s, pack.observation_clip)
        return x

    mean = network(pack.layers, obs, pack.observation_mean, pack.observation_inverse_std)
    A = mean.shape[1]
    bias = pack.action_bias if pack.action_bias.size else np.zeros(A, np.float32)
    scale = pack.action_scale if pack.action_scale.size else np.ones(A, np.float32)
    acts, lat, lp = ol.policy_sample(mean, bias, scale, pack.action_log_std if pack.action_log_std.size else None,
                                     pack.action_clip, seed, pack.revision, states, control_step)
    values = None
    if critic_obs is not None:
        values = network(pack.critic_layers, critic_obs, pack.critic_observation_mean,
                         pack.critic_observation_inverse_std)[:, 0]
    return acts, lat, lp, values


@needs_g1
@pytest.mark.gpu
def test_gpu_reference_policy_matches_oracle(gpu):
    import torch

    pack = mr.PolicyPack.read(G1_PACK)
    E = 96
    rng = np.random.default_rng(8)
    print(f"G1 policy pack on {E} environments", flush=True)
    obs = (pack.observation_mean + rng.standard_normal((E, 240)) / pack.observation_inverse_std).astype(np.float32)
    obs[:4] *= 50.0  # far outside the training distribution: the observation clip binds
    cobs = (pack.critic_observation_mean + rng.standard_normal((E, 98)) /
            pack.critic_observation_inverse_std).astype(np.float32)
    states = np.zeros(E, mr.TASK_STATE_DTYPE)
    states["episode"][:, 0] = rng.integers(0, 50, E)
    states["episode"][:, 1] = np.arange(E)
    policy = mr.Policy(pack, E)
    assert (policy.actor_input, policy.critic_input, policy.actions, policy.revision) == (240, 98, 29, 108)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    st = dev(states.view(np.uint8).reshape(E, -1))
    acts, lat, lp, val = policy.act(dev(obs), st, seed=12345, control_step=3, critic_observations=dev(cobs))
    torch.cuda.synchronize()
    racts, rlat, rlp, rval = oracle_policy(pack, obs, states, 12345, 3, cobs)
    # ELU evaluates expf on the device and the host: agreement to float rounding through four layers
    np.testing.assert_allclose(lat.cpu().numpy(), rlat, rtol=1e-5, atol=2e-5)
    np.testing.assert_allclose(acts.cpu().numpy(), racts, rtol=1e-5, atol=2e-5)
    np.testing.assert_allclose(lp.cpu().numpy(), rlp, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(val.cpu().numpy(), rval, rtol=1e-5, atol=1e-4)
    assert np.isfinite(rlat).all() and np.abs(racts).max() > 0.0


@pytest.mark.gpu
def test_gpu_policy_drives_the_task_like_the_oracle(gpu, tmp_path):
    import torch

    E, steps = 8, 6
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, E, 5)
    cfg = model.default_config()
    prog = tk.humanoid_locomotion_program(model, seed=9, response_time=0.03, actuator_delay=1)
    prog["max_episode_steps"] = 4
    world = mr.World(model, cfg)
    task = mr.LocomotionTask(world, prog)
    otask = ol.OracleTask(model, cfg, prog)
    O = task.history_length * task.observation_count
    # ReLU and identity layers sum fused multiply-adds in one order on both sides: bit-exact
    rng = np.random.default_rng(10)
    pack = small_pack(rng, inputs=O, hidden=32, actions=task.action_count, critic_inputs=O, stochastic=False,
                      activation=pk.ACTIVATION_RELU)
    pack.revision = 41
    print(f"policy-driven task: {E} humanoid worlds, {steps} steps", flush=True)
    with pytest.raises(mr.MetalRoboError):
        task.step(None)  # no policy installed, no actions
    task.set_policy(pack)
    assert task.policy_revision() == 41
    otask.policy_revision = 41
    obs_log, lat_log, lp_log, tr_log = [], [], [], []
    for k in range(steps):
        o_before = otask.observations.reshape(E, -1).copy()
        acts, lat, lp, _ = oracle_policy(pack, o_before, otask.states.copy(), int(prog["seed"]), 0)
        task.step()
        otask.step(acts, control_step=0)
        g_acts, g_lat, g_lp = task.read_policy_outputs()
        assert np.array_equal(g_acts.view(np.uint32), acts.view(np.uint32)), f"step {k}: actions"
        obs, tr, st = task.read()
        q, v, _ = world.state()
        assert np.array_equal(q.view(np.uint32), otask.physics.q.view(np.uint32)), f"step {k}: q"
        assert np.array_equal(obs.view(np.uint32), otask.observations.view(np.uint32)), f"step {k}: observations"
        np.testing.assert_array_equal(tr["termination"], otask.transitions["termination"])
        assert (tr["policy_revision"] == 41).all() and (otask.transitions["policy_revision"] == 41).all()
        obs_log.append(o_before)
        lat_log.append(g_lat)
        lp_log.append(g_lp)
        tr_log.append(tr)
    # one rollout pack of the whole batch
    actor = np.stack(obs_log)
    path = str(tmp_path / "task.rolloutpack")
    task.write_policy_rollout_pack(path, "task-batch", steps, actor, actor, np.stack(lat_log), np.stack(lp_log),
                                   np.zeros((steps, E), np.float32), np.zeros(E, np.float32), np.concatenate(tr_log))
    got = mr.read_policy_rollout_pack(path)
    assert (got["environment_count"], got["control_step_count"], got["actor_observation_count"],
            got["action_count"], got["policy_revision"]) == (E, steps, O, task.action_count, 41)
    assert got["task_fingerprint"] != 0 and got["policy_fingerprint"] != 0
    assert np.array_equal(got["actor_observations"], actor.reshape(-1))
    trs = np.concatenate(tr_log)
    assert np.array_equal(got["transitions"]["reward_and_bootstrap"][:, 0], trs["reward_and_state"][:, 0])
    assert np.array_equal(got["transitions"]["termination"], trs["termination"])
    assert (got["transitions"]["policy_revision"] == 41).all()
    task.clear_policy()
    assert task.policy_revision() == 0
    with pytest.raises(mr.MetalRoboError):
        task.write_policy_rollout_pack(path, "x", steps, actor, actor, np.stack(lat_log), np.stack(lp_log),
                                       np.zeros((steps, E), np.float32), np.zeros(E, np.float32), trs)


@pytest.mark.gpu
def test_gpu_task_refuses_a_policy_of_the_wrong_shape(gpu):
    model = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 2, 5)
    prog = tk.humanoid_locomotion_program(model, seed=1)
    task = mr.LocomotionTask(mr.World(model, model.default_config()), prog)
    rng = np.random.default_rng(11)
    good = small_pack(rng, inputs=task.history_length * task.observation_count, hidden=8,
                      actions=task.action_count, critic_inputs=3, stochastic=False, activation=pk.ACTIVATION_RELU)
    for mutate, match in ((lambda p: setattr(p, "layers", [p.layers[0]]), "final layer"),
                          (lambda p: setattr(p.layers[1], "activation", pk.ACTIVATION_TANH), "final layer"),
                          (lambda p: setattr(p, "action_bias", p.action_bias[:-1]), "action bias"),
                          (lambda p: setattr(p, "observation_mean", p.observation_mean[:-1]), "normalisation"),
                          (lambda p: setattr(p, "action_log_std", np.full(task.action_count, 3.0, np.float32)),
                           r"\[-5, 2\]")):
        p = small_pack(np.random.default_rng(11), inputs=task.history_length * task.observation_count, hidden=8,
                       actions=task.action_count, critic_inputs=3, stochastic=False, activation=pk.ACTIVATION_RELU)
        mutate(p)
        with pytest.raises(mr.MetalRoboError, match=match):
            task.set_policy(p)
        assert task.policy_revision() == 0
    task.set_policy(good)
    assert task.policy_revision() == good.revision
