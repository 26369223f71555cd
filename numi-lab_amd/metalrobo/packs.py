"""Learning packs and compiled policies through the C ABI (include/mr_packs.h).

PolicyPack is the host form of a policy pack (numpy tables), read and written
by the native library with the reference's wire format; Policy is one
compiled into device memory for a fixed environment count; the rollout-pack
writer publishes a batch of rollout samples."""

from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .native import (LEARNING_TRANSITION_DTYPE, MetalRoboError, PolicyDenseLayerC, PolicyOutcomeDescriptorC,
                     PolicyPackC, PolicyRolloutPackC, load)

ACTIVATION_IDENTITY, ACTIVATION_RELU, ACTIVATION_TANH, ACTIVATION_ELU, ACTIVATION_SILU = range(5)


@dataclass
class DenseLayer:
    weights: np.ndarray  # [output][input] float32
    bias: np.ndarray     # [output] float32
    activation: int = ACTIVATION_IDENTITY

    @property
    def input_count(self) -> int:
        return int(self.weights.shape[1])

    @property
    def output_count(self) -> int:
        return int(self.weights.shape[0])


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))


@dataclass
class PolicyPack:
    """PolicyPack (PolicyProgram.hpp:39-76): actor and critic dense networks,
    observation normalisation, action distribution and the semantic contract
    (version, world, task, observation, action fingerprints)."""
    id: str
    revision: int
    layers: list
    contract: tuple = (0, 0, 0, 0, 0)
    observation_mean: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    observation_inverse_std: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    critic_observation_mean: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    critic_observation_inverse_std: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    critic_layers: list = field(default_factory=list)
    action_log_std: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    action_bias: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    action_scale: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    observation_clip: float = 100.0
    action_clip: float = float(np.finfo(np.float32).max)
    format_version: int = 4
    content_hash: int = 0

    def to_c(self):
        """(PolicyPackC, keep-alive list): the C record pointing into this pack's arrays."""
        keep = []

        def table(arr):
            a = _f32(arr)
            keep.append(a)
            return (a.ctypes.data if a.size else None), a.size

        def layers(ls):
            recs = (PolicyDenseLayerC * max(len(ls), 1))()
            for i, l in enumerate(ls):
                w, nw = table(l.weights)
                b, nb = table(l.bias)
                recs[i] = PolicyDenseLayerC(int(l.input_count), int(l.output_count), int(l.activation), w, nw, b, nb)
            keep.append(recs)
            return (ct.addressof(recs) if ls else None), len(ls)

        c = PolicyPackC()
        c.id = self.id.encode()
        keep.append(c.id)
        c.revision = int(self.revision)
        (c.contract_version, c.world_fingerprint, c.task_fingerprint, c.observation_fingerprint,
         c.action_fingerprint) = (int(x) for x in self.contract)
        for name, arr in (("observation_mean", self.observation_mean),
                          ("observation_inverse_standard_deviation", self.observation_inverse_std),
                          ("critic_observation_mean", self.critic_observation_mean),
                          ("critic_observation_inverse_standard_deviation", self.critic_observation_inverse_std),
                          ("action_log_standard_deviation", self.action_log_std),
                          ("action_bias", self.action_bias), ("action_scale", self.action_scale)):
            p, n = table(arr)
            setattr(c, name, p)
            setattr(c, name + "_count", n)
        c.layers, c.layer_count = layers(self.layers)
        c.critic_layers, c.critic_layer_count = layers(self.critic_layers)
        c.observation_clip = float(self.observation_clip)
        c.action_clip = float(self.action_clip)
        return c, keep

    def write(self, path: str) -> None:
        """mr_write_policy_pack: validate and publish at path."""
        b = load()
        c, _keep = self.to_c()
        b.check(b.lib.mr_write_policy_pack(ct.byref(c), str(path).encode()))

    @classmethod
    def read(cls, path: str) -> "PolicyPack":
        """mr_read_policy_pack: read and validate a pack of wire version 3 or 4."""
        b = load()
        h = b.lib.mr_read_policy_pack(str(path).encode())
        if not h:
            raise MetalRoboError(b.last_error())
        try:
            v = b.lib.mr_policy_pack_view(h).contents

            def floats(ptr, n):
                return np.ctypeslib.as_array(ct.cast(ptr, ct.POINTER(ct.c_float)), (n,)).copy() if n else \
                    np.zeros(0, np.float32)

            def layers(ptr, n):
                out = []
                if n:
                    recs = ct.cast(ptr, ct.POINTER(PolicyDenseLayerC))
                    for i in range(n):
                        r = recs[i]
                        out.append(DenseLayer(floats(r.weights, r.weight_count).reshape(r.output_count, r.input_count),
                                              floats(r.bias, r.bias_count), int(r.activation)))
                return out

            return cls(
                id=v.id.decode(), revision=int(v.revision),
                contract=(int(v.contract_version), int(v.world_fingerprint), int(v.task_fingerprint),
                          int(v.observation_fingerprint), int(v.action_fingerprint)),
                observation_mean=floats(v.observation_mean, v.observation_mean_count),
                observation_inverse_std=floats(v.observation_inverse_standard_deviation,
                                               v.observation_inverse_standard_deviation_count),
                layers=layers(v.layers, v.layer_count),
                critic_observation_mean=floats(v.critic_observation_mean, v.critic_observation_mean_count),
                critic_observation_inverse_std=floats(v.critic_observation_inverse_standard_deviation,
                                                      v.critic_observation_inverse_standard_deviation_count),
                critic_layers=layers(v.critic_layers, v.critic_layer_count),
                action_log_std=floats(v.action_log_standard_deviation, v.action_log_standard_deviation_count),
                action_bias=floats(v.action_bias, v.action_bias_count),
                action_scale=floats(v.action_scale, v.action_scale_count),
                observation_clip=float(v.observation_clip), action_clip=float(v.action_clip),
                format_version=int(b.lib.mr_policy_pack_format_version(h)),
                content_hash=int(b.lib.mr_policy_pack_content_hash(h)))
        finally:
            b.lib.mr_policy_pack_destroy(h)


def learning_pack_content_hash(payload: bytes) -> int:
    """mr_learning_pack_content_hash: the wire hash of one payload."""
    b = load()
    buf = (ct.c_uint8 * max(len(payload), 1)).from_buffer_copy(payload or b"\0")
    return int(b.lib.mr_learning_pack_content_hash(buf if payload else None, len(payload)))


class Policy:
    """mr_policy_*: a policy compiled into device memory for `environments`
    environments. act() takes contiguous float32 GPU tensors."""

    def __init__(self, pack: PolicyPack, environments: int) -> None:
        self._b = load()
        c, keep = pack.to_c()
        self._h = self._b.lib.mr_policy_create(ct.byref(c), int(environments))
        del keep
        if not self._h:
            raise MetalRoboError(self._b.last_error())
        self.environments = int(environments)
        self.revision = int(self._b.lib.mr_policy_revision(self._h))
        self.actor_input = int(self._b.lib.mr_policy_actor_input_count(self._h))
        self.critic_input = int(self._b.lib.mr_policy_critic_input_count(self._h))
        self.actions = int(self._b.lib.mr_policy_action_count(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._b.lib.mr_policy_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def act(self, actor_observations, task_states, seed: int, control_step: int = 0, critic_observations=None):
        """(actions, latents, log_probabilities, values or None) as new GPU tensors."""
        import torch

        E, A = self.environments, self.actions
        dev = actor_observations.device
        actions = torch.empty((E, A), dtype=torch.float32, device=dev)
        latents = torch.empty((E, A), dtype=torch.float32, device=dev)
        logp = torch.empty(E, dtype=torch.float32, device=dev)
        values = torch.zeros(E, dtype=torch.float32, device=dev) if critic_observations is not None else None
        stream = torch.cuda.current_stream().cuda_stream
        self._b.check(self._b.lib.mr_policy_act(
            self._h, actor_observations.data_ptr(),
            None if critic_observations is None else critic_observations.data_ptr(), task_states.data_ptr(),
            int(control_step), int(seed), actions.data_ptr(), latents.data_ptr(), logp.data_ptr(),
            None if values is None else values.data_ptr(), stream))
        return actions, latents, logp, values


def write_policy_rollout_pack(path: str, *, id: str, task_fingerprint: int, policy_fingerprint: int,
                              policy_revision: int, environment_count: int, control_step_count: int,
                              actor_observation_count: int, critic_observation_count: int, action_count: int,
                              actor_observations, critic_observations, latents, log_probabilities, values,
                              bootstrap_values, transitions, motion_feature_count: int = 0, motion_features=None,
                              teacher_actions=None, outcomes=(), outcome_values=None) -> None:
    """mr_write_policy_rollout_pack: tables are step-major numpy arrays;
    transitions use LEARNING_TRANSITION_DTYPE; outcomes are (id, unit, direction)."""
    b = load()
    keep = []

    def table(arr):
        if arr is None:
            return None, 0
        a = _f32(arr)
        keep.append(a)
        return (a.ctypes.data if a.size else None), a.size

    c = PolicyRolloutPackC()
    c.id = id.encode()
    keep.append(c.id)
    c.task_fingerprint, c.policy_fingerprint, c.policy_revision = task_fingerprint, policy_fingerprint, policy_revision
    (c.environment_count, c.control_step_count, c.actor_observation_count, c.critic_observation_count,
     c.action_count, c.motion_feature_count) = (environment_count, control_step_count, actor_observation_count,
                                                critic_observation_count, action_count, motion_feature_count)
    for name, count, arr in (("actor_observations", "actor_observation_count_total", actor_observations),
                             ("critic_observations", "critic_observation_count_total", critic_observations),
                             ("motion_features", "motion_feature_count_total", motion_features),
                             ("teacher_actions", "teacher_action_count", teacher_actions),
                             ("latents", "latent_count", latents),
                             ("log_probabilities", "log_probability_count", log_probabilities),
                             ("values", "value_count", values),
                             ("bootstrap_values", "bootstrap_value_count", bootstrap_values),
                             ("outcome_values", "outcome_value_count", outcome_values)):
        p, n = table(arr)
        setattr(c, name, p)
        setattr(c, count, n)
    outs = (PolicyOutcomeDescriptorC * max(len(outcomes), 1))()
    for i, (oid, unit, direction) in enumerate(outcomes):
        outs[i] = PolicyOutcomeDescriptorC(oid.encode(), unit.encode(), int(direction))
    keep.append(outs)
    c.outcomes, c.outcome_count = (ct.addressof(outs) if outcomes else None), len(outcomes)
    tr = np.ascontiguousarray(transitions, LEARNING_TRANSITION_DTYPE)
    keep.append(tr)
    c.transitions, c.transition_count = (tr.ctypes.data if tr.size else None), tr.size
    b.check(b.lib.mr_write_policy_rollout_pack(ct.byref(c), str(path).encode()))


def read_policy_rollout_pack(path: str) -> dict:
    """Parses a rollout pack (header, content hash, every table) in Python;
    raises ValueError on any inconsistency. An independent reader of the wire
    format (serializePolicyRollout, LearningPacks.cpp:1818-1904)."""
    import struct

    data = open(path, "rb").read()
    magic, version, kind, nbytes, chash = struct.unpack_from("<8sIIQQ", data, 0)
    if magic != b"MRLEARN\0" or kind != 3 or version != 7 or nbytes != len(data) - 32:
        raise ValueError("not a version-7 rollout pack")
    payload = data[32:]
    h = 14695981039346656037
    for byte in payload:
        h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    if (h or 1) != chash:
        raise ValueError("content hash mismatch")
    at = 0

    def take(fmt):
        nonlocal at
        v = struct.unpack_from(fmt, payload, at)
        at += struct.calcsize(fmt)
        return v[0] if len(v) == 1 else v

    def text():
        nonlocal at
        n = take("<Q")
        s = payload[at:at + n].decode()
        at += n
        return s

    def table(dtype):
        nonlocal at
        n = take("<Q")
        a = np.frombuffer(payload, dtype, n, at).copy()
        at += n * np.dtype(dtype).itemsize
        return a

    out = {"content_hash": chash, "id": text()}
    out["task_fingerprint"], out["policy_fingerprint"], out["policy_revision"] = take("<QQQ")
    for k in ("environment_count", "control_step_count", "actor_observation_count", "critic_observation_count",
              "action_count", "motion_feature_count"):
        out[k] = take("<I")
    for k in ("actor_observations", "critic_observations", "motion_features", "teacher_actions", "latents",
              "log_probabilities", "values", "bootstrap_values"):
        out[k] = table("<f4")
    outcomes = []
    for _ in range(take("<Q")):
        oid, unit = text(), text()
        outcomes.append((oid, unit, take("<I")))
    out["outcomes"] = outcomes
    out["outcome_values"] = table("<f4")
    out["transitions"] = table(LEARNING_TRANSITION_DTYPE)
    if at != len(payload):
        raise ValueError("trailing bytes in rollout pack")
    return out
