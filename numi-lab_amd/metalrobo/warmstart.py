"""Learned warm start for the cone solve: an 8 -> 256 -> 256 -> 3 MLP with BF16
weights on the gfx950 matrix cores, trained by the library's own kernels
(include/mr_mlp.h). PyTorch tensors only hold device memory and give the
stream; every forward, loss, backward and optimizer step is a call into the
C ABI.

Features per contact, at the start of a microstep's cone solve:
    normal relative velocity, penetration depth, the Euclidean norms of the
    normal and two tangent Jacobian rows, and the previous-step impulse
    (normal, tangent u, tangent v) the plain path warm-starts from.
The network predicts the correction to the previous-step impulse; the
initial impulse is lambda_prev + out_scale * y_hat (the solver projects it onto
the cone), trained on the impulse the plain solve converges to.
"""

from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import native as nat

FEATURES = ("normal_velocity", "penetration", "row_norm_n", "row_norm_u", "row_norm_v",
            "previous_lambda_n", "previous_lambda_u", "previous_lambda_v")
IN, HIDDEN, OUT = 8, 256, 3
SHAPES = {"w1": (HIDDEN, IN), "b1": (HIDDEN,), "w2": (HIDDEN, HIDDEN), "b2": (HIDDEN,), "w3": (OUT, HIDDEN),
          "b3": (OUT,)}
MATRICES = ("w1", "w2", "w3")


def _torch():
    import torch  # the Python edge only

    return torch


def _stream():
    return ct.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _p(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr())


@dataclass
class TrainReport:
    steps: int
    first_loss: float
    final_loss: float
    losses: list
    kept: int = 0
    dropped: int = 0


class WarmStartMLP:
    """Weights (FP32 master + BF16 copies), Adam state and normalization on one GPU."""

    def __init__(self, seed: int = 0, device: str = "cuda") -> None:
        torch = _torch()
        self._b = nat.load()
        self.device = device
        rng = np.random.default_rng(seed)
        init = {}
        for name, shape in SHAPES.items():
            if name.startswith("w"):
                fan_in = shape[1]
                init[name] = (rng.standard_normal(shape) * np.sqrt(2.0 / fan_in)).astype(np.float32)
            else:
                init[name] = np.zeros(shape, np.float32)
        self.master = {k: torch.from_numpy(v).to(device) for k, v in init.items()}
        self.bf16 = {k: self.master[k].to(torch.bfloat16).contiguous() for k in MATRICES}
        self.m1 = {k: torch.zeros_like(v) for k, v in self.master.items()}
        self.m2 = {k: torch.zeros_like(v) for k, v in self.master.items()}
        self.in_shift = torch.zeros(IN, dtype=torch.float32, device=device)
        self.in_scale = torch.ones(IN, dtype=torch.float32, device=device)
        self.out_scale = torch.ones(OUT, dtype=torch.float32, device=device)
        self.step = 0
        self.min_correction = 2.0  # impulse units; see MRMlpWeightsC.min_correction
        # a converged learned solve with a larger normal residual is re-solved from the plain warm
        # start (MRMlpWeightsC.max_normal_residual); the bound evaluate() calls healthy
        self.max_normal_residual = 1.0
        self._bufs = {}

    # -------------------------------------------------------------- views
    def weights(self) -> nat.MlpWeightsC:
        return nat.MlpWeightsC(_p(self.bf16["w1"]), _p(self.master["b1"]), _p(self.bf16["w2"]),
                               _p(self.master["b2"]), _p(self.bf16["w3"]), _p(self.master["b3"]),
                               _p(self.in_shift), _p(self.in_scale), _p(self.out_scale),
                               float(self.min_correction), float(self.max_normal_residual))

    def set_normalization(self, features, targets) -> None:
        """Feature mean / inverse std and target std from training data (rows x 8, rows x 3)."""
        torch = _torch()
        f = torch.as_tensor(features, dtype=torch.float32, device=self.device)
        t = torch.as_tensor(targets, dtype=torch.float32, device=self.device)
        self.in_shift.copy_(f.mean(0))
        self.in_scale.copy_(1.0 / f.std(0).clamp_min(1.0e-6))
        self.out_scale.copy_(t.std(0).clamp_min(1.0e-6))

    # -------------------------------------------------------------- kernels
    def _buffers(self, n: int):
        torch = _torch()
        if self._bufs.get("n") != n:
            d = self.device
            self._bufs = {
                "n": n,
                "xq": torch.empty((n, IN), dtype=torch.bfloat16, device=d),
                "z1": torch.empty((n, HIDDEN), dtype=torch.float32, device=d),
                "h1": torch.empty((n, HIDDEN), dtype=torch.bfloat16, device=d),
                "z2": torch.empty((n, HIDDEN), dtype=torch.float32, device=d),
                "h2": torch.empty((n, HIDDEN), dtype=torch.bfloat16, device=d),
                "y": torch.empty((n, OUT), dtype=torch.float32, device=d),
                "dy": torch.empty((n, OUT), dtype=torch.float32, device=d),
                "dz1": torch.empty((n, HIDDEN), dtype=torch.float32, device=d),
                "dz2": torch.empty((n, HIDDEN), dtype=torch.float32, device=d),
                "loss": torch.zeros(1, dtype=torch.float32, device=d),
                "scratch": torch.empty(self._b.lib.mr_mlp_loss_scratch_floats(n), dtype=torch.float32, device=d),
                "grads": {k: torch.empty_like(v) for k, v in self.master.items()},
            }
        return self._bufs

    def _trace(self, b) -> nat.MlpTraceC:
        return nat.MlpTraceC(_p(b["xq"]), _p(b["z1"]), _p(b["h1"]), _p(b["z2"]), _p(b["h2"]))

    def forward(self, x, trace: bool = False):
        """y_hat (normalized units) of features x [n][8] (device FP32)."""
        torch = _torch()
        x = x.contiguous()
        n = x.shape[0]
        b = self._buffers(n)
        w = self.weights()
        tr = self._trace(b) if trace else None
        y = b["y"] if trace else torch.empty((n, OUT), dtype=torch.float32, device=self.device)
        self._b.check(self._b.lib.mr_mlp_forward(ct.byref(w), _p(x), n, _p(y), ct.byref(tr) if tr else None,
                                                 _stream()))
        return y

    def predict(self, x):
        """Initial impulses (physical units, before the cone projection) for features x."""
        return x[:, 5:8] + self.forward(x) * self.out_scale

    def fit(self, features, targets, steps: int = 1500, batch: int = 8192, lr: float = 1.0e-3,
            bound: float = 1.0e3, seed: int = 0, huber_delta: float = 0.1) -> TrainReport:
        """Trains on captured samples (features, converged impulses): rows with a non-finite or
        larger-than-`bound` value are dropped, the normalization is fitted to the rest, and the
        network learns the correction converged - previous impulse with a Huber loss (the
        corrections are heavy-tailed: most contacts need almost none)."""
        f = np.asarray(features, np.float32)
        t = np.asarray(targets, np.float32)
        keep = np.all(np.isfinite(f), 1) & np.all(np.isfinite(t), 1)
        keep &= np.all(np.abs(f) <= bound, 1) & np.all(np.abs(t) <= bound, 1)
        f, t = f[keep], t[keep]
        r = t - f[:, 5:8]
        self.set_normalization(f, r)
        rep = self.train(f, r, steps=steps, batch=batch, lr=lr, seed=seed, huber_delta=huber_delta)
        rep.kept = int(keep.sum())
        rep.dropped = int((~keep).sum())
        return rep

    def loss_and_grads(self, x, t, huber_delta: float = 0.0):
        """Forward with trace, loss gradient (MSE, or Huber with huber_delta > 0) and backward;
        gradients land in the grads buffers."""
        n = x.shape[0]
        b = self._buffers(n)
        y = self.forward(x, trace=True)
        t = t.contiguous()
        self._b.check(self._b.lib.mr_mlp_loss(_p(y), _p(t), _p(self.out_scale), n, huber_delta, _p(b["dy"]),
                                              _p(b["loss"]), _p(b["scratch"]), _stream()))
        w = self.weights()
        tr = self._trace(b)
        g = b["grads"]
        gc = nat.MlpGradsC(_p(g["w1"]), _p(g["b1"]), _p(g["w2"]), _p(g["b2"]), _p(g["w3"]), _p(g["b3"]))
        self._b.check(self._b.lib.mr_mlp_backward(ct.byref(w), ct.byref(tr), _p(b["dy"]), n, _p(b["dz2"]),
                                                  _p(b["dz1"]), ct.byref(gc), _stream()))
        return b["loss"], g

    def adam_step(self, grads, lr=1.0e-3, beta1=0.9, beta2=0.999, eps=1.0e-8) -> None:
        self.step += 1
        for k, w in self.master.items():
            wb = _p(self.bf16[k]) if k in self.bf16 else None
            self._b.check(self._b.lib.mr_mlp_adam(_p(w), wb, _p(grads[k]), _p(self.m1[k]), _p(self.m2[k]),
                                                  w.numel(), lr, beta1, beta2, eps, self.step, _stream()))

    def train(self, features, targets, steps: int = 400, batch: int = 8192, lr: float = 1.0e-3,
              seed: int = 0, huber_delta: float = 0.0) -> TrainReport:
        """Minibatch Adam on (features, targets) resident on the GPU."""
        torch = _torch()
        f = torch.as_tensor(features, dtype=torch.float32, device=self.device)
        t = torch.as_tensor(targets, dtype=torch.float32, device=self.device)
        n = f.shape[0]
        batch = min(batch, n)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for s in range(steps):
            idx = torch.randint(0, n, (batch,), generator=gen).to(self.device)
            loss, grads = self.loss_and_grads(f[idx], t[idx], huber_delta)
            self.adam_step(grads, lr=lr)
            if s == 0 or (s + 1) % 50 == 0 or s + 1 == steps:
                losses.append(float(loss.item()))
        return TrainReport(steps, losses[0], losses[-1], losses)

    # -------------------------------------------------------------- I/O
    def numpy_state(self) -> dict:
        out = {k: v.detach().cpu().numpy() for k, v in self.master.items()}
        out.update(in_shift=self.in_shift.cpu().numpy(), in_scale=self.in_scale.cpu().numpy(),
                   out_scale=self.out_scale.cpu().numpy())
        return out

    def save(self, path: str | Path) -> None:
        np.savez(path, **self.numpy_state())

    def load_state(self, state: dict) -> None:
        torch = _torch()
        for k in self.master:
            self.master[k].copy_(torch.as_tensor(np.asarray(state[k], np.float32)))
        for k in MATRICES:
            self.bf16[k].copy_(self.master[k].to(torch.bfloat16))
        self.in_shift.copy_(torch.as_tensor(np.asarray(state["in_shift"], np.float32)))
        self.in_scale.copy_(torch.as_tensor(np.asarray(state["in_scale"], np.float32)))
        self.out_scale.copy_(torch.as_tensor(np.asarray(state["out_scale"], np.float32)))

    @classmethod
    def load(cls, path: str | Path, device: str = "cuda") -> "WarmStartMLP":
        m = cls(device=device)
        with np.load(path, allow_pickle=False) as z:
            m.load_state({k: z[k] for k in z.files})
        return m


# ------------------------------------------------------------------ workflows

def _settled_world(model, cfg, settle_steps: int):
    """A world advanced `settle_steps` control steps with `cfg` from the model's initial state."""
    from .native import World

    w = World(model, cfg)
    if settle_steps > 0:
        w.advance(model.commands(0, settle_steps))
    return w


def collect_samples(model, cfg, control_steps: int, settle_steps: int = 40, max_samples: int = 10_000_000,
                    capture_substep: int = 1):
    """(features [n][8], targets [n][3]) from the plain solve: one microstep per control step is
    captured (its features at the start of the solve and the impulse it converged to)."""
    cfg = cfg.copy()
    cfg.capture_substep = capture_substep
    w = _settled_world(model, cfg, settle_steps)
    w.set_mlp_capture(True)
    cmds = model.commands(settle_steps, control_steps)
    feats, targs, total = [], [], 0
    for t in range(control_steps):
        w.advance(cmds[t:t + 1])
        rows = w.mlp_samples().reshape(-1, 20)
        rows = rows[rows[:, 11] == 1.0]
        feats.append(rows[:, :8].copy())
        targs.append(rows[:, 8:11].copy())
        total += len(rows)
        if total >= max_samples:
            break
    f = np.concatenate(feats)[:max_samples]
    t = np.concatenate(targs)[:max_samples]
    w.close()
    return f, t


def evaluate(model, cfg, mlp, control_steps: int, settle_steps: int = 40, healthy_residual: float = 1.0) -> dict:
    """Cone sweeps per control step at the config's residual tolerance, from the same settled
    state, with the plain warm start and with `mlp`. Worlds whose plain run ever shows a normal
    residual above `healthy_residual` (diverging worlds) are reported apart."""
    torch = _torch()
    import time
    from .native import World

    cmds = model.commands(settle_steps, control_steps)
    base = _settled_world(model, cfg, settle_steps)
    q0, v0, s0 = base.state()
    h0, p0, c0 = base.manifolds()
    base.close()
    runs = {}
    for name, net in (("plain", None), ("mlp", mlp)):
        w = World(model, cfg)
        w.set_state(q0, v0, s0)
        w.write_manifolds(h0, p0, c0)
        if net is not None:
            w.set_warm_start_mlp(net)
        sweeps = np.zeros(model.environment_count)
        fallbacks = np.zeros(model.environment_count)
        worst = np.zeros(model.environment_count)
        ms = []
        for t in range(control_steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            w.advance(cmds[t:t + 1])
            ms.append((time.perf_counter() - t0) * 1e3)
            cs = w.contact_statuses()[-1]
            sweeps += cs["residuals"][:, 3]
            fallbacks += cs["diagnostics"][:, 2]
            worst = np.maximum(worst, np.nan_to_num(cs["residuals"][:, 1], nan=np.inf))
        runs[name] = (sweeps, fallbacks, worst, float(np.median(ms)))
        w.close()
    healthy = runs["plain"][2] <= healthy_residual
    out = {"worlds": int(model.environment_count), "healthy_worlds": int(healthy.sum()),
           "control_steps": int(control_steps)}
    for name, (sweeps, fallbacks, worst, ms) in runs.items():
        out[name] = {
            "sweeps_per_step": float(sweeps[healthy].mean() / control_steps),
            "sweeps_per_step_all": float(sweeps.mean() / control_steps),
            "fallbacks_per_step": float(fallbacks[healthy].mean() / control_steps),
            "normal_residual_max_healthy": float(worst[healthy].max()) if healthy.any() else 0.0,
            # worlds healthy under the plain warm start whose run here leaves that bound (a chaotic
            # world can take another trajectory from a different, equally converged warm start)
            "healthy_worlds_leaving_bound": int((worst[healthy] > healthy_residual).sum()),
            "ms_per_step": ms,
        }
    return out
