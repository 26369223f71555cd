"""Learned warm start for the cone solve: an 8 -> 256 -> 256 -> 3 MLP with BF16
weights on the gfx950 matrix cores, trained by the library's own kernels
(include/mr_mlp.h). PyTorch tensors only hold device memory and give the
stream; every forward, loss, backward and optimizer step is a call into the
C ABI.

Features per contact, at the start of a microstep's cone solve:
    normal relative velocity, penetration depth, the Euclidean norms of the
    normal and two tangent Jacobian rows, and the previous-step impulse
    (normal, tangent u, tangent v) the plain path warm-starts from.
Target: the impulse the plain solve converges to in that microstep.
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
        self._bufs = {}

    # -------------------------------------------------------------- views
    def weights(self) -> nat.MlpWeightsC:
        return nat.MlpWeightsC(_p(self.bf16["w1"]), _p(self.master["b1"]), _p(self.bf16["w2"]),
                               _p(self.master["b2"]), _p(self.bf16["w3"]), _p(self.master["b3"]),
                               _p(self.in_shift), _p(self.in_scale), _p(self.out_scale))

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
        """Impulses (physical units) predicted for features x."""
        return self.forward(x) * self.out_scale

    def loss_and_grads(self, x, t):
        """Forward with trace, MSE gradient and backward; gradients land in the grads buffers."""
        n = x.shape[0]
        b = self._buffers(n)
        y = self.forward(x, trace=True)
        t = t.contiguous()
        self._b.check(self._b.lib.mr_mlp_loss(_p(y), _p(t), _p(self.out_scale), n, _p(b["dy"]), _p(b["loss"]),
                                              _p(b["scratch"]), _stream()))
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
              seed: int = 0) -> TrainReport:
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
            loss, grads = self.loss_and_grads(f[idx], t[idx])
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
