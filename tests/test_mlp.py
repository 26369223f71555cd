"""Warm-start MLP kernels (include/mr_mlp.h) against the CPU restatement in
oracle/mlp_oracle.py.

CPU tests pin the restatement itself: BF16 rounding known answers, closed-form
GELU values, backward() and the Huber gradient against central finite
differences of the float64 loss, the Huber limits, and the first Adam step's
known magnitude. GPU tests compare each
kernel with the restatement on the same inputs:
  * forward: h1/h2 BF16 bits agree except where the FP32 and float64 GELU
    values straddle a BF16 rounding boundary (<= 0.5 % of elements, one BF16
    ulp), z and y_hat within 2e-3 relative of their scale;
  * loss/dy (squared error and Huber) within 1e-5 relative;
  * backward from the GPU's own trace: every gradient within 2e-3 of its
    scale (BF16 operands, FP32 accumulation in a different order);
  * Adam: 1e-6 relative;
  * determinism (two passes bit-identical) and a short training run that
    lowers the loss on a smooth regression task.
"""

import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import mlp_oracle as mo  # noqa: E402


def random_params(seed, quantize=True):
    rng = np.random.default_rng(seed)
    p = {}
    for k, shape in (("w1", (256, 8)), ("w2", (256, 256)), ("w3", (3, 256))):
        w = (rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])).astype(np.float32)
        p[k] = mo.bf16(w) if quantize else w
    for k, n in (("b1", 256), ("b2", 256), ("b3", 3)):
        p[k] = (0.1 * rng.standard_normal(n)).astype(np.float32)
    p["in_shift"] = rng.standard_normal(8).astype(np.float32) * 0.1
    p["in_scale"] = (1.0 + 0.2 * rng.random(8)).astype(np.float32)
    p["out_scale"] = np.array([2.0, 0.5, 0.5], np.float32)
    return p


# ------------------------------------------------------------------ CPU pins

def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 2.0 - 2.0 ** -8, -1.5, 3.0e-39], np.float32)
    got = mo.bf16(vals)
    # halfway cases go to the even mantissa; the last one carries into the exponent
    assert got[0] == one and got[1] == one
    assert got[2] == np.float32(1.0 + 2.0 ** -6)
    assert got[3] == np.float32(2.0)
    assert got[4] == np.float32(-1.5)
    assert mo.from_bf16_bits(mo.bf16_bits(vals))[2] == got[2]


def test_gelu_closed_forms():
    assert mo.gelu(0.0) == 0.0
    assert abs(mo.gelu(1.0) - 0.8413447460685429) < 1e-15
    assert abs(mo.gelu(-1.0) + 0.15865525393145707) < 1e-15
    assert abs(mo.gelu_grad(0.0) - 0.5) < 1e-15
    x = np.linspace(-4, 4, 41)
    fd = (mo.gelu(x + 1e-6) - mo.gelu(x - 1e-6)) / 2e-6
    np.testing.assert_allclose(mo.gelu_grad(x), fd, rtol=1e-7, atol=1e-8)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(4, quantize=False)
    p = {k: v.astype(np.float64) for k, v in p.items()}
    x = rng.standard_normal((5, 8))
    t = rng.standard_normal((5, 3))

    def loss_of(pp):
        tr = mo.forward(pp, x, quantize=False)
        return mo.mse(tr["y_hat"], t, pp["out_scale"])[0]

    tr = mo.forward(p, x, quantize=False)
    _, dy = mo.mse(tr["y_hat"], t, p["out_scale"])
    g, _, _ = mo.backward(p, tr, dy, quantize=False)
    for name, idx in (("w1", (7, 3)), ("b1", (11,)), ("w2", (5, 200)), ("b2", (17,)), ("w3", (2, 40)),
                      ("b3", (1,))):
        pp = {k: v.copy() for k, v in p.items()}
        pm = {k: v.copy() for k, v in p.items()}
        pp[name][idx] += 1e-6
        pm[name][idx] -= 1e-6
        fd = (loss_of(pp) - loss_of(pm)) / 2e-6
        assert abs(g[name][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, g[name][idx], fd)


def test_huber_loss_gradient_and_limits():
    rng = np.random.default_rng(12)
    y = rng.standard_normal((6, 3)) * 2.0
    t = rng.standard_normal((6, 3))
    s = np.array([1.0, 2.0, 0.5])
    loss, g = mo.mse(y, t, s, huber_delta=0.3)
    for idx in ((0, 0), (3, 1), (5, 2)):
        yp, ym = y.copy(), y.copy()
        yp[idx] += 1e-7
        ym[idx] -= 1e-7
        fd = (mo.mse(yp, t, s, 0.3)[0] - mo.mse(ym, t, s, 0.3)[0]) / 2e-7
        assert abs(g[idx] - fd) <= 1e-6, (idx, g[idx], fd)
    # tiny residuals: Huber is half the squared error; huge ones grow linearly
    l_small, _ = mo.mse(np.full((1, 3), 1e-3), np.zeros((1, 3)), np.ones(3), huber_delta=1.0)
    assert abs(l_small - 0.5e-6) < 1e-15
    l_big, g_big = mo.mse(np.full((1, 3), 100.0), np.zeros((1, 3)), np.ones(3), huber_delta=1.0)
    assert abs(l_big - 99.5) < 1e-12 and np.allclose(g_big, 1.0 / 3.0)


def test_adam_first_step_moves_by_the_learning_rate():
    g = np.array([3.0, -0.02, 1e-3])
    w, m1, m2 = mo.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(w, -0.01 * np.sign(g), rtol=1e-6)


# ------------------------------------------------------------------ GPU

def _upload(p, torch):
    from metalrobo import warmstart as ws

    m = ws.WarmStartMLP()
    m.load_state({**{k: p[k] for k in ("w1", "b1", "w2", "b2", "w3", "b3")}, "in_shift": p["in_shift"],
                  "in_scale": p["in_scale"], "out_scale": p["out_scale"]})
    return m


def _close(got, ref, rel):
    scale = max(float(np.abs(ref).max()), 1e-12)
    err = float(np.abs(np.asarray(got, np.float64) - ref).max())
    assert err <= rel * scale, (err, scale)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(gpu):
    import torch

    p = random_params(1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((100, 8)).astype(np.float32)  # not a multiple of 16
    m = _upload(p, torch)
    xt = torch.from_numpy(x).cuda()
    y = m.forward(xt, trace=True)
    torch.cuda.synchronize()
    b = m._bufs
    ref = mo.forward(p, x)
    for k in ("h1", "h2"):
        got = b[k].float().cpu().numpy()
        mism = np.mean(got != ref[k])
        assert mism <= 5e-3, (k, mism)
        _close(got, ref[k], 1e-2)
    np.testing.assert_array_equal(b["xq"].float().cpu().numpy(), ref["xq"].astype(np.float32))
    _close(b["z1"].cpu().numpy(), ref["z1"], 2e-3)
    _close(b["z2"].cpu().numpy(), ref["z2"], 2e-3)
    _close(y.cpu().numpy(), ref["y_hat"], 2e-3)
    # inference-only pass gives the same bits
    y2 = m.forward(xt)
    assert torch.equal(y2, y)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [300, 1000, 8192])
def test_gpu_loss_and_backward_match_oracle(gpu, n):
    """Rows not a multiple of the 32-row chunks (300, 1000) and the training batch (8192),
    where the weight gradients split the rows over 16 wavefronts per tile."""
    import torch

    p = random_params(5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((n, 8)).astype(np.float32)
    t = rng.standard_normal((n, 3)).astype(np.float32)
    m = _upload(p, torch)
    loss, g = m.loss_and_grads(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    torch.cuda.synchronize()
    b = m._bufs
    # the oracle backward from the GPU's own activations
    tr = {"xq": b["xq"].float().cpu().numpy().astype(np.float64), "z1": b["z1"].cpu().numpy().astype(np.float64),
          "h1": b["h1"].float().cpu().numpy().astype(np.float64), "z2": b["z2"].cpu().numpy().astype(np.float64),
          "h2": b["h2"].float().cpu().numpy().astype(np.float64)}
    y = b["y"].cpu().numpy()
    ref_loss, ref_dy = mo.mse(y, t, p["out_scale"])
    assert abs(float(loss.item()) - ref_loss) <= 1e-5 * ref_loss
    _close(b["dy"].cpu().numpy(), ref_dy, 1e-5)
    ref_g, ref_dz2, ref_dz1 = mo.backward(p, tr, b["dy"].cpu().numpy().astype(np.float64))
    _close(b["dz2"].cpu().numpy(), ref_dz2, 2e-3)
    _close(b["dz1"].cpu().numpy(), ref_dz1, 2e-3)
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        _close(g[k].cpu().numpy(), ref_g[k], 2e-3)


@pytest.mark.gpu
def test_gpu_huber_loss_matches_oracle(gpu):
    import torch

    p = random_params(13)
    rng = np.random.default_rng(14)
    x = rng.standard_normal((200, 8)).astype(np.float32)
    t = (rng.standard_normal((200, 3)) * 3.0).astype(np.float32)
    m = _upload(p, torch)
    loss, _ = m.loss_and_grads(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), huber_delta=0.5)
    torch.cuda.synchronize()
    b = m._bufs
    ref_loss, ref_dy = mo.mse(b["y"].cpu().numpy(), t, p["out_scale"], huber_delta=0.5)
    assert abs(float(loss.item()) - ref_loss) <= 1e-5 * ref_loss
    _close(b["dy"].cpu().numpy(), ref_dy, 1e-5)


@pytest.mark.gpu
def test_gpu_adam_matches_oracle_and_training_is_deterministic(gpu):
    import torch

    p = random_params(8)
    m = _upload(p, torch)
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.standard_normal((256, 8)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.standard_normal((256, 3)).astype(np.float32)).cuda()
    _, g = m.loss_and_grads(x, t)
    grads = {k: v.clone() for k, v in g.items()}
    before = {k: v.cpu().numpy().astype(np.float64) for k, v in m.master.items()}
    m.adam_step(grads, lr=1e-3)
    torch.cuda.synchronize()
    for k in m.master:
        ref, _, _ = mo.adam(before[k], grads[k].cpu().numpy().astype(np.float64), 0.0, 0.0, 1e-3, 0.9, 0.999,
                            1e-8, 1)
        _close(m.master[k].cpu().numpy(), ref, 1e-6)
        if k in m.bf16:
            np.testing.assert_array_equal(m.bf16[k].float().cpu().numpy(), mo.bf16(m.master[k].cpu().numpy()))
    # the same step from the same state repeats bit for bit
    m2 = _upload(p, torch)
    _, g2 = m2.loss_and_grads(x, t)
    for k in grads:
        assert torch.equal(g2[k], grads[k]), k


@pytest.mark.gpu
def test_gpu_training_lowers_the_loss(gpu):
    import torch
    from metalrobo import warmstart as ws

    rng = np.random.default_rng(11)
    x = rng.standard_normal((20000, 8)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0] + 0.5 * x[:, 5]), 0.3 * x[:, 6] * x[:, 2], np.sin(x[:, 7])], 1).astype(np.float32)
    m = ws.WarmStartMLP(seed=1)
    m.set_normalization(x, t)
    rep = m.train(x, t, steps=300, batch=2048, lr=2e-3)
    assert rep.final_loss < 0.1 * rep.first_loss, rep.losses
