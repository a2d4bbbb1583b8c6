"""Pins of the CNN oracle (oracle/cnn_oracle.c) on the CPU.

The learned correction's network has no reference counterpart; the oracle is
pinned against an independent implementation: torch's float64 conv3d (cross-
correlation with zero padding 1) for the forward and torch autograd for every
parameter gradient, plus torch's bfloat16 cast for the rounding function.
"""
import numpy as np
import pytest
import torch

import pyoracle

OFF = dict(W1=0, b1=432, W2=448, b2=7360, W3=7376, b3=7808)


def _params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(pyoracle.CNN_NPARAM)).astype(np.float32)


def _torch_net(p):
    p = torch.tensor(p.astype(np.float64))
    w1 = p[OFF["W1"]:OFF["b1"]].reshape(3, 3, 3, 16).permute(3, 0, 1, 2)[:, None]      # (16,1,3,3,3)
    b1 = p[OFF["b1"]:OFF["W2"]]
    w2 = p[OFF["W2"]:OFF["b2"]].reshape(3, 3, 3, 16, 16).permute(4, 3, 0, 1, 2)        # (co,ci,3,3,3)
    b2 = p[OFF["b2"]:OFF["W3"]]
    w3 = p[OFF["W3"]:OFF["b3"]].reshape(3, 3, 3, 16).permute(3, 0, 1, 2)[None]          # (1,16,3,3,3)
    b3 = p[OFF["b3"]:]
    return [t.clone().requires_grad_(True) for t in (w1, b1, w2, b2, w3, b3)]


def _torch_forward(ws, x):
    F = torch.nn.functional
    w1, b1, w2, b2, w3, b3 = ws
    h = F.relu(F.conv3d(x[:, None], w1, b1, padding=1))
    h = F.relu(F.conv3d(h, w2, b2, padding=1))
    return F.conv3d(h, w3, b3, padding=1)[:, 0]


def _inputs(n=2, seed=1):
    return np.random.default_rng(seed).standard_normal((n, 16, 16, 16)).astype(np.float32)


def test_forward_equals_torch_conv3d():
    c = pyoracle.CnnOracle()
    p, x = _params(), _inputs()
    out, h1, h2 = c.forward(p, x, bf16=False)
    ref = _torch_forward(_torch_net(p), torch.tensor(x.astype(np.float64))).detach().numpy()
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_backward_equals_torch_autograd():
    c = pyoracle.CnnOracle()
    p, x = _params(2), _inputs(2, 3)
    out, h1, h2 = c.forward(p, x, bf16=False)
    gout = np.random.default_rng(4).standard_normal(out.shape)
    g = c.backward(p, x, h1, h2, gout, bf16=False)
    ws = _torch_net(p)
    y = _torch_forward(ws, torch.tensor(x.astype(np.float64)))
    (y * torch.tensor(gout)).sum().backward()
    tw = [w.grad.numpy() for w in ws]
    ref = np.concatenate([tw[0][:, 0].transpose(1, 2, 3, 0).ravel(), tw[1],
                          tw[2].transpose(2, 3, 4, 1, 0).ravel(), tw[3],
                          tw[4][0].transpose(1, 2, 3, 0).ravel(), tw[5]])
    for name, a, b in (("W1", 0, 432), ("b1", 432, 448), ("W2", 448, 7360), ("b2", 7360, 7376),
                       ("W3", 7376, 7808), ("b3", 7808, 7809)):
        assert np.max(np.abs(g[a:b] - ref[a:b])) <= 1e-10 * np.max(np.abs(ref[a:b])), name


def test_round_bf16_matches_torch():
    c = pyoracle.CnnOracle()
    v = np.concatenate([np.random.default_rng(5).standard_normal(2000).astype(np.float32) * 10.0 ** np.arange(-10, 10, 0.01)[:2000].astype(np.float32),
                        np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0], np.float32)])
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(c.round_bf16(v), ref)


def test_bf16_forward_close_to_fp64():
    c = pyoracle.CnnOracle()
    p, x = _params(6), _inputs(1, 7)
    a, _, _ = c.forward(p, x, bf16=False)
    b, h1, _ = c.forward(p, x, bf16=True)
    assert np.max(np.abs(a - b)) < 3e-2 * np.max(np.abs(a))
    # stored activations are exactly representable in bfloat16
    assert np.array_equal(c.round_bf16(h1[0, :64].astype(np.float32)), h1[0, :64].astype(np.float32))
