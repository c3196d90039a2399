"""Pins of the warm-start network's CPU restatement (oracle/wsnet_oracle.py)
against torch's conv3d (circular padding, dilation) and autograd on the CPU,
plus the Adam update against torch.optim.Adam."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import wsnet_oracle as W


def _torch_net(weights, biases, dils, x):
    a = torch.tensor(x[None], dtype=torch.float64)
    Ws = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w in weights]
    bs = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for b in biases]
    for li, (w, b, d) in enumerate(zip(Ws, bs, dils)):
        pad = d * (w.shape[2] // 2)
        a = F.conv3d(F.pad(a, (pad,) * 6, mode="circular") if pad else a, w, b, dilation=d)
        if li < len(Ws) - 1:
            a = F.leaky_relu(a, 0.1)
    return a[0], Ws, bs


@pytest.fixture(scope="module")
def small_net():
    rng = np.random.default_rng(0)
    chans = [3, 5, 4, 6, 2]
    ks = [1, 3, 3, 1]
    dils = [1, 2, 3, 1]
    weights = [rng.normal(0, 0.3, (chans[i + 1], chans[i], k, k, k)) for i, k in enumerate(ks)]
    biases = [rng.normal(0, 0.1, chans[i + 1]) for i in range(4)]
    x = rng.normal(size=(3, 7, 6, 9))
    t = rng.normal(size=(2, 7, 6, 9))
    mask = (rng.uniform(size=(1, 7, 6, 9)) > 0.3).astype(np.float64)
    return weights, biases, dils, x, t, mask


def test_forward_matches_torch(small_net):
    weights, biases, dils, x, _, _ = small_net
    y = W.Net(weights, biases, dils).forward(x)[-1]
    yt, _, _ = _torch_net(weights, biases, dils, x)
    assert np.abs(y - yt.detach().numpy()).max() < 1e-12


def test_gradients_match_autograd(small_net):
    weights, biases, dils, x, t, mask = small_net
    loss, gW, gb, gx = W.Net(weights, biases, dils).loss_and_grads(x, t, mask)
    yt, Ws, bs = _torch_net(weights, biases, dils, x)
    m = torch.tensor(np.broadcast_to(mask, t.shape).copy())
    lt = (m * (yt - torch.tensor(t)) ** 2).sum() / m.sum()
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12
    for a, b in zip(gW, Ws):
        assert np.abs(a - b.grad.numpy()).max() < 1e-12
    for a, b in zip(gb, bs):
        assert np.abs(a - b.grad.numpy()).max() < 1e-12


def test_input_gradient_by_finite_differences(small_net):
    weights, biases, dils, x, t, mask = small_net
    net = W.Net(weights, biases, dils)
    _, _, _, gx = net.loss_and_grads(x, t, mask)
    rng = np.random.default_rng(1)
    for _ in range(5):
        idx = tuple(rng.integers(0, s) for s in x.shape)
        h = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (net.loss_and_grads(xp, t, mask)[0] - net.loss_and_grads(xm, t, mask)[0]) / (2 * h)
        assert abs(fd - gx[idx]) < 1e-6 * max(1.0, abs(fd))


def test_adam_matches_torch():
    rng = np.random.default_rng(2)
    p = rng.normal(size=20)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2)
    m, v = np.zeros(20), np.zeros(20)
    for step in range(1, 6):
        g = rng.normal(size=20)
        W.adam_step(p, g, m, v, step, 1e-2)
        pt.grad = torch.tensor(g)
        opt.step()
    assert np.abs(p - pt.detach().numpy()).max() < 1e-12
