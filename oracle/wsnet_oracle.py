"""CPU restatement (float64, numpy) of the warm-start network -- test
infrastructure only: tests compare the HIP forward/backward kernels with it.

The reference has no learned component; this restates the network that the
HIP kernels implement (mdft-dev_general_amd/csrc/k_net.hip):

    a_0 = x                                   [C0][nz][ny][nx]
    z_l = conv3d(a_{l-1}; W_l, dilation d_l) + b_l   (k x k x k taps, k in {1, 3}, periodic)
    a_l = leaky_relu(z_l) = max(z_l, 0.1 z_l)   for l < L,   a_L = z_L  (linear output)
    loss = sum(mask * (a_L - t)^2) / (sum(mask) C_L)

Weights W_l[co][ci][tz][ty][tx], tap offsets (t - (k-1)/2) * d_l along each
axis, periodic wrap (the simulation box is periodic).  Pinned in the tests
against torch's conv3d with circular padding and autograd on the CPU.
"""
from __future__ import annotations

import numpy as np

LEAK = 0.1  # slope of the leaky ReLU for negative inputs


def _taps(k):
    h = (k - 1) // 2
    return [(tz, ty, tx) for tz in range(k) for ty in range(k) for tx in range(k)], h


def conv3d(x, w, b, dil):
    """x [Ci][nz][ny][nx], w [Co][Ci][k][k][k] -> [Co][nz][ny][nx] (periodic)."""
    co = w.shape[0]
    out = np.zeros((co,) + x.shape[1:], dtype=np.float64)
    taps, h = _taps(w.shape[2])
    for tz, ty, tx in taps:
        # out[p] += w * x[p + (t-h) d]
        xs = np.roll(x, shift=(-(tz - h) * dil, -(ty - h) * dil, -(tx - h) * dil), axis=(1, 2, 3))
        out += np.einsum("oi,izyx->ozyx", w[:, :, tz, ty, tx], xs)
    return out + b[:, None, None, None]


def conv3d_bwd_data(g, w, dil):
    """dL/dx given dL/dz = g [Co][...]: x[p] feeds z[p - (t-h) d]."""
    ci = w.shape[1]
    out = np.zeros((ci,) + g.shape[1:], dtype=np.float64)
    taps, h = _taps(w.shape[2])
    for tz, ty, tx in taps:
        gs = np.roll(g, shift=((tz - h) * dil, (ty - h) * dil, (tx - h) * dil), axis=(1, 2, 3))
        out += np.einsum("oi,ozyx->izyx", w[:, :, tz, ty, tx], gs)
    return out


def conv3d_bwd_weight(g, x, dil, k=3):
    co, ci = g.shape[0], x.shape[0]
    dw = np.zeros((co, ci, k, k, k))
    taps, h = _taps(k)
    for tz, ty, tx in taps:
        xs = np.roll(x, shift=(-(tz - h) * dil, -(ty - h) * dil, -(tx - h) * dil), axis=(1, 2, 3))
        dw[:, :, tz, ty, tx] = np.einsum("ozyx,izyx->oi", g, xs)
    return dw, g.sum(axis=(1, 2, 3))


class Net:
    def __init__(self, weights, biases, dilations):
        self.W = [np.asarray(w, dtype=np.float64) for w in weights]
        self.b = [np.asarray(b, dtype=np.float64) for b in biases]
        self.d = list(dilations)

    def forward(self, x):
        acts = [np.asarray(x, dtype=np.float64)]
        for li, (w, b, d) in enumerate(zip(self.W, self.b, self.d)):
            z = conv3d(acts[-1], w, b, d)
            acts.append(np.where(z > 0.0, z, LEAK * z) if li < len(self.W) - 1 else z)
        return acts

    def loss_and_grads(self, x, t, mask=None):
        acts = self.forward(x)
        y = acts[-1]
        m = np.ones_like(y) if mask is None else np.broadcast_to(mask, y.shape)
        cnt = m.sum()   # mask broadcast over channels: sum(mask) * C_L
        r = y - t
        loss = float((m * r * r).sum() / cnt)
        g = 2.0 * m * r / cnt
        gW, gb = [None] * len(self.W), [None] * len(self.W)
        for li in range(len(self.W) - 1, -1, -1):
            if li < len(self.W) - 1:
                g = g * np.where(acts[li + 1] > 0.0, 1.0, LEAK)
            gW[li], gb[li] = conv3d_bwd_weight(g, acts[li], self.d[li], self.W[li].shape[2])
            g = conv3d_bwd_data(g, self.W[li], self.d[li])
        return loss, gW, gb, g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected), in place."""
    m *= beta1
    m += (1 - beta1) * g
    v *= beta2
    v += (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    p -= lr * mh / (np.sqrt(vh) + eps)
