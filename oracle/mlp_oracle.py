"""CPU oracle of the value net (test infrastructure only: imported by tests/
and never by the product path).

A plain float64/float32 numpy statement of the network the device trains:
dense layers y = x W + b (W stored in x out), GELU (exact erf form) on every
hidden layer, a linear output layer, the masked Huber loss averaged over the
batch's valid outputs (the reference's masked_huber_loss,
cfvnet/src/model/loss.rs:8-29), and AdamW with decoupled weight
decay (torch.optim.AdamW semantics).  Parameters are flattened as W_1, b_1,
W_2, b_2, ... like psr_mlp_get_params.
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover - scipy is in the image
    _erf = np.vectorize(math.erf)

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def unflatten(params, dims):
    out, off = [], 0
    for i in range(len(dims) - 1):
        nw = dims[i] * dims[i + 1]
        W = params[off:off + nw].reshape(dims[i], dims[i + 1])
        off += nw
        b = params[off:off + dims[i + 1]]
        off += dims[i + 1]
        out.append((W, b))
    assert off == len(params)
    return out


def forward(params, dims, x, keep=False):
    layers = unflatten(params, dims)
    h, zs, hs = x, [], [x]
    for i, (W, b) in enumerate(layers):
        z = h @ W + b
        zs.append(z)
        h = gelu(z) if i + 1 < len(layers) else z
        hs.append(h)
    return (h, zs, hs) if keep else h


def huber(d, delta):
    ad = np.abs(d)
    return np.where(ad <= delta, 0.5 * d * d, delta * (ad - 0.5 * delta))


def huber_grad(d, delta):
    return np.where(np.abs(d) <= delta, d, delta * np.sign(d))


def loss_and_grads(params, dims, x, y, mask, delta):
    """Masked Huber loss averaged over the batch's valid outputs (sum over
    the batch / max(sum of the mask, 1), cfvnet/src/model/loss.rs:8-29) and
    its gradient with respect to every parameter (flattened)."""
    layers = unflatten(params, dims)
    pred, zs, hs = forward(params, dims, x, keep=True)
    scale = 1.0 / max(float(np.asarray(mask, np.float64).sum()), 1.0)
    d = pred - y
    loss = float((mask * huber(d, delta)).sum() * scale)
    g = mask * huber_grad(d, delta) * scale
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        gW = hs[i].T @ g
        gb = g.sum(0)
        grads[i] = (gW, gb)
        if i > 0:
            g = (g @ W.T) * gelu_grad(zs[i - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def clip_grad_norm(grads, max_norm):
    """torch.nn.utils.clip_grad_norm_ on the flattened gradient: scaled by
    min(1, max_norm / (||g||_2 + 1e-6)); returns (clipped, ||g||_2)."""
    norm = float(np.sqrt(np.sum(np.asarray(grads, np.float64) ** 2)))
    return grads * min(1.0, max_norm / (norm + 1e-6)), norm


class AdamW:
    """torch.optim.AdamW: m, v moments, bias correction, decoupled decay
    (weights only, as the device applies it; biases undecayed)."""

    def __init__(self, dims, lr, weight_decay, beta1=0.9, beta2=0.999, eps=1e-8):
        self.dims, self.lr, self.wd = dims, lr, weight_decay
        self.b1, self.b2, self.eps = beta1, beta2, eps
        self.t = 0
        n = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.decay = np.concatenate([np.concatenate([np.ones(dims[i] * dims[i + 1]), np.zeros(dims[i + 1])])
                                     for i in range(len(dims) - 1)])

    def step(self, params, grads):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * grads
        self.v = self.b2 * self.v + (1 - self.b2) * grads * grads
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return params - self.lr * (mh / (np.sqrt(vh) + self.eps) + self.wd * self.decay * params)
