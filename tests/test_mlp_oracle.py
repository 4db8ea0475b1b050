"""Pins for the value-net oracle (oracle/mlp_oracle.py): its gradients
against central finite differences and against torch autograd (CPU, float64)
of the same network, loss and AdamW step."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

DIMS = [7, 5, 5, 4]


def setup(seed=0, n=6):
    rng = np.random.default_rng(seed)
    n_params = sum(DIMS[i] * DIMS[i + 1] + DIMS[i + 1] for i in range(len(DIMS) - 1))
    p = rng.normal(0, 0.6, n_params)
    x = rng.normal(0, 1, (n, DIMS[0]))
    y = rng.normal(0, 1.5, (n, DIMS[-1]))
    mask = (rng.random((n, DIMS[-1])) < 0.8).astype(np.float64)
    return p, x, y, mask


@pytest.mark.parametrize("delta", [0.5, 3.0])
def test_gradients_match_finite_differences(delta):
    p, x, y, mask = setup()
    _, g = M.loss_and_grads(p, DIMS, x, y, mask, delta)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(len(p)):
        q = p.copy()
        q[i] += eps
        lp, _ = M.loss_and_grads(q, DIMS, x, y, mask, delta)
        q[i] -= 2 * eps
        lm, _ = M.loss_and_grads(q, DIMS, x, y, mask, delta)
        num[i] = (lp - lm) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-8)


def test_matches_torch_autograd_and_adamw():
    torch = pytest.importorskip("torch")
    p, x, y, mask = setup(1)
    loss, g = M.loss_and_grads(p, DIMS, x, y, mask, 1.0)
    layers = M.unflatten(p, DIMS)
    tw = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in layers]
    h = torch.tensor(x)
    for i, (W, b) in enumerate(tw):
        z = h @ W + b
        h = torch.nn.functional.gelu(z) if i + 1 < len(tw) else z
    # the reference's masked_huber_loss: masked sum / max(mask sum, 1)
    tm = torch.tensor(mask)
    tl = (tm * torch.nn.functional.huber_loss(h, torch.tensor(y), reduction="none", delta=1.0)).sum() \
        / tm.sum().clamp_min(1.0)
    tl.backward()
    tg = np.concatenate([np.concatenate([W.grad.numpy().ravel(), b.grad.numpy()]) for W, b in tw])
    assert abs(float(tl.detach()) - loss) < 1e-12 and np.allclose(tg, g, rtol=1e-10, atol=1e-14)
    # AdamW: torch on the weights (decay) and biases (no decay)
    opt = torch.optim.AdamW([{"params": [w for w, _ in tw], "weight_decay": 0.01},
                             {"params": [b for _, b in tw], "weight_decay": 0.0}], lr=1e-2)
    opt.step()
    ours = M.AdamW(DIMS, 1e-2, 0.01).step(p, g)
    tp = np.concatenate([np.concatenate([W.detach().numpy().ravel(), b.detach().numpy()]) for W, b in tw])
    assert np.allclose(tp, ours, rtol=1e-12, atol=1e-14)


def test_empty_mask_gives_zero_loss_and_gradient():
    """All outputs masked out: the normaliser clamps at 1 and nothing flows."""
    p, x, y, mask = setup(2)
    loss, g = M.loss_and_grads(p, DIMS, x, y, np.zeros_like(mask), 1.0)
    assert loss == 0.0 and not np.any(g)


def test_loss_is_the_mean_over_valid_outputs():
    """Appending fully masked rows leaves the loss and gradient unchanged:
    the loss is a mean over the valid outputs, not over all rows."""
    p, x, y, mask = setup(3)
    _, xo, yo, _ = setup(4)
    l1, g1 = M.loss_and_grads(p, DIMS, x, y, mask, 1.0)
    l2, g2 = M.loss_and_grads(p, DIMS, np.vstack([x, xo]), np.vstack([y, yo]),
                              np.vstack([mask, np.zeros_like(mask)]), 1.0)
    assert abs(l1 - l2) < 1e-12 and np.allclose(g1, g2, rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("max_norm", [1e-3, 0.05, 100.0])
def test_clip_grad_norm_matches_torch(max_norm):
    """Global-norm clipping against torch.nn.utils.clip_grad_norm_ on the
    network's per-layer gradients (clipping and non-clipping cases)."""
    torch = pytest.importorskip("torch")
    p, x, y, mask = setup(5)
    _, g = M.loss_and_grads(p, DIMS, x, y, mask, 1.0)
    ours, norm = M.clip_grad_norm(g, max_norm)
    ts = [torch.zeros(a.shape, dtype=torch.float64, requires_grad=True) for a in
          [t for W, b in M.unflatten(g, DIMS) for t in (W, b)]]
    for t, a in zip(ts, [t for W, b in M.unflatten(g, DIMS) for t in (W, b)]):
        t.grad = torch.tensor(a)
    tn = torch.nn.utils.clip_grad_norm_(ts, max_norm)
    tg = np.concatenate([t.grad.numpy().ravel() for t in ts])
    assert abs(float(tn) - norm) < 1e-12 * norm
    assert np.allclose(ours, tg, rtol=1e-12, atol=0)
    assert (norm > max_norm) == (not np.allclose(ours, g))
