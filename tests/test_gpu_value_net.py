"""Value-net kernels on the GPU (BF16 MFMA GEMM, GELU, masked Huber, AdamW)
against the numpy oracle (oracle/mlp_oracle.py, itself pinned by finite
differences and torch autograd in test_mlp_oracle.py).

Tolerances: operands are BF16 (8-bit mantissa) with FP32 accumulation, so
the forward pass is compared with the oracle evaluated on the same
BF16-rounded weights and inputs (relative 2e-2 of the largest output), and
gradients to 5e-2 of the largest gradient.
"""
import os
import sys

import numpy as np
import pytest

from poker_solver_rust_amd.value_net import DEFAULT_DIMS, ValueNet

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_weights(p, dims):
    out, off = [], 0
    for i in range(len(dims) - 1):
        nw = dims[i] * dims[i + 1]
        out.append(bf16(p[off:off + nw]))
        out.append(p[off + nw:off + nw + dims[i + 1]])
        off += nw + dims[i + 1]
    return np.concatenate(out)


@pytest.mark.parametrize("n_in,n_out,n", [(96, 80, 70), (99, 1500, 4096), (101, 2600, 4096)])
def test_mfma_tile_layout_single_layer(n_in, n_out, n):
    """One linear layer (no activation) with exactly representable BF16
    operands: small integers, so every product and sum is exact in FP32 and
    the device result must equal x @ W + b bit for bit -- checks the MFMA
    operand / accumulator lane layout, the tile edges and the K padding
    (input widths not a multiple of 8).  The shapes select the 64 x 64,
    128 x 64 and 128 x 128 tiles (128-row tiles only where they still give
    600 or more workgroups)."""
    net = ValueNet((n_in, n_out), max_batch=n, seed=1)
    rng = np.random.default_rng(0)
    W = rng.integers(-3, 4, (n_in, n_out)).astype(np.float32)
    b = rng.integers(-3, 4, n_out).astype(np.float32)
    net.set_params(np.concatenate([W.ravel(), b]))
    x = rng.integers(-3, 4, (n, n_in)).astype(np.float32)
    assert np.array_equal(net.forward(x), x @ W + b)


def test_weight_gradient_split_k_exact():
    """Gradients of one linear layer on 4096 rows, summed over K = 4096 in 8
    slices: with integer inputs and targets one to three units off the
    prediction, the Huber (delta 1) gradient is -sign(e) / (rows * outputs)
    = +-2^-22 or 0 (exact in BF16) and every partial sum is an exact FP32
    multiple of 2^-22, so dW = x^T dZ and db = sum_rows dZ are bit-exact."""
    n_in, n_out, n = 96, 1024, 4096
    net = ValueNet((n_in, n_out), max_batch=n, seed=2)
    rng = np.random.default_rng(5)
    W = rng.integers(-3, 4, (n_in, n_out)).astype(np.float32)
    b = rng.integers(-3, 4, n_out).astype(np.float32)
    net.set_params(np.concatenate([W.ravel(), b]))
    x = rng.integers(-3, 4, (n, n_in)).astype(np.float32)
    e = rng.integers(-3, 4, (n, n_out)).astype(np.float32)
    y = x @ W + b + e
    _, g = net.gradients(x, y, np.ones_like(y), 1.0)
    dz = np.clip(-e, -1, 1).astype(np.float64) * 2.0 ** -22
    assert np.array_equal(g[:n_in * n_out].reshape(n_in, n_out), (x.astype(np.float64).T @ dz).astype(np.float32))
    assert np.array_equal(g[n_in * n_out:], dz.sum(0).astype(np.float32))


@pytest.mark.parametrize("n", [2048])
def test_gradients_match_oracle_full_width(n):
    """The 2706 -> 4 x 512 -> 2652 net on 2048 rows: 128 x 128 tiles on the
    output layer, K split in four for the hidden weight gradients."""
    dims = DEFAULT_DIMS
    net = ValueNet(dims, max_batch=n, seed=7)
    p = net.get_params()
    rng = np.random.default_rng(8)
    x = rng.normal(0, 1, (n, dims[0])).astype(np.float32)
    y = rng.normal(0, 0.3, (n, dims[-1])).astype(np.float32)
    mask = (rng.random((n, dims[-1])) < 0.8).astype(np.float32)
    loss, g = net.gradients(x, y, mask, 1.0)
    lo, go = M.loss_and_grads(p.astype(np.float64), dims, x.astype(np.float64), y.astype(np.float64),
                              mask.astype(np.float64), 1.0)
    assert abs(loss - lo) <= 2e-2 * abs(lo)
    assert np.abs(g - go).max() <= 5e-2 * np.abs(go).max()


@pytest.mark.parametrize("dims,n", [((70, 40, 33, 19), 37), ((70, 40, 33, 19), 1), (DEFAULT_DIMS, 48)])
def test_forward_matches_oracle(dims, n):
    net = ValueNet(dims, max_batch=64, seed=3)
    p = net.get_params()
    x = np.random.default_rng(1).normal(0, 1, (n, dims[0])).astype(np.float32)
    got = net.forward(x)
    ref = M.forward(bf16_weights(p, dims).astype(np.float64), dims, bf16(x).astype(np.float64))
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-9)
    assert err <= 2e-2, err


@pytest.mark.parametrize("dims,n", [((70, 40, 33, 19), 37), ((70, 40, 33, 19), 1), ((70, 40, 33, 19), 2),
                                    (DEFAULT_DIMS, 64)])
def test_gradients_match_oracle(dims, n):
    net = ValueNet(dims, max_batch=64, seed=4)
    p = net.get_params()
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (n, dims[0])).astype(np.float32)
    y = rng.normal(0, 1, (n, dims[-1])).astype(np.float32)
    mask = (rng.random((n, dims[-1])) < 0.8).astype(np.float32)
    loss, g = net.gradients(x, y, mask, 1.0)
    lo, go = M.loss_and_grads(p.astype(np.float64), dims, x.astype(np.float64), y.astype(np.float64),
                              mask.astype(np.float64), 1.0)
    assert abs(loss - lo) <= 2e-2 * abs(lo)
    assert np.abs(g - go).max() <= 5e-2 * np.abs(go).max()


def test_adamw_step_matches_oracle():
    dims = (70, 40, 33, 19)
    net = ValueNet(dims, max_batch=64, seed=5)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (32, dims[0])).astype(np.float32)
    y = rng.normal(0, 1, (32, dims[-1])).astype(np.float32)
    mask = np.ones((32, dims[-1]), np.float32)
    net.set_dataset(x, y, mask)
    net.train(1, 32, lr=1e-2, weight_decay=0.01, delta=1.0, seed=0)
    _, go = M.loss_and_grads(p, dims, x.astype(np.float64), y.astype(np.float64), mask.astype(np.float64), 1.0)
    ref = M.AdamW(dims, 1e-2, 0.01).step(p, go)
    big = np.abs(go) > 1e-2 * np.abs(go).max()
    assert np.abs(net.get_params() - ref)[big].max() <= 1e-3


def test_clipped_adamw_step_matches_oracle():
    """Gradient-norm clipping: max_norm puts the median clipped gradient at
    AdamW's eps, so the step's size depends on the clip coefficient.  The
    device step matches the oracle's clipped step and misses its unclipped
    one; the reported norm is the oracle's gradient norm (BF16 tolerance)."""
    dims = (70, 40, 33, 19)
    net = ValueNet(dims, max_batch=64, seed=6)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(6)
    x = rng.normal(0, 1, (32, dims[0])).astype(np.float32)
    y = rng.normal(0, 1, (32, dims[-1])).astype(np.float32)
    mask = (rng.random((32, dims[-1])) < 0.8).astype(np.float32)
    _, go = M.loss_and_grads(p, dims, x.astype(np.float64), y.astype(np.float64), mask.astype(np.float64), 1.0)
    max_norm = 1e-8 * np.linalg.norm(go) / np.median(np.abs(go))
    gc, norm = M.clip_grad_norm(go, max_norm)
    net.set_dataset(x, y, mask)
    net.set_grad_clip(max_norm)
    net.train(1, 32, lr=1e-2, weight_decay=0.01, delta=1.0, seed=0)
    d = net.get_params().astype(np.float64)
    big = np.abs(go) > 1e-2 * np.abs(go).max()
    assert np.abs(d - M.AdamW(dims, 1e-2, 0.01).step(p, gc))[big].max() <= 1e-3
    assert np.abs(d - M.AdamW(dims, 1e-2, 0.01).step(p, go))[big].max() >= 3e-3
    assert abs(net.last_grad_norm() - norm) <= 2e-2 * norm


def test_training_reduces_loss():
    """A learnable target (a fixed random two-layer function of the input):
    the loss after training is a small fraction of the initial loss."""
    dims = (64, 128, 128, 32)
    rng = np.random.default_rng(4)
    A = rng.normal(0, 0.3, (64, 32))
    x = rng.normal(0, 1, (4096, 64)).astype(np.float32)
    y = np.tanh(x @ A).astype(np.float32)
    net = ValueNet(dims, max_batch=256, seed=6)
    net.set_dataset(x, y, np.ones_like(y))
    losses = net.train(30, 256, lr=1e-3, weight_decay=1e-4, delta=1.0, seed=1)
    assert losses[-1] < 0.25 * losses[0], losses
    assert np.all(np.diff(losses) < 0), losses
