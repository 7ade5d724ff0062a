"""Pins the CPU oracle of the MLP surrogate (oracle/mlp_oracle.cpp).

quant=0 (plain FP64) is checked against torch float64 autograd (an independent
forward and backward) and against central finite differences. quant=1 (bfloat16
at the GPU's storage points) is checked against torch's own bfloat16 conversion
placed at those points, with a straight-through forward rounding and a
gradient-rounding hook on the pre-activation products; torch's autograd supplies
the backward.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def _setup(in_dim, width, hidden, n, seed, trig=False):
    rng = np.random.default_rng(seed)
    P = O.mlp_num_params(in_dim, width, hidden, trig)
    p = (rng.normal(size=P) * 0.4).astype(np.float32).astype(np.float64)
    x = rng.normal(size=(n, in_dim)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    return p, x, y


def _bf16(t):
    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _RoundGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t):
        return t.clone()

    @staticmethod
    def backward(ctx, g):
        return _bf16(g)


def _torch_mlp(in_dim, width, hidden, p, x, y, quant, trig=False):
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    q = (lambda t: t + (_bf16(t) - t).detach()) if quant else (lambda t: t)
    xt = torch.tensor(x, dtype=torch.float64)
    if trig:  # features 2j, 2j+1 = cos x_j, sin x_j
        xt = torch.stack([torch.cos(xt), torch.sin(xt)], dim=2).reshape(xt.shape[0], -1)
    h = q(xt)
    f0 = 2 * in_dim if trig else in_dim
    off = 0
    for layer in range(hidden + 1):
        fi = f0 if layer == 0 else width
        fo = 1 if layer == hidden else width
        W = pt[off:off + fo * fi].reshape(fo, fi)
        off += fo * fi
        b = pt[off:off + fo]
        off += fo
        if layer < hidden:
            prod = h @ q(W).T
            z = (_RoundGrad.apply(prod) if quant else prod) + b
            h = q(torch.nn.functional.gelu(z))
        else:
            h = h @ W.T + b
    out = h[:, 0]
    loss = ((out - torch.tensor(y, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    return out.detach().numpy(), loss.item(), pt.grad.numpy()


@pytest.mark.parametrize("shape", [(5, 7, 1, 9), (6, 8, 3, 11), (16, 64, 2, 33)])
def test_mlp_oracle_fp64_vs_autograd(shape):
    in_dim, width, hidden, n = shape
    p, x, y = _setup(in_dim, width, hidden, n, seed=sum(shape))
    out, loss, g = O.mlp_loss_grad(in_dim, width, hidden, p, x, y, quant=False)
    tout, tloss, tg = _torch_mlp(in_dim, width, hidden, p, x, y, quant=False)
    assert np.max(np.abs(out - tout)) < 1e-12
    assert abs(loss - tloss) < 1e-12 * max(1.0, tloss)
    assert np.max(np.abs(g - tg)) < 1e-12


def test_mlp_oracle_fp64_finite_differences():
    in_dim, width, hidden, n = 4, 5, 2, 6
    p, x, y = _setup(in_dim, width, hidden, n, seed=3)
    _, _, g = O.mlp_loss_grad(in_dim, width, hidden, p, x, y, quant=False)
    eps = 1e-6
    fd = np.zeros_like(p)
    for k in range(p.size):
        a, b = p.copy(), p.copy()
        a[k] += eps
        b[k] -= eps
        fd[k] = (O.mlp_loss_grad(in_dim, width, hidden, a, x, y, quant=False)[1]
                 - O.mlp_loss_grad(in_dim, width, hidden, b, x, y, quant=False)[1]) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("shape", [(16, 64, 2, 33), (32, 64, 4, 20)])
def test_mlp_oracle_bf16_points_vs_torch(shape):
    in_dim, width, hidden, n = shape
    p, x, y = _setup(in_dim, width, hidden, n, seed=7 + n)
    out, loss, g = O.mlp_loss_grad(in_dim, width, hidden, p, x, y, quant=True)
    tout, tloss, tg = _torch_mlp(in_dim, width, hidden, p, x, y, quant=True)
    assert np.max(np.abs(out - tout)) < 1e-12
    assert abs(loss - tloss) < 1e-12 * max(1.0, tloss)
    assert np.max(np.abs(g - tg)) < 1e-12
    # and the bf16 network stays close to the fp64 one
    _, loss64, g64 = O.mlp_loss_grad(in_dim, width, hidden, p, x, y, quant=False)
    assert abs(loss - loss64) < 0.05 * loss64
    assert np.linalg.norm(g - g64) < 0.1 * np.linalg.norm(g64)


@pytest.mark.parametrize("quant", [False, True])
def test_mlp_oracle_trig_encoding_vs_torch(quant):
    in_dim, width, hidden, n = 8, 64, 2, 17
    p, x, y = _setup(in_dim, width, hidden, n, seed=21, trig=True)
    x = (x * 2).astype(np.float32)
    out, loss, g = O.mlp_loss_grad(in_dim, width, hidden, p, x, y, quant=quant, trig=True)
    tout, tloss, tg = _torch_mlp(in_dim, width, hidden, p, x, y, quant=quant, trig=True)
    assert np.max(np.abs(out - tout)) < 1e-12
    assert abs(loss - tloss) < 1e-12 * max(1.0, tloss)
    assert np.max(np.abs(g - tg)) < 1e-12


def test_mlp_oracle_gelu_special_values():
    # one hidden unit with weight 1 and bias 0 on one input, output weight 1: out = gelu(x)
    in_dim, width, hidden = 1, 1, 1
    p = np.array([1.0, 0.0, 1.0, 0.0])
    xs = np.array([[0.0], [8.0], [-8.0], [1.0]], dtype=np.float32)
    out, _, _ = O.mlp_loss_grad(in_dim, width, hidden, p, xs, np.zeros(4, np.float32), quant=False)
    assert out[0] == 0.0
    assert abs(out[1] - 8.0) < 1e-12
    assert abs(out[2]) < 1e-12
    # gelu(1) = Phi(1) = 0.8413447460685429...
    assert abs(out[3] - 0.8413447460685429) < 1e-15


def test_mlp_num_params():
    assert O.mlp_num_params(288, 256, 4) == 288 * 256 + 256 + 3 * (256 * 256 + 256) + 257
    assert O.mlp_num_params(288, 256, 4, trig=True) == 576 * 256 + 256 + 3 * (256 * 256 + 256) + 257


def test_best_k_matches_stable_argsort():
    """The screen's selection of the k best predictions equals the first k of a
    stable argsort, ties (many equal values, the k-th among them) included."""
    from yaohip.surrogate import best_k
    rng = np.random.default_rng(0)
    for trial in range(50):
        n = int(rng.integers(1, 400))
        v = rng.integers(0, 6, size=n).astype(np.float64) if trial % 2 else rng.normal(size=n)
        for k in (1, 2, n // 3 + 1, n):
            assert np.array_equal(best_k(v, k), np.argsort(v, kind="stable")[:k])
