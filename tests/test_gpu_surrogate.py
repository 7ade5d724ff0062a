"""MLP surrogate on the GPU: forward/backward kernels against the CPU oracle, the
Adam step against its closed form, training, and the screening gain over random
selection of VQE parameters.

Bars: gradients within 2e-3 relative L2 of the oracle with bfloat16 at the GPU's
storage points (FP32 accumulation order is the only other difference); loss and
predictions within 1e-4 relative.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from yaohip.surrogate import Mlp, SurrogateScreen

pytestmark = pytest.mark.gpu


def _layers(in_dim, width, hidden, trig=False):
    out, off = [], 0
    for layer in range(hidden + 1):
        fi = (2 * in_dim if trig else in_dim) if layer == 0 else width
        fo = 1 if layer == hidden else width
        out.append((f"W{layer}", off, off + fo * fi))
        off += fo * fi
        out.append((f"b{layer}", off, off + fo))
        off += fo
    return out


def _check_grads(m, x, y, tag):
    p = m.get_params().astype(np.float64)
    loss, g = m.loss_grad(x, y)
    pred = m.predict(x)
    wout, wloss, wg = O.mlp_loss_grad(m.in_dim, m.width, m.hidden, p, x, y, quant=True, trig=m.trig)
    assert abs(loss - wloss) <= 1e-4 * wloss, (tag, loss, wloss)
    assert np.max(np.abs(pred - wout)) <= 1e-4 * max(1.0, np.max(np.abs(wout))), tag
    worst = 0.0
    for name, a, b in _layers(m.in_dim, m.width, m.hidden, m.trig):
        e = np.linalg.norm(g[a:b] - wg[a:b]) / max(np.linalg.norm(wg[a:b]), 1e-30)
        worst = max(worst, e)
        assert e < 2e-3, (tag, name, e)
    print(f"{tag}: loss {loss:.6e} oracle {wloss:.6e}; worst per-layer grad rel L2 {worst:.2e}")


@pytest.mark.parametrize("shape", [(32, 64, 2, 100), (64, 128, 3, 64), (288, 256, 4, 256)])
def test_mlp_loss_grad_vs_oracle(gpu, shape):
    in_dim, width, hidden, n = shape
    m = Mlp(in_dim, width, hidden, max_batch=512, seed=11)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, size=(n, in_dim)).astype(np.float32)
    y = np.sin(x[:, :4].sum(axis=1)).astype(np.float32)
    _check_grads(m, x, y, f"{shape}")


@pytest.mark.parametrize("shape", [(8, 64, 2, 50), (288, 256, 4, 192)])
def test_mlp_trig_encoding_loss_grad_vs_oracle(gpu, shape):
    in_dim, width, hidden, n = shape
    m = Mlp(in_dim, width, hidden, max_batch=256, seed=12, encoding="trig")
    rng = np.random.default_rng(n + 1)
    x = rng.uniform(-math.pi, math.pi, size=(n, in_dim)).astype(np.float32)
    y = np.cos(x[:, :3].sum(axis=1)).astype(np.float32)
    _check_grads(m, x, y, f"trig {shape}")


def test_mlp_rejects_bad_shapes(gpu):
    import yaohip
    for args in [(20, 64, 2, 64, 0, "linear"), (12, 64, 2, 64, 0, "trig"), (16, 96, 2, 64, 0, "linear"),
                 (16, 64, 0, 64, 0, "linear"), (16, 64, 2, 100, 0, "linear")]:
        with pytest.raises(yaohip.YhError):
            Mlp(*args[:5], encoding=args[5])
    m = Mlp(16, 64, 1, max_batch=64)
    with pytest.raises(yaohip.YhError):
        m.loss_grad(np.zeros((65, 16)), np.zeros(65))
    with pytest.raises(yaohip.YhError):
        m.train(np.zeros((100, 16)), np.zeros(100), batch=96)


def test_mlp_ragged_batches_after_larger(gpu):
    # a 100-row batch then a 70-row one: padded rows of the second must not carry
    # values from the first into the weight gradients
    m = Mlp(32, 64, 2, max_batch=256, seed=5)
    rng = np.random.default_rng(1)
    for n in (100, 70, 129, 64, 1):
        x = rng.uniform(-1, 1, size=(n, 32)).astype(np.float32)
        y = rng.normal(size=n).astype(np.float32)
        _check_grads(m, x, y, f"n={n}")


def test_mlp_predict_chunked_and_params(gpu):
    m = Mlp(16, 64, 1, max_batch=64, seed=3)
    p = m.get_params()
    bound = 1 / math.sqrt(16)
    assert p.size == O.mlp_num_params(16, 64, 1)
    assert np.all(np.abs(p[:64 * 16 + 64]) <= bound)
    assert np.array_equal(Mlp(16, 64, 1, max_batch=64, seed=3).get_params(), p)
    x = np.random.default_rng(2).normal(size=(200, 16)).astype(np.float32)
    want, _, _ = O.mlp_loss_grad(16, 64, 1, p.astype(np.float64), x, np.zeros(200, np.float32))
    got = m.predict(x)
    assert np.max(np.abs(got - want)) < 1e-4 * max(1.0, np.max(np.abs(want)))
    p2 = p * 0.5
    m.set_params(p2)
    assert np.array_equal(m.get_params(), p2)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 1000])
def test_mlp_predict_fp64_rows_match_fp32(gpu, n):
    """yh_mlp_predict_f64 rounds each input to FP32 as it stages the rows: its
    predictions equal those of the same rows rounded by numpy, bit for bit, for row
    counts that end inside a chunk, on a chunk edge and across staging parts; and a
    call after a larger one reuses the staging."""
    m = Mlp(16, 64, 2, max_batch=64, seed=5)
    x64 = np.random.default_rng(n).normal(size=(n, 16))
    got = m.predict(x64)
    want = m.predict(x64.astype(np.float32))
    assert got.dtype == np.float32 and got.shape == (n,)
    assert np.array_equal(got, want)
    p = m.get_params().astype(np.float64)
    ref, _, _ = O.mlp_loss_grad(16, 64, 2, p, x64.astype(np.float32), np.zeros(n, np.float32))
    assert np.max(np.abs(got - ref)) < 1e-4 * max(1.0, np.max(np.abs(ref)))


def test_adam_first_step_closed_form(gpu):
    # first bias-corrected Adam step: p1 = p0 - lr * g / (|g| + eps)
    m = Mlp(32, 64, 2, max_batch=128, seed=9)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(128, 32)).astype(np.float32)
    y = rng.normal(size=128).astype(np.float32)
    p0 = m.get_params().astype(np.float64)
    _, g = m.loss_grad(x, y)
    lr = 1e-3
    m.train(x, y, batch=128, epochs=1, lr=lr, seed=0)
    step = (p0 - m.get_params()) / lr
    big = np.abs(g) > 1e-4
    want = np.sign(g[big]) * np.abs(g[big]) / (np.abs(g[big]) + 1e-8)
    assert big.sum() > 0.5 * g.size
    assert np.max(np.abs(step[big] - want)) < 2e-2


def test_mlp_training_reduces_loss(gpu):
    rng = np.random.default_rng(0)
    n, d = 8192, 32
    x = rng.uniform(-1, 1, size=(n, d)).astype(np.float32)
    y = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2]).astype(np.float32)
    m = Mlp(d, 128, 2, max_batch=1024, seed=1)
    l0, _ = m.loss_grad(x[:1024], y[:1024])
    losses = m.train(x, y, batch=256, epochs=8, lr=2e-3, seed=2)
    l1, _ = m.loss_grad(x[:1024], y[:1024])
    print("epoch losses", losses)
    assert np.all(np.isfinite(losses))
    assert l1 < 0.1 * l0


def test_surrogate_screening_beats_random_selection(gpu):
    """Trained on simulator energies (8 qubits, depth 3, Rz-Ry-Rz blocks, angles
    uniform in [-pi/4, pi/4]), the surrogate's best 1% of 8192 fresh candidates has
    a lower mean exact energy than a random 1% of them."""
    s = SurrogateScreen(n=8, depth=3, apq=3, width=256, hidden=4, max_batch=2048, seed=0)
    train = s.sample_angles(1 << 15, seed=1)
    e = s.exact_energies(train, dtype="c64")
    losses = s.fit(train, e, epochs=10, batch=512, lr=1e-3, seed=3)
    cand = s.sample_angles(8192, seed=2)
    idx, exact, pred, _ = s.screen(cand, keep_frac=0.01, dtype="c128")
    pool = s.exact_energies(cand, dtype="c128")
    assert np.max(np.abs(pool[idx] - exact)) < 1e-10
    mae = float(np.mean(np.abs(pred - pool)))
    rand = np.random.default_rng(5).choice(cand.shape[0], size=idx.size, replace=False)
    print(f"losses {losses}; MAE {mae:.4f} (pool std {pool.std():.4f}); top-1% mean {exact.mean():.4f}, "
          f"random-1% mean {pool[rand].mean():.4f}, pool min {pool.min():.4f}")
    assert mae < 0.8 * float(np.mean(np.abs(pool - pool.mean())))
    assert exact.mean() < pool[rand].mean() - pool.std()
