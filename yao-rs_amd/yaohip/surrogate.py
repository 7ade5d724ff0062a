"""Learned energy surrogate for VQE parameter screening (yh_mlp_* over the C ABI).

An MLP (in_dim -> hidden x [width, GELU] -> 1; BF16 hidden weights and activations,
FP32 accumulation, FP32 master weights with Adam) is trained on the GPU, by the
library's own forward and backward kernels, on (angles, energy) pairs that the
batched simulator generates. It then ranks candidate angle vectors; the best
fraction is re-evaluated exactly by the simulator, which also judges the
surrogate's error. The surrogate has no counterpart in the reference.
"""
import ctypes
import math
import time

import numpy as np

from . import _abi as A
from .vqe import VqeBatch, fixed_instances

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def best_k(values, k):
    """The indices of the k smallest values, in the order of a stable sort (ties by
    index): np.argsort(values, kind="stable")[:k] without sorting them all."""
    values = np.asarray(values)
    if k >= values.size:
        return np.argsort(values, kind="stable")
    kth = np.partition(values, k - 1)[k - 1]
    sel = np.flatnonzero(values <= kth)  # ascending indices: ties keep index order
    return sel[np.argsort(values[sel], kind="stable")][:k]


class Mlp:
    """yh_mlp handle. encoding "linear" feeds the inputs as features (in_dim % 16 == 0);
    "trig" feeds (cos x_j, sin x_j) pairs (in_dim % 8 == 0). width % 64 == 0,
    max_batch % 64 == 0."""

    def __init__(self, in_dim, width=256, hidden=4, max_batch=4096, seed=0, encoding="linear"):
        A.init()
        self.in_dim, self.width, self.hidden, self.max_batch = in_dim, width, hidden, max_batch
        self.trig = encoding == "trig"
        enc = {"linear": A.MLP_LINEAR, "trig": A.MLP_TRIG}[encoding]
        self._h = ctypes.c_void_p()
        A.check(A.lib().yh_mlp_create(in_dim, enc, width, hidden, max_batch, seed, ctypes.byref(self._h)))
        n = ctypes.c_size_t()
        A.check(A.lib().yh_mlp_num_params(self._h, ctypes.byref(n)))
        self.num_params = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and A._lib is not None:
            A._lib.yh_mlp_destroy(h)
            self._h = None

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        A.check(A.lib().yh_mlp_get_params(self._h, p.ctypes.data_as(_fp), p.size))
        return p

    def set_params(self, p):
        p = _f32(p)
        A.check(A.lib().yh_mlp_set_params(self._h, p.ctypes.data_as(_fp), p.size))

    def predict(self, x):
        """FP64 rows go to the library as they are (it rounds them to FP32 while staging
        them), anything else as FP32."""
        if isinstance(x, np.ndarray) and x.dtype == np.float64:
            x = np.ascontiguousarray(x).reshape(-1, self.in_dim)
            y = np.zeros(x.shape[0], dtype=np.float32)
            A.check(A.lib().yh_mlp_predict_f64(self._h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               x.shape[0], y.ctypes.data_as(_fp)))
            return y
        x = _f32(x).reshape(-1, self.in_dim)
        y = np.zeros(x.shape[0], dtype=np.float32)
        A.check(A.lib().yh_mlp_predict(self._h, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, x, y):
        """Mean squared error and its gradient over one batch (<= max_batch rows)."""
        x = _f32(x).reshape(-1, self.in_dim)
        y = _f32(y).reshape(-1)
        loss = ctypes.c_float()
        g = np.zeros(self.num_params, dtype=np.float32)
        A.check(A.lib().yh_mlp_loss_grad(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp), y.size,
                                         ctypes.byref(loss), g.ctypes.data_as(_fp)))
        return float(loss.value), g

    def train(self, x, y, batch=1024, epochs=1, lr=1e-3, seed=0):
        """Adam over shuffled mini-batches; returns the mean batch loss of each epoch."""
        x = _f32(x).reshape(-1, self.in_dim)
        y = _f32(y).reshape(-1)
        losses = np.zeros(epochs, dtype=np.float32)
        A.check(A.lib().yh_mlp_train(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp), y.size, batch, epochs,
                                     lr, seed, losses.ctypes.data_as(_fp)))
        return losses


class SurrogateScreen:
    """Trains an Mlp on simulator energies of one ansatz shape (n qubits, depth,
    apq angles per qubit per layer) and screens candidate angle vectors. The
    network reads the angles through the trig encoding: each rotation contributes
    terms of first degree in (cos, sin) of its angle to the state's amplitudes,
    so the energy is a polynomial in these features."""

    def __init__(self, n, depth, apq=3, width=256, hidden=4, max_batch=4096, seed=0, J=1.0, h=1.0,
                 sim_dtype="c64"):
        self.n, self.depth, self.apq, self.J, self.h = n, depth, apq, J, h
        self.n_angles = n * depth * apq
        self.in_dim = self.n_angles
        self.mlp = Mlp(self.in_dim, width, hidden, max_batch, seed, encoding="trig")
        self.sim_dtype = sim_dtype
        self._sims = {}
        self.mu, self.sigma = 0.0, 1.0

    def sim(self, dtype):
        if dtype not in self._sims:
            self._sims[dtype] = VqeBatch(self.chunk, self.n, dtype)
        return self._sims[dtype]

    chunk = 65536

    def exact_energies(self, angles, dtype="c128"):
        """Batched simulator energies of angle rows (the judge of the surrogate)."""
        angles = np.asarray(angles, dtype=np.float64)
        out = np.zeros(angles.shape[0])
        s = self.sim(dtype)
        for i in range(0, angles.shape[0], self.chunk):
            a = angles[i:i + self.chunk]
            out[i:i + a.shape[0]] = s.energies(fixed_instances(a, self.n, self.depth, self.apq), self.J, self.h)
        return out

    def sample_angles(self, count, seed, lo=-math.pi / 4, hi=math.pi / 4):
        return np.random.default_rng(seed).uniform(lo, hi, size=(count, self.n_angles))

    def fit(self, angles, energies, epochs=4, batch=1024, lr=1e-3, seed=0, schedule=None):
        """Standardises the targets and trains; returns the per-epoch losses (standardised units).
        schedule: optional [(epochs, lr), ...] phases run in order (Adam's state carries
        over, so a later phase is a learning-rate drop); it replaces epochs and lr."""
        e = np.asarray(energies, dtype=np.float64)
        self.mu, self.sigma = float(e.mean()), float(e.std() or 1.0)
        y = (e - self.mu) / self.sigma
        x = _f32(angles)
        phases = schedule if schedule is not None else [(epochs, lr)]
        losses = [self.mlp.train(x, y, batch, ep, rate, seed + k) for k, (ep, rate) in enumerate(phases)]
        return np.concatenate(losses)

    def predict(self, angles):
        return self.mlp.predict(np.asarray(angles)).astype(np.float64) * self.sigma + self.mu

    def screen(self, candidates, keep_frac=0.01, dtype="c128"):
        """Ranks candidates by predicted energy, re-evaluates the best keep_frac exactly.
        Returns (kept indices, exact energies of the kept, predicted energies of all, timings)."""
        t0 = time.perf_counter()
        pred = self.predict(candidates)
        t1 = time.perf_counter()
        k = max(1, int(round(keep_frac * len(candidates))))
        idx = best_k(pred, k)
        exact = self.exact_energies(np.asarray(candidates)[idx], dtype)
        t2 = time.perf_counter()
        return idx, exact, pred, {"predict_s": t1 - t0, "reeval_s": t2 - t1}
