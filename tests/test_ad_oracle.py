"""Pins the oracle's expect_grad (adjoint differentiation, src/ad.rs:75-156) with
the reference's own known answers (src/unit_tests/ad.rs: closed forms for Rx, Ry,
Rz, Phase; parameter ordering; channel rejection; the finite-difference cross
check on its 2-qubit circuit) and with central differences of the oracle's value
on random circuits of every parametric gate kind, with controls."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from yaohip.circuit import Circuit, Gate, NoiseChannel, Op, OperatorPolynomial, PositionedGate, channel, control, lower, put


def zero(n):
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    return z


def grad_of(n, els, h, psi0=None):
    c = Circuit.qubits(n, els)
    return O.expect_grad(n, zero(n) if psi0 is None else psi0, lower(els), h, c.num_params())


def hz(s):
    return OperatorPolynomial.single(s, Op.Z, 1.0)


def hx(s):
    return OperatorPolynomial.single(s, Op.X, 1.0)


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.7, 1.57, 2.0, -0.5])
def test_rx_on_zero_vs_z(theta):
    v, g = grad_of(1, [put([0], Gate.Rx(theta))], hz(0))
    assert abs(v - math.cos(theta)) < 1e-10 and abs(g[0] + math.sin(theta)) < 1e-10


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.7, 1.57, -0.2])
def test_ry_on_zero_vs_x(theta):
    v, g = grad_of(1, [put([0], Gate.Ry(theta))], hx(0))
    assert abs(v - math.sin(theta)) < 1e-10 and abs(g[0] - math.cos(theta)) < 1e-10


@pytest.mark.parametrize("gate", ["Rz", "Phase"])
@pytest.mark.parametrize("theta", [0.1, 0.9, 1.3, -0.4])
def test_rz_phase_on_plus_vs_x(gate, theta):
    g_ = Gate.Rz(theta) if gate == "Rz" else Gate.Phase(theta)
    v, g = grad_of(1, [put([0], Gate.H), put([0], g_)], hx(0))
    assert abs(v - math.cos(theta)) < 1e-10 and abs(g[0] + math.sin(theta)) < 1e-10


def test_gradient_length_and_order():
    els = [put([0], Gate.Rx(0.1)), put([1], Gate.Ry(0.2)), control([0], [1], Gate.Rz(0.3))]
    c = Circuit.qubits(2, els)
    assert c.parameters() == [0.1, 0.2, 0.3]
    _, g = grad_of(2, els, hz(1))
    assert g.shape == (3,)


def test_rejects_channels():
    els = [put([0], Gate.Rx(0.3)), channel([0], NoiseChannel.BitFlip(0.01))]
    with pytest.raises(O.OracleError):
        O.expect_grad(1, zero(1), lower(els), hz(0), 1)


def _fd(n, els, h, psi0, eps=1e-5):
    c = Circuit.qubits(n, els)
    p = c.parameters()
    out = []
    for i in range(len(p)):
        vals = []
        for s in (+1, -1):
            q = list(p)
            q[i] += s * eps
            c2 = Circuit.qubits(n, els)
            c2.dispatch(q)
            vals.append(O.expect(n, O.apply(n, psi0, lower(c2.elements)), h).real)
        out.append((vals[0] - vals[1]) / (2 * eps))
    return np.array(out)


def test_reference_finite_difference_circuit():
    els = [put([0], Gate.H), put([1], Gate.H), put([0], Gate.Rx(0.31)), put([1], Gate.Ry(-0.47)),
           control([0], [1], Gate.Rz(0.22)), put([0, 1], Gate.FSim(0.15, 0.37)), control([1], [0], Gate.Phase(-0.63))]
    h = OperatorPolynomial.single(0, Op.Z, 1.0) + OperatorPolynomial.single(1, Op.X, 0.5)
    _, g = grad_of(2, els, h)
    assert np.max(np.abs(g - _fd(2, els, h, zero(2)))) < 1e-6


def random_param_circuit(n, count, seed, with_custom=True):
    rng = np.random.default_rng(seed)
    names1 = ["H", "X", "Y", "Z", "S", "T", "SqrtX", "SqrtY", "SqrtW"]
    els = []
    for _ in range(count):
        r = rng.random()
        locs = [int(x) for x in rng.permutation(n)]
        if r < 0.45:
            kind = ["Rx", "Ry", "Rz", "Phase"][rng.integers(4)]
            g = getattr(Gate, kind)(rng.uniform(-math.pi, math.pi))
            nt = 1
        elif r < 0.55 and n >= 2:
            g = Gate.FSim(*rng.uniform(-math.pi, math.pi, size=2))
            nt = 2
        elif r < 0.6 and n >= 2:
            g = [Gate.ISWAP, Gate.SWAP][rng.integers(2)]
            nt = 2
        elif r < 0.7 and with_custom:
            nt = int(rng.integers(1, min(2, n) + 1))
            q, _ = np.linalg.qr(rng.standard_normal((1 << nt, 1 << nt)) + 1j * rng.standard_normal((1 << nt, 1 << nt)))
            g = Gate.Custom(q, False)
        else:
            g = getattr(Gate, names1[rng.integers(len(names1))])
            nt = 1
        nc = int(rng.integers(0, min(2, n - nt) + 1))
        els.append(PositionedGate(g, locs[:nt], locs[nt:nt + nc], [bool(rng.integers(2)) for _ in range(nc)]))
    return els


def random_poly(n, seed):
    rng = np.random.default_rng(seed)
    ops = [Op.X, Op.Y, Op.Z]
    h = OperatorPolynomial.single(0, Op.Z, 0.7)
    for _ in range(4):
        a, b = rng.permutation(n)[:2] if n > 1 else (0, 0)
        term = OperatorPolynomial.single(int(a), ops[rng.integers(3)], complex(rng.normal(), 0))
        h = h + term
    return h


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (4, 3)])
def test_random_circuits_vs_finite_differences(n, seed):
    els = random_param_circuit(n, 14, seed)
    rng = np.random.default_rng(seed)
    psi0 = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi0 /= np.linalg.norm(psi0)
    h = random_poly(n, seed)
    v, g = grad_of(n, els, h, psi0)
    assert abs(v - O.expect(n, O.apply(n, psi0, lower(els)), h).real) < 1e-12
    assert np.max(np.abs(g - _fd(n, els, h, psi0))) < 1e-6
