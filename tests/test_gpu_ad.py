"""expect_grad on the GPU (yh_expect_grad) against the oracle's adjoint sweep
(pinned by the reference's own AD tests and finite differences). Bars: value and
gradient within 1e-10 (absolute, unit-norm states, O(1) observables) in
complex128; complex64 within 1e-4."""
import math

import numpy as np
import pytest

from conftest import seeded_state
from oracle import oracle as O
from test_ad_oracle import random_param_circuit, random_poly
from yaohip.circuit import Circuit, Gate, NoiseChannel, Op, OperatorPolynomial, channel, control, lower, put, tfim_hamiltonian

pytestmark = pytest.mark.gpu


def both(gpu, n, els, h, psi0, dtype="c128"):
    c = Circuit.qubits(n, els)
    reg = gpu.ArrayReg.from_vec(n, psi0, dtype)
    v, g = gpu.expect_grad(h, c, reg)
    wv, wg = O.expect_grad(n, psi0, lower(els), h, c.num_params())
    return v, g, wv, wg, reg


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (5, 3), (8, 4), (11, 5)])
def test_expect_grad_random_circuits_vs_oracle(gpu, n, seed):
    els = random_param_circuit(n, 40, 100 + seed)
    h = random_poly(n, seed)
    psi0 = seeded_state(n, seed)
    print(f"[case] expect_grad n={n} seed={seed}", flush=True)
    v, g, wv, wg, reg = both(gpu, n, els, h, psi0)
    assert abs(v - wv) < 1e-10
    assert g.shape == wg.shape and np.max(np.abs(g - wg), initial=0.0) < 1e-10
    # the input register is left unchanged
    assert np.array_equal(reg.state_vec(), A_roundtrip(psi0))


def A_roundtrip(psi):
    return np.ascontiguousarray(psi, dtype=np.complex128)


def test_expect_grad_reference_cases(gpu):
    for theta in (0.0, 0.3, 0.7, 1.57, 2.0, -0.5):
        c = Circuit.qubits(1, [put([0], Gate.Rx(theta))])
        v, g = gpu.expect_grad(OperatorPolynomial.single(0, Op.Z, 1.0), c, gpu.ArrayReg.zero_state(1))
        assert abs(v - math.cos(theta)) < 1e-10 and abs(g[0] + math.sin(theta)) < 1e-10
    els = [put([0], Gate.H), put([1], Gate.H), put([0], Gate.Rx(0.31)), put([1], Gate.Ry(-0.47)),
           control([0], [1], Gate.Rz(0.22)), put([0, 1], Gate.FSim(0.15, 0.37)), control([1], [0], Gate.Phase(-0.63))]
    h = OperatorPolynomial.single(0, Op.Z, 1.0) + OperatorPolynomial.single(1, Op.X, 0.5)
    z = np.zeros(4, dtype=complex)
    z[0] = 1
    v, g, wv, wg, _ = both(gpu, 2, els, h, z)
    assert abs(v - wv) < 1e-12 and np.max(np.abs(g - wg)) < 1e-12


def test_expect_grad_vqe_ansatz_16q(gpu):
    """A 16-qubit hardware-efficient ansatz (Ry, Rz per qubit per layer + CZ chain,
    3 layers) on the TFIM Hamiltonian."""
    n = 16
    rng = np.random.default_rng(7)
    els = []
    for _ in range(3):
        for q in range(n):
            els += [put([q], Gate.Ry(rng.uniform(-math.pi, math.pi))), put([q], Gate.Rz(rng.uniform(-math.pi, math.pi)))]
        els += [control([q], [q + 1], Gate.Z) for q in range(n - 1)]
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    v, g, wv, wg, _ = both(gpu, n, els, tfim_hamiltonian(n), z)
    assert abs(v - wv) < 1e-10 and np.max(np.abs(g - wg)) < 1e-10


def test_expect_grad_complex64(gpu):
    n = 8
    els = random_param_circuit(n, 30, 77)
    h = random_poly(n, 77)
    psi0 = seeded_state(n, 77)
    v, g, wv, wg, _ = both(gpu, n, els, h, psi0, "c64")
    print(f"c64 expect_grad: |dv| {abs(v - wv):.2e}, max |dg| {np.max(np.abs(g - wg)):.2e}")
    assert abs(v - wv) < 1e-4 and np.max(np.abs(g - wg)) < 1e-4


def test_expect_grad_errors(gpu):
    import yaohip
    c = Circuit.qubits(1, [put([0], Gate.Rx(0.3)), channel([0], NoiseChannel.BitFlip(0.01))])
    with pytest.raises(yaohip.YhError):
        gpu.expect_grad(OperatorPolynomial.single(0, Op.Z, 1.0), c, gpu.ArrayReg.zero_state(1))
    c = Circuit.qubits(2, [put([0], Gate.Rx(0.3))])
    with pytest.raises(yaohip.YhError):
        gpu.expect_grad(OperatorPolynomial.single(0, Op.Z, 1.0), c, gpu.ArrayReg.zero_state(3))
    # no parameters: value only
    c = Circuit.qubits(2, [put([0], Gate.H)])
    v, g = gpu.expect_grad(OperatorPolynomial.single(0, Op.X, 1.0), c, gpu.ArrayReg.zero_state(2))
    assert abs(v - 1.0) < 1e-12 and g.size == 0
