"""Register and density-matrix constructors on the device against the oracle
(src/register.rs:28-84, src/density_matrix.rs:37-61, 266-272) and the
reference's own known answers. Bars: 1e-12 relative for amplitudes and entries;
exact where the device computes the same products in the same order."""
import ctypes
import math

import numpy as np
import pytest

from conftest import rel_err, seeded_state
from oracle import oracle as O
from yaohip import _abi as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_product_uniform_ghz(gpu, dtype):
    tol = 1e-15 if dtype == "c128" else 1e-7
    # tests/suites/integration.rs:79,102,125,415: BitStr::<N>::new(k) is basis index k
    for n, k in ((2, 1), (2, 2), (2, 3), (4, 6), (3, 1)):
        s = gpu.ArrayReg.product_state(n, k, dtype).state_vec()
        want = np.zeros(1 << n)
        want[k] = 1
        assert np.array_equal(s, want)
    assert np.array_equal(gpu.ArrayReg.product_state(4, "0110", dtype).state_vec(), np.eye(16)[6])
    with pytest.raises(A.YhError):
        gpu.ArrayReg.product_state(3, 8, dtype)
    u = gpu.ArrayReg.uniform_state(5, dtype).state_vec()
    assert np.max(np.abs(u - 1 / math.sqrt(32))) < tol
    g = gpu.ArrayReg.ghz_state(6, dtype).state_vec()
    want = np.zeros(64)
    want[0] = want[63] = 1 / math.sqrt(2)
    assert np.max(np.abs(g - want)) < tol


@pytest.mark.parametrize("n", [4, 13, 20])
def test_deterministic_state(gpu, n):
    """register.rs:188-194 (norm 1, amplitudes 0 and 1 differ) and the oracle."""
    reg = gpu.ArrayReg.deterministic_state(n)
    s = reg.state_vec()
    assert abs(reg.norm() - 1) < 1e-12 and abs(s[0] - s[1]) > 1e-6
    assert rel_err(s, O.deterministic_state(n)) < 1e-12
    s64 = gpu.ArrayReg.deterministic_state(n, "c64").state_vec()
    assert rel_err(s64, O.deterministic_state(n)) < 1e-6


@pytest.mark.parametrize("n", [3, 16])
def test_rand_state_same_stream(gpu, n):
    """rand_state (register.rs:35-45): with the same stream of f64 uniforms as the
    reference's loop draws, the state equals the oracle's."""
    u = np.random.default_rng(n).random(2 << n)
    reg = gpu.ArrayReg.rand_state(n, lambda k: u[:k])
    assert rel_err(reg.state_vec(), O.rand_state(n, u)) < 1e-12
    # a numpy Generator as the stream
    reg2 = gpu.ArrayReg.rand_state(n, np.random.default_rng(n))
    assert np.array_equal(reg2.state_vec(), reg.state_vec())
    h = ctypes.c_void_p()
    with pytest.raises(A.YhError):  # one uniform short
        A.check(A.lib().yh_reg_rand_state(n, A.C128, A.dptr(u), (2 << n) - 1, ctypes.byref(h)))


def test_mixed(gpu):
    """DensityMatrix::mixed (density_matrix.rs:37-61): the reference's known answers
    (purity 1/2 :304-310, entropy ln 2 :322-330, diagonal 0.25 / 0.75 of
    tests/suites/measure.rs:610-619) and the oracle, bit for bit, on 9 qubits."""
    e0 = gpu.ArrayReg.from_vec(1, [1, 0])
    e1 = gpu.ArrayReg.from_vec(1, [0, 1])
    dm = gpu.DensityMatrix.mixed([0.5, 0.5], [e0, e1])
    assert abs(dm.purity() - 0.5) < 1e-12
    assert abs(dm.von_neumann_entropy() - math.log(2)) < 1e-8
    np.testing.assert_allclose(gpu.probs(gpu.DensityMatrix.mixed([0.25, 0.75], [e0, e1])), [0.25, 0.75], atol=1e-15)
    n = 9
    psis = [seeded_state(n, 40 + j) for j in range(3)]
    w = [0.2, 0.5, 0.3]
    dm = gpu.DensityMatrix.mixed(w, [gpu.ArrayReg.from_vec(n, p) for p in psis])
    assert np.array_equal(dm.state_data(), O.dm_mixed(n, w, psis))
    assert abs(dm.trace() - 1) < 1e-12
    with pytest.raises(A.YhError):
        gpu.DensityMatrix.mixed([0.5, 0.5], [e0, gpu.ArrayReg.from_vec(2, [1, 0, 0, 0])])


@pytest.mark.parametrize("n,locs", [(6, [0, 2, 3]), (10, [9, 1]), (12, [0, 1, 2, 3, 4, 5, 6]), (5, [0, 1, 2, 3, 4])])
def test_density_matrix_from_reg(gpu, n, locs):
    """density_matrix_from_reg (density_matrix.rs:266-272): the reduced matrix computed
    from the amplitudes equals the oracle's from_reg + partial_tr bit for bit (the same
    products, summed in the same order)."""
    psi = seeded_state(n, n)
    reg = gpu.ArrayReg.from_vec(n, psi)
    dm = gpu.density_matrix_from_reg(reg, locs)
    k = len(locs)
    assert dm.nbits() == k
    want = O.density_matrix_from_reg(n, psi, locs)
    assert np.array_equal(dm.state_data().reshape(1 << k, 1 << k), want)
    assert abs(dm.trace() - 1) < 1e-12
    if k == n:
        assert np.array_equal(dm.state_data(), gpu.DensityMatrix.from_reg(reg).state_data())
