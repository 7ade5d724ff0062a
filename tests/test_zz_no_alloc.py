"""yh_expect_noalloc (no reference counterpart: expect_arrayreg, src/expect.rs:7-11,
for callers that cannot allocate device memory) against yh_expect on the same
device state: two calls in a row, the second on a smaller register with fewer
terms, complex128 and complex64, and the refusal beyond the static scratch's 256
terms. Both paths run the same reduction kernels in the same order, so the bar is
1e-12 absolute (the values are O(1)); the complex128 result is also checked
against the oracle."""
import ctypes

import numpy as np
import pytest

from conftest import seeded_state
from oracle import oracle as O
from yaohip import _abi as A
from yaohip.circuit import OperatorPolynomial, OperatorString, tfim_hamiltonian

pytestmark = pytest.mark.gpu

def _fn():
    return A.lib().yh_expect_noalloc


def noalloc(reg, poly):
    T, coeffs, tl, sites, ops = poly.flat()
    ptr = ctypes.c_void_p()
    A.check(A.lib().yh_reg_device_ptr(reg._h, ctypes.byref(ptr)))
    out = np.zeros(2)
    st = _fn()(ptr, reg.nqubits(), A.C128 if reg.dtype == "c128" else A.C64, T, A.dptr(coeffs), A.i32ptr(tl),
               A.i32ptr(sites), A.i32ptr(ops), None, A.dptr(out))
    A.check(st)
    return complex(out[0], out[1])


def random_poly(n, nterms, seed):
    rng = np.random.default_rng(seed)
    coeffs, strings = [], []
    for _ in range(nterms):
        k = int(rng.integers(1, min(n, 4) + 1))
        locs = rng.choice(n, size=k, replace=False)
        strings.append(OperatorString([(int(l), int(rng.integers(1, 8))) for l in locs]))
        coeffs.append(complex(rng.standard_normal(), rng.standard_normal()))
    return OperatorPolynomial(coeffs, strings)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_matches_yh_expect_twice(gpu, dtype):
    for n, nterms, seed in ((16, 37, 1), (9, 5, 2)):  # the second call is smaller
        print(f"[case] expect_noalloc dtype={dtype} n={n} terms={nterms}", flush=True)
        psi = seeded_state(n, seed)
        reg = gpu.ArrayReg.from_vec(n, psi, dtype)
        for poly in (random_poly(n, nterms, seed), tfim_hamiltonian(n, 1.0, 0.7)):
            want = gpu.expect_arrayreg(reg, poly)
            got = noalloc(reg, poly)
            assert abs(got - want) < 1e-12, (n, got, want)
            if dtype == "c128":
                assert abs(got - O.expect(n, psi, poly)) < 1e-12


def test_noalloc_refuses_beyond_scratch(gpu):
    n = 6
    reg = gpu.ArrayReg.from_vec(n, seeded_state(n, 3))
    with pytest.raises(A.YhError) as e:
        noalloc(reg, random_poly(n, 257, 4))
    assert e.value.status == 1 and "256" in e.value.message
    # a call at the limit still works after the refusal
    poly = random_poly(n, 256, 5)
    assert abs(noalloc(reg, poly) - gpu.expect_arrayreg(reg, poly)) < 1e-12
