"""Pins the oracle's partial_tr and von_neumann_entropy (density_matrix.rs:80-122).

partial_tr: against numpy's reshape-and-trace on the row-major matrix, with the
reference's convention that traced locations are bit positions (LSB-first) and
the reduced index keeps the remaining bits in increasing order; closed forms on a
Bell state and on a product state. Entropy: against numpy.linalg.eigvalsh with the
reference's clamp (v < 0 -> 0) and filter (v > 1e-15); closed forms ln 2 and 0.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def random_rho(n, seed, rank=None):
    rng = np.random.default_rng(seed)
    d = 1 << n
    k = rank or d
    a = rng.standard_normal((d, k)) + 1j * rng.standard_normal((d, k))
    r = a @ a.conj().T
    return r / np.trace(r).real


def numpy_partial_tr(rho, n, traced_bits):
    # axis j of the reshaped tensor is bit n-1-j of the index
    t = rho.reshape([2] * (2 * n))
    kept = [b for b in range(n) if b not in traced_bits]
    # move traced bits to be contracted: row axis n-1-b pairs with column axis 2n-1-b
    letters = "abcdefghijklmnopqrstuvwxyz"
    row = [letters[j] for j in range(n)]
    col = [letters[n + j].upper() for j in range(n)]
    for b in traced_bits:
        col[n - 1 - b] = row[n - 1 - b]
    out_row = [row[n - 1 - b] for b in reversed(kept)]
    out_col = [col[n - 1 - b] for b in reversed(kept)]
    spec = "".join(row) + "".join(col) + "->" + "".join(out_row) + "".join(out_col)
    red = np.einsum(spec, t)
    dk = 1 << len(kept)
    return red.reshape(dk, dk)


@pytest.mark.parametrize("n,traced", [(2, [0]), (2, [1]), (3, [0, 2]), (4, [1]), (4, [3, 0]), (5, [2, 4, 1]), (3, [])])
def test_partial_tr_vs_numpy(n, traced):
    rho = random_rho(n, 10 * n + len(traced))
    got = O.dm_partial_tr(n, rho, traced)
    want = numpy_partial_tr(rho, n, sorted(set(traced)))
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 1e-14


def test_partial_tr_closed_forms():
    bell = np.zeros(4, dtype=complex)
    bell[0] = bell[3] = 1 / math.sqrt(2)
    rho = np.outer(bell, bell.conj())
    assert np.allclose(O.dm_partial_tr(2, rho, [0]), np.eye(2) / 2, atol=1e-15)
    # product |0><0| (bit 1) x |+><+| (bit 0): tracing bit 1 leaves |+><+|
    plus = np.array([1, 1]) / math.sqrt(2)
    prod = np.kron(np.diag([1.0, 0.0]), np.outer(plus, plus))
    assert np.allclose(O.dm_partial_tr(2, prod, [1]), np.outer(plus, plus), atol=1e-15)
    assert np.allclose(O.dm_partial_tr(2, prod, [0]), np.diag([1.0, 0.0]), atol=1e-15)


def numpy_entropy(rho):
    ev = np.linalg.eigvalsh(rho)
    ev = np.maximum(ev, 0.0)
    ev = ev[ev > 1e-15]
    return float(-np.sum(ev * np.log(ev)))


@pytest.mark.parametrize("n,rank", [(1, None), (2, 1), (3, 2), (4, None), (5, 3), (6, None)])
def test_entropy_vs_numpy(n, rank):
    rho = random_rho(n, 100 + n, rank)
    assert abs(O.dm_entropy(n, rho) - numpy_entropy(rho)) < 1e-10


def test_entropy_closed_forms():
    assert abs(O.dm_entropy(1, np.eye(2) / 2) - math.log(2)) < 1e-14
    assert abs(O.dm_entropy(3, np.eye(8) / 8) - 3 * math.log(2)) < 1e-13
    psi = np.random.default_rng(1).standard_normal(8) + 0j
    psi /= np.linalg.norm(psi)
    assert abs(O.dm_entropy(3, np.outer(psi, psi.conj()))) < 1e-12
