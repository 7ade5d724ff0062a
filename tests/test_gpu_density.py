"""GPU parity of the density-matrix path (gates, Kraus channels, trace, purity,
expect_dm, probs) against the oracle and the reference's noise fixtures.
Bar: entries within 1e-12 relative; trace preserved to 1e-12."""
import math

import numpy as np
import pytest

from conftest import load_golden, rel_err, seeded_state
from fixture_circuits import noisy_brickwork, random_circuit
from oracle import oracle as O
from test_oracle import KRAUS_FIXTURES
from yaohip.circuit import (Circuit, Gate, NoiseChannel, OperatorPolynomial, OperatorString, Op, channel, control,
                            lower, put)

pytestmark = pytest.mark.gpu


def _mat(d):
    return np.array(d["re"]) + 1j * np.array(d["im"])


@pytest.mark.parametrize("fused", [False, True, "blocks"])
def test_dm_cases_match_julia(gpu, fused):
    fx = load_golden("noise.json")["dm_cases"]
    cases = {
        "h_phaseflip_0.1": (1, [put([0], Gate.H), channel([0], NoiseChannel.PhaseFlip(0.1))]),
        "bitflip_0.1_on_zero": (1, [channel([0], NoiseChannel.BitFlip(0.1))]),
        "h_depolarizing_0.1": (1, [put([0], Gate.H), channel([0], NoiseChannel.Depolarizing(1, 0.1))]),
        "bell_depolarizing_q0_0.1": (2, [put([0], Gate.H), control([0], [1], Gate.X),
                                         channel([0], NoiseChannel.Depolarizing(1, 0.1))]),
        "h_amplitude_damping_0.3": (1, [put([0], Gate.H), channel([0], NoiseChannel.AmplitudeDamping(0.3, 0.0))]),
    }
    for name, (n, els) in cases.items():
        dm = gpu.DensityMatrix.zero_state(n)
        dm.apply(Circuit.qubits(n, els), fused=fused)
        assert np.max(np.abs(dm.state_data() - _mat(fx[name]["rho"]).reshape(-1))) < 1e-10, name


def test_dm_expectation_fixtures(gpu):
    fx = load_golden("noise.json")["expectations"]
    cases = {
        "z_h_pure": ([put([0], Gate.H)], Op.Z),
        "x_h_depolarizing_0.1": ([put([0], Gate.H), channel([0], NoiseChannel.Depolarizing(1, 0.1))], Op.X),
        "z_h_phaseflip_0.1": ([put([0], Gate.H), channel([0], NoiseChannel.PhaseFlip(0.1))], Op.Z),
    }
    for name, (els, op) in cases.items():
        dm = gpu.DensityMatrix.zero_state(1)
        dm.apply(Circuit.qubits(1, els))
        v = gpu.expect_dm(dm, OperatorPolynomial.single(0, op, 1.0))
        assert abs(v - complex(fx[name]["value_re"], fx[name]["value_im"])) < 1e-10, name


def random_noisy(n, ngates, seed):
    rng = np.random.default_rng(seed)
    chans = list(KRAUS_FIXTURES.values()) + [NoiseChannel.Reset(0.1, 0.2), NoiseChannel.Reset(0.3, 0.0)]
    out = []
    for e in random_circuit(n, ngates, seed, max_ctrls=1, custom_targets=(1, 2, 3)):
        out.append(e)
        if rng.random() < 0.6:
            q = int(rng.integers(n))
            out.append(channel([q], chans[rng.integers(len(chans))]))
        if rng.random() < 0.1 and n >= 2:
            a, b = [int(x) for x in rng.permutation(n)[:2]]
            out.append(channel([a, b], NoiseChannel.Depolarizing(2, 0.05)))
    return out


@pytest.mark.parametrize("fused", [False, True, "blocks"])
@pytest.mark.parametrize("n,seed", [(1, 0), (3, 1), (5, 2), (7, 3), (9, 4)])
def test_random_noisy_circuits_vs_oracle(gpu, n, seed, fused):
    els = random_noisy(n, 40, 100 + seed)
    psi = seeded_state(n, seed)
    rho0 = O.dm_from_reg(n, psi)
    want = O.dm_apply(n, rho0, lower(els))
    dm = gpu.DensityMatrix.from_reg(gpu.ArrayReg.from_vec(n, psi))
    assert rel_err(dm.state_data(), rho0) < 1e-15
    dm.apply(Circuit.qubits(n, els), fused=fused)
    got = dm.state_data()
    assert rel_err(got, want) < 1e-12
    assert abs(dm.trace() - O.dm_trace(n, want)) < 1e-12
    assert abs(dm.trace() - 1.0) < 1e-12
    assert abs(dm.purity() - O.dm_purity(n, want)) < 1e-12
    H = OperatorPolynomial([1.0, 0.5j, -0.3], [OperatorString([(0, Op.Z)] + ([(1, Op.X)] if n > 1 else [])),
                                              OperatorString([(n - 1, Op.Y)]), OperatorString([(0, Op.Pu)])])
    assert abs(gpu.expect_dm(dm, H) - O.dm_expect(n, want, H)) < 1e-12
    # probabilities of the same rho are bit-exact; against the oracle's rho, 1e-12
    assert np.array_equal(gpu.probs(dm), O.dm_probs(n, got))
    assert rel_err(gpu.probs(dm), O.dm_probs(n, want)) < 1e-12
    locs = [n - 1, 0] if n > 1 else [0]
    assert np.array_equal(gpu.probs(dm, locs), O.dm_probs(n, got, locs))


@pytest.mark.parametrize("fused", [False, True, "blocks"])
def test_noisy_brickwork_8q(gpu, fused):
    """The noisy brickwork workload at 8 qubits, depth 3."""
    n = 8
    els = noisy_brickwork(n, 3, 17)
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    want = O.dm_apply(n, O.dm_from_reg(n, z), lower(els))
    dm = gpu.DensityMatrix.zero_state(n)
    dm.apply(Circuit.qubits(n, els), fused=fused)
    assert rel_err(dm.state_data(), want) < 1e-12
    assert abs(dm.trace() - 1) < 1e-12
    for i in range(n - 1):
        zz = OperatorPolynomial([1.0], [OperatorString([(i, Op.Z), (i + 1, Op.Z)])])
        assert abs(gpu.expect_dm(dm, zz) - O.dm_expect(n, want, zz)) < 1e-12


def test_dm_unitary_matches_pure(gpu):
    n = 6
    psi = seeded_state(n, 44)
    els = random_circuit(n, 50, 44, max_ctrls=2, custom_targets=(1, 2, 3, 4))
    reg = gpu.ArrayReg.from_vec(n, psi)
    dm = gpu.DensityMatrix.from_reg(reg)
    dm.apply(Circuit.qubits(n, els))
    gpu.apply_inplace(Circuit.qubits(n, els), reg, 0)
    out = reg.state_vec()
    assert rel_err(dm.state_data(), np.outer(out, out.conj()).reshape(-1)) < 1e-12
    assert abs(dm.purity() - 1) < 1e-12


def test_dm_large_trace_and_purity_14q(gpu):
    """14 qubits (2^28 entries): trace kept at 1, purity 1 for a unitary layer, and
    below 1 after noise."""
    n = 14
    dm = gpu.DensityMatrix.zero_state(n)
    dm.apply(Circuit.qubits(n, noisy_brickwork(n, 1, 3)[:60]), fused=True)
    assert abs(dm.trace() - 1) < 1e-12
    p = dm.purity()
    assert 0.9 < p < 1.0


def test_dm_blocks_match_fused_14q(gpu):
    """The superoperator block path (12-bit tile passes at 2^28 entries) against the
    per-gate fused path on the noisy brickwork: trace, purity and every ZZ."""
    n = 14
    els = noisy_brickwork(n, 2, 3)
    print("[case] dm blocks vs fused 14q", flush=True)
    a = gpu.DensityMatrix.zero_state(n)
    a.apply(Circuit.qubits(n, els), fused=True)
    b = gpu.DensityMatrix.zero_state(n)
    b.apply(Circuit.qubits(n, els), fused="blocks")
    assert abs(b.trace() - 1) < 1e-12
    assert abs(b.purity() - a.purity()) < 1e-12
    for i in range(n - 1):
        zz = OperatorPolynomial([1.0], [OperatorString([(i, Op.Z), (i + 1, Op.Z)])])
        assert abs(gpu.expect_dm(b, zz) - gpu.expect_dm(a, zz)) < 1e-12


def reference_noisy_circuit(nq):
    """build_noisy_circuit of the reference's validation suite: H on all, CNOT
    chain, Depolarizing(n=1, p=0.01) on all, Rz(0.3) on all, AmplitudeDamping(0.05)."""
    els = [put([q], Gate.H) for q in range(nq)]
    els += [control([q], [q + 1], Gate.X) for q in range(nq - 1)]
    els += [channel([q], NoiseChannel.Depolarizing(1, 0.01)) for q in range(nq)]
    els += [put([q], Gate.Rz(0.3)) for q in range(nq)]
    els += [channel([q], NoiseChannel.AmplitudeDamping(0.05, 0.0)) for q in range(nq)]
    return els


@pytest.mark.parametrize("nq", [4, 6, 8, 10])
def test_partial_tr_and_entropy_reference_circuit(gpu, nq):
    els = reference_noisy_circuit(nq)
    d, dk = 1 << nq, 1 << (nq - 1)
    dm = gpu.DensityMatrix.zero_state(nq)
    dm.apply(Circuit.qubits(nq, els))
    rho = dm.state_data().reshape(d, d)
    red = dm.partial_tr([nq - 1])
    got = red.state_data().reshape(dk, dk)
    # bit-identical to the oracle's partial trace of the same matrix (same summation order)
    assert np.array_equal(got, O.dm_partial_tr(nq, rho, [nq - 1]))
    # and close to the oracle's evolution
    z = np.zeros((d, d), dtype=complex)
    z[0, 0] = 1
    want_rho = O.dm_apply(nq, z, lower(els)).reshape(d, d)
    assert np.max(np.abs(got - O.dm_partial_tr(nq, want_rho, [nq - 1]))) < 1e-12
    s = red.von_neumann_entropy()
    if nq <= 8:
        want_s = O.dm_entropy(nq - 1, got)
    else:  # the oracle's Jacobi sweep is slow at 512 x 512; numpy's eigvalsh pins it
        ev = np.maximum(np.linalg.eigvalsh(got), 0.0)
        ev = ev[ev > 1e-15]
        want_s = float(-np.sum(ev * np.log(ev)))
    print(f"nq={nq}: reduced entropy {s:.12f} oracle {want_s:.12f}")
    assert abs(s - want_s) < 1e-10


def test_partial_tr_edge_cases(gpu):
    import yaohip
    n = 3
    rng = np.random.default_rng(3)
    a = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    rho = a @ a.conj().T
    rho /= np.trace(rho).real
    dm = gpu.DensityMatrix.from_matrix(n, rho)
    # every qubit traced: the 1 x 1 trace; no qubit traced: the matrix itself;
    # traced locations at or beyond nbits trace nothing
    full = dm.partial_tr([0, 1, 2])
    assert full.nbits() == 0 and abs(full.state_data()[0] - np.trace(rho)) < 1e-14
    assert full.von_neumann_entropy() == 0.0
    assert np.array_equal(dm.partial_tr([]).state_data().reshape(8, 8), rho)
    assert np.array_equal(dm.partial_tr([5]).state_data().reshape(8, 8), rho)
    assert np.array_equal(dm.partial_tr([2, 0, 2]).state_data().reshape(2, 2), O.dm_partial_tr(n, rho, [2, 0]))
    with pytest.raises(yaohip.YhError):
        dm.partial_tr([-1])
    # entropy of the mixed 3-qubit matrix against numpy's eigvalsh
    ev = np.maximum(np.linalg.eigvalsh(rho), 0.0)
    ev = ev[ev > 1e-15]
    assert abs(dm.von_neumann_entropy() - float(-np.sum(ev * np.log(ev)))) < 1e-10


def random_kraus(nq, nk, seed):
    """nk Kraus operators on nq qubits with sum K^dagger K = I (a random isometry cut
    into blocks)."""
    rng = np.random.default_rng(seed)
    d = 1 << nq
    a = rng.standard_normal((nk * d, d)) + 1j * rng.standard_normal((nk * d, d))
    q, _ = np.linalg.qr(a)
    return [q[k * d:(k + 1) * d] for k in range(nk)]


@pytest.mark.parametrize("fused", [False, True, "blocks"])
def test_channels_on_three_or_more_qubits(gpu, fused):
    """apply_channel (density_matrix.rs:181-224) for channels the 2-qubit kernels do not
    cover: Depolarizing on 3 qubits (noise.rs:462-508, 64 Kraus operators), a
    3-qubit Custom Kraus set, a 4-qubit Depolarizing channel (256 operators) and a
    1-qubit Custom channel of 70 operators; against the oracle at 1e-12, trace kept."""
    n = 6
    psi = seeded_state(n, 71)
    els = [put([q], Gate.H) for q in range(n)] + [control([0], [1], Gate.X), control([2], [4], Gate.X)]
    els += [channel([1, 3, 5], NoiseChannel.Depolarizing(3, 0.2)), put([2], Gate.Ry(0.7))]
    els += [channel([4, 0, 2], NoiseChannel.Custom(random_kraus(3, 5, 1))), control([3], [5], Gate.X)]
    els += [channel([5, 1, 0, 3], NoiseChannel.Depolarizing(4, 0.1))]
    els += [channel([2], NoiseChannel.Custom(random_kraus(1, 70, 2))), put([1], Gate.Rx(0.4))]
    print(f"[case] dm channels >= 3 qubits fused={fused}", flush=True)
    dm = gpu.DensityMatrix.from_reg(gpu.ArrayReg.from_vec(n, psi))
    dm.apply(Circuit.qubits(n, els), fused=fused)
    want = O.dm_apply(n, O.dm_from_reg(n, psi), lower(els))
    assert rel_err(dm.state_data(), want) < 1e-12
    assert abs(dm.trace() - 1) < 1e-12


def test_channel_three_qubits_on_a_large_matrix(gpu):
    """A 3-qubit depolarizing channel on a 12-qubit density matrix (2^24 entries): the
    Kraus branch sum (noise.rs:462-508 weights 1 - p + p/64 on the identity and p/64
    on the other Pauli strings) against the Pauli-twirl closed form
    rho -> (1 - p) rho + p tr_S(rho) (x) I/8 on the channel's qubits S."""
    n, p = 12, 0.3
    els = [put([q], Gate.Ry(0.1 + 0.2 * q)) for q in range(n)] + [control([q], [q + 1], Gate.X) for q in range(n - 1)]
    dm = gpu.DensityMatrix.zero_state(n)
    dm.apply(Circuit.qubits(n, els))
    rho = dm.state_data()
    S = [2, 7, 9]
    dm.apply(Circuit.qubits(n, [channel(S, NoiseChannel.Depolarizing(3, p))]))
    got = dm.state_data()
    # tensor axes: row of loc l is axis l, column of loc l is axis n + l
    t = rho.reshape([2] * (2 * n))
    lab = list(range(2 * n))
    tr_in = [l if (l >= n and l - n in S) is False else l - n for l in lab]
    out = [l for l in lab if l not in S and l - n not in S]
    tr = np.einsum(t, tr_in, out)
    eye = np.eye(8).reshape([2] * 6)
    full = np.einsum(tr, out, eye, S + [n + l for l in S], lab) / 8
    want = (1 - p) * rho + p * full.reshape(-1)
    assert np.max(np.abs(got - want)) < 1e-12
    assert abs(dm.trace() - 1) < 1e-12


@pytest.mark.parametrize("n", [11, 12])
def test_entropy_beyond_ten_qubits(gpu, n):
    """von_neumann_entropy (density_matrix.rs:113-122) without a size cap: above 10
    qubits the eigenvalues come from the device eigensolver; against numpy's eigvalsh
    (which pins the oracle's eigensolver at small sizes) and the closed form of a
    mixture of orthonormal states, -sum w ln w."""
    rng = np.random.default_rng(n)
    k = 6
    w = rng.random(k)
    w /= w.sum()
    q, _ = np.linalg.qr(rng.standard_normal((1 << n, k)) + 1j * rng.standard_normal((1 << n, k)))
    regs = [gpu.ArrayReg.from_vec(n, q[:, j]) for j in range(k)]
    dm = gpu.DensityMatrix.mixed(w, regs)
    s = dm.von_neumann_entropy()
    want = float(-np.sum(w * np.log(w)))
    print(f"[case] entropy n={n}: {s:.14f} closed form {want:.14f}", flush=True)
    assert abs(s - want) < 1e-10
    # the same matrix after noise, against numpy
    dm.apply(Circuit.qubits(n, [channel([q], NoiseChannel.Depolarizing(1, 0.05)) for q in range(3)]))
    ev = np.maximum(np.linalg.eigvalsh(dm.state_data().reshape(1 << n, 1 << n)), 0.0)
    ev = ev[ev > 1e-15]
    assert abs(dm.von_neumann_entropy() - float(-np.sum(ev * np.log(ev)))) < 1e-10
