"""GPU parity of the state-vector path against the oracle and the reference's golden
vectors. Bars: amplitudes within 1e-12 relative (max-abs error / max-abs amplitude)
in complex128; probabilities of the full distribution and sampled outcomes
bit-exact; complex64 within 2e-5 relative (its error is printed)."""
import ctypes
import math

import numpy as np
import pytest

from conftest import load_golden, rel_err, seeded_state
from fixture_circuits import (NAMED, brickwork, elements_from_case, parse_matrix, parse_state, qft_with_swaps,
                              random_circuit)
from oracle import oracle as O
from yaohip import _abi as A
from yaohip.circuit import (Circuit, Gate, OperatorPolynomial, OperatorString, Op, PositionedGate, control, lower,
                            put, tfim_hamiltonian)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run_gpu(gpu, n, psi, els, dtype="c128", fused=0):
    reg = gpu.ArrayReg.from_vec(n, psi, dtype)
    gpu.apply_inplace(Circuit.qubits(n, els), reg, fused)
    return reg.state_vec()


def test_instruct_fixtures(gpu):
    tested = 0
    for case in load_golden("instruct.json")["cases"]:
        label = case["label"]
        if "locs" not in case or "regression" in label:
            continue
        n = case["nbits"]
        cl = case.get("ctrl_locs", [])
        cb = [bool(b) for b in case.get("ctrl_bits", [])]
        if "gate_matrix" in case:
            m = parse_matrix(case["gate_matrix"])
            gates = [Gate.Custom(m, False)]
            if np.count_nonzero(m - np.diag(np.diag(m))) == 0:
                gates.append(Gate.Custom(m, True))   # the diagonal kernel as well
            elss = [[PositionedGate(g, case["locs"], cl, cb)] for g in gates]
        elif case["gate_name"] in NAMED:
            g = NAMED[case["gate_name"]]([case.get("theta", 0.0)])
            if g.num_sites() == 2:
                elss = [[PositionedGate(g, case["locs"], cl, cb)]]
            else:
                elss = [[PositionedGate(g, [l], cl, cb) for l in case["locs"]]]
        else:
            continue
        inp, want = parse_state(case["input_state"]), parse_state(case["output_state"])
        for els in elss:
            got = run_gpu(gpu, n, inp, els)
            assert np.max(np.abs(got - want)) < 1e-10, label
            assert rel_err(got, O.apply(n, inp, lower(els))) < TOL, label
        tested += 1
    assert tested >= 40


def test_instruct_regression_fixture(gpu):
    case = [c for c in load_golden("instruct.json")["cases"] if "regression" in c["label"]][0]
    m = parse_matrix(case["gate_matrix"])
    els = [put(pair, Gate.Custom(m, False)) for pair in case["gate_pairs"]]
    got = run_gpu(gpu, case["nbits"], parse_state(case["input_state"]), els)
    assert np.max(np.abs(got - parse_state(case["output_state"]))) < 1e-8


def test_apply_fixtures(gpu):
    tested = 0
    for case in load_golden("apply.json")["cases"]:
        if any(d != 2 for d in case["dims"]):
            continue
        n = len(case["dims"])
        if case.get("input_state_re") is not None:
            inp = np.array(case["input_state_re"]) + 1j * np.array(case["input_state_im"])
        else:
            inp = np.zeros(1 << n, dtype=complex)
            inp[0] = 1
        els = elements_from_case(case)
        got = run_gpu(gpu, n, inp, els)
        want = np.array(case["output_state_re"]) + 1j * np.array(case["output_state_im"])
        assert np.max(np.abs(got - want)) < 1e-10, case["label"]
        assert rel_err(got, O.apply(n, inp, lower(els))) < TOL, case["label"]
        tested += 1
    assert tested >= 60


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (7, 3), (10, 4), (12, 5), (13, 6)])
def test_random_circuits_all_dispatch_classes(gpu, n, seed):
    custom = (1, 2, 3, 4, 5, 6, 7) if n >= 7 else tuple(range(1, n + 1))
    els = random_circuit(n, 160, 500 + seed, max_ctrls=min(3, n - 1), custom_targets=custom)
    psi = seeded_state(n, seed)
    print(f"[case] random circuit n={n} seed={seed}", flush=True)
    got = run_gpu(gpu, n, psi, els)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


@pytest.mark.parametrize("n", [6, 11])
def test_complex64_error_vs_fp64_oracle(gpu, n):
    els = random_circuit(n, 120, 900 + n, max_ctrls=2, custom_targets=(1, 2, 3))
    psi = seeded_state(n, n)
    got = run_gpu(gpu, n, psi, els, "c64")
    err = rel_err(got, O.apply(n, psi, lower(els)))
    print(f"complex64 relative error n={n}: {err:.3e}")
    assert err < 2e-5


def test_brickwork_20q_vs_oracle(gpu):
    n = 20
    els = brickwork(n, 4, 31)
    psi = seeded_state(n, 31)
    got = run_gpu(gpu, n, psi, els)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


def dagger(els):
    inv = {"Rx": Gate.Rx, "Ry": Gate.Ry, "Rz": Gate.Rz, "Phase": Gate.Phase}
    out = []
    for e in reversed(els):
        g = e.gate
        from yaohip import _abi as A
        name = A.GATE_NAMES[g.kind]
        g2 = inv[name](-g.params[0]) if name in inv else g
        out.append(PositionedGate(g2, e.target_locs, e.control_locs, e.control_configs))
    return out


def test_brickwork_26q_round_trip_and_norm(gpu):
    """Full-size property: U then U^dagger returns the input; the norm is kept."""
    n = 26
    els = brickwork(n, 2, 5)
    psi = seeded_state(n, 9)
    reg = gpu.ArrayReg.from_vec(n, psi)
    gpu.apply_inplace(Circuit.qubits(n, els), reg, 0)
    assert abs(reg.norm() - 1.0) < 1e-12
    gpu.apply_inplace(Circuit.qubits(n, dagger(els)), reg, 0)
    assert rel_err(reg.state_vec(), psi) < 1e-12


def test_norm_normalize_fidelity(gpu):
    n = 9
    psi = seeded_state(n, 1) * 3.0
    reg = gpu.ArrayReg.from_vec(n, psi)
    assert abs(reg.norm() - O.norm(n, psi)) < 1e-13
    other = gpu.ArrayReg.from_vec(n, seeded_state(n, 2))
    assert abs(reg.fidelity(other) - O.fidelity(n, psi, seeded_state(n, 2))) < 1e-12
    reg.normalize()
    assert rel_err(reg.state_vec(), psi / 3.0) < 1e-15
    z = gpu.ArrayReg.zero_state(5)
    s = z.state_vec()
    assert s[0] == 1 and np.count_nonzero(s) == 1


def test_expectations_vs_oracle(gpu):
    n = 11
    psi = seeded_state(n, 4)
    reg = gpu.ArrayReg.from_vec(n, psi)
    H = tfim_hamiltonian(n, 1.0, 1.0)
    assert abs(gpu.expect_arrayreg(reg, H) - O.expect(n, psi, H)) < 1e-12
    rng = np.random.default_rng(3)
    coeffs, strings = [], []
    for _ in range(21):
        k = int(rng.integers(1, 5))
        sites = rng.permutation(n)[:k]
        strings.append(OperatorString([(int(s), int(rng.integers(0, 8))) for s in sites]))
        coeffs.append(complex(rng.standard_normal(), rng.standard_normal()))
    P = OperatorPolynomial(coeffs, strings)
    assert abs(gpu.expect_arrayreg(reg, P) - O.expect(n, psi, P)) < 1e-12


def test_probs_bit_exact(gpu):
    n = 14
    psi = seeded_state(n, 8)
    reg = gpu.ArrayReg.from_vec(n, psi)
    assert np.array_equal(gpu.probs(reg), O.probs(n, psi))
    for locs in ([0], [3, 1], [13, 0, 7], list(range(n))[::-1], [5, 9, 2, 11]):
        assert np.array_equal(gpu.probs(reg, locs), O.probs(n, psi, locs)), locs


def test_probs_chunked_marginal_close(gpu):
    """Beyond 2^18 members per bucket the sum is split into ordered chunks."""
    n = 21
    psi = seeded_state(n, 12)
    reg = gpu.ArrayReg.from_vec(n, psi)
    got = gpu.probs(reg, [4, 17])
    assert rel_err(got, O.probs(n, psi, [4, 17])) < 1e-13


def test_sample_bit_exact_with_same_uniforms(gpu):
    n = 12
    psi = seeded_state(n, 21)
    reg = gpu.ArrayReg.from_vec(n, psi)
    rng = np.random.default_rng(42)
    u = rng.random(4096)
    # adversarial uniforms: exactly at the sequential partial sums
    cs, acc = [], 0.0
    for p in O.probs(n, psi):
        acc += p
        cs.append(acc)
    u = np.concatenate([u, np.array(cs[::97]), [0.0, 1.0 - 2 ** -53, cs[-1]]])
    for locs in (None, [0, 5], [11, 3, 7]):
        got = gpu.sample(reg, locs, u)
        want = O.sample(n, psi, locs, u)
        assert np.array_equal(got, want), locs


@pytest.mark.parametrize("n", [1, 3, 13, 22])
def test_sample_bit_exact_cdf_chunks(gpu, n):
    """sample_from_probs (measure.rs:140-150) over distributions of 2 .. 2^22 bins:
    the device's sequential cumulative sum is folded in LDS chunks of 4096 with
    ragged ends; every shot, including uniforms exactly at partial sums at chunk
    edges, matches the oracle's sequential loop."""
    psi = seeded_state(n, 60 + n)
    reg = gpu.ArrayReg.from_vec(n, psi)
    p = O.probs(n, psi)
    cs = np.cumsum(p)  # numpy's cumsum is the same sequential left fold:
    acc, seq = 0.0, []
    for x in p[:9000]:
        acc += x
        seq.append(acc)
    assert np.array_equal(np.array(seq), cs[:len(seq)])
    edges = [i for i in (0, 1, 4094, 4095, 4096, 4097, 8191, 8192, len(p) - 2, len(p) - 1) if 0 <= i < len(p)]
    u = np.concatenate([np.random.default_rng(n).random(3000), cs[edges], np.nextafter(cs[edges], 0.0)])
    got = gpu.sample(reg, None, u)
    want = np.searchsorted(cs, u, side="right")
    want[want >= len(p)] = len(p) - 1
    assert np.array_equal(got, want)
    if n <= 13:
        assert np.array_equal(got, O.sample(n, psi, None, u))


def _adversarial_distributions():
    """States whose probabilities |a|^2 are exact dyadic numbers, chosen against the
    parallel form of the sequential cumulative sum: basis states (one bucket of 1,
    the rest zero), GHZ, uniform, runs of exact half-ulp ties before and after the sum
    changes binade (also at a chunk edge of 4096), subnormal probabilities and
    probabilities near 2^600; up to 2^14 buckets (one workgroup) and above."""
    out = []
    n = 12
    N = 1 << n
    for k in (0, 5, N - 1):
        z = np.zeros(N, complex)
        z[k] = 1
        out.append((f"basis {k}", n, z))
    z = np.zeros(N, complex)
    z[0] = z[-1] = 0.5 + 0.5j  # p = 0.5 each
    out.append(("ghz", n, z))
    out.append(("uniform", n, np.full(N, 2.0 ** -6 + 0j)))
    z = np.full(N, 2.0 ** -27 + 2.0 ** -27 * 1j)  # p = 2^-53: half an ulp of 1.0
    z[0] = 1
    out.append(("ties after 1", n, z))
    # p in {0, 2^-54, 2^-53, 2^-52, 2^-50 + 2^-54}: ties at 2^-54 while the sum is in
    # [0.5, 1), at 2^-53 in [1, 2) and at 2^-52 in [2, 4); the 0.5 and 1 buckets move it
    n = 13
    N = 1 << n
    vals = np.array([0, 2.0 ** -27, 2.0 ** -27 + 2.0 ** -27 * 1j, 2.0 ** -26, 2.0 ** -25 + 2.0 ** -27 * 1j])
    rng = np.random.default_rng(77)
    z = vals[rng.integers(0, len(vals), N)].astype(complex)
    z[0] = 0.5 + 0.5j
    z[4096] = 0.5 + 0.5j
    z[6000] = 1
    out.append(("mixed ties", n, z))
    z = np.zeros(N, complex)
    z[rng.integers(0, N, 300)] = 2.0 ** -530  # p = 2^-1060, subnormal
    z[rng.integers(0, N, 30)] = 2.0 ** -511
    z[-5] = 2.0 ** -20
    out.append(("subnormal", n, z))
    out.append(("scaled 2^600", 12, seeded_state(12, 5) * 2.0 ** 300))
    z = seeded_state(14, 6)
    z[rng.random(1 << 14) < 0.3] = 0  # zero runs across chunks
    out.append(("random with zeros", 14, z))
    # above 16384 buckets the chunks are planned together and walked by one workgroup
    n = 16
    N = 1 << n
    z = vals[rng.integers(0, len(vals), N)].astype(complex)
    z[0] = 0.5 + 0.5j
    z[4096 * 5] = 0.5 + 0.5j
    z[4096 * 9 + 17] = 1
    out.append(("mixed ties, 16 chunks", n, z))
    z = np.full(1 << 15, 2.0 ** -27 + 2.0 ** -27 * 1j)  # every chunk folded bucket by bucket
    z[3] = 1
    out.append(("ties after 1, 8 chunks", 15, z))
    z = np.zeros(1 << 16, complex)
    z[(1 << 16) - 1] = 1
    out.append(("basis last, 16 chunks", 16, z))
    z = seeded_state(17, 7)
    z[rng.random(1 << 17) < 0.5] = 0
    z[: 1 << 13] = 0
    out.append(("random with zeros, 32 chunks", 17, z))
    return out


@pytest.mark.parametrize("case", range(14))
def test_sample_bit_exact_adversarial(gpu, case):
    """Every partial sum of the device cumulative sum equals the sequential fold's:
    uniforms at each partial sum and just below it pick the same bucket as the
    oracle's loop (measure.rs:140-150)."""
    name, n, z = _adversarial_distributions()[case]
    print(f"[case] cdf adversarial '{name}' n={n}", flush=True)
    reg = gpu.ArrayReg.from_vec(n, z)
    p = O.probs(n, z)
    cs = np.empty(len(p))
    acc = 0.0
    for i, x in enumerate(p):
        acc += x
        cs[i] = acc
    u = np.concatenate([cs, np.nextafter(cs, 0.0), np.random.default_rng(case).random(500) * cs[-1]])
    got = gpu.sample(reg, None, u)
    want = np.searchsorted(cs, u, side="right")
    want[want >= len(p)] = len(p) - 1
    assert np.array_equal(got, want)
    assert np.array_equal(got, O.sample(n, z, None, u))


@pytest.mark.parametrize("kind", ["seeded", "uniform", "alternating chunks"])
def test_sample_every_partial_sum_many_chunks(gpu, kind):
    """Every partial sum of a 2^18-bucket cumulative sum (64 chunks of 4,096, whose
    records the cross-chunk walk takes in groups of one binade) equals the sequential
    fold's (measure.rs:140-150): uniforms at each partial sum and just below it pick
    the bucket of the sequential loop. The uniform distribution changes binade
    exactly at chunk edges; the third alternates chunk magnitudes by 2^20."""
    n = 18
    N = 1 << n
    if kind == "seeded":
        z = seeded_state(n, 77)
    elif kind == "uniform":
        z = np.full(N, 2.0 ** -9 + 0j)
    else:
        amp = np.where((np.arange(N) // 4096) % 2 == 0, 1.0, 2.0 ** -10)
        z = amp * np.exp(1j * 0.37 * np.arange(N))
        z = z / np.linalg.norm(z)
    print(f"[case] cdf every partial sum '{kind}' n={n}", flush=True)
    reg = gpu.ArrayReg.from_vec(n, z)
    p = O.probs(n, z)
    cs = np.empty(N)
    acc = 0.0
    for i, x in enumerate(p.tolist()):
        acc += x
        cs[i] = acc
    u = np.concatenate([cs, np.nextafter(cs, 0.0)])
    got = gpu.sample(reg, None, u)
    want = np.searchsorted(cs, u, side="right")
    want[want >= N] = N - 1
    assert np.array_equal(got, want)


def test_sample_marginal_and_complex64_many_buckets(gpu):
    """The cumulative sum over more than one 4,096-bucket chunk reached through a
    marginal distribution (13 of 15 qubits, 8,192 buckets) and through a complex64
    register (14 qubits): every partial sum, and just below it, picks the oracle's
    bucket."""
    n = 15
    psi = seeded_state(n, 81)
    locs = [14, 0, 3, 7, 1, 12, 5, 9, 2, 11, 6, 10, 4]
    reg = gpu.ArrayReg.from_vec(n, psi)
    p = O.probs(n, psi, locs)
    cs = np.cumsum(p)
    u = np.concatenate([cs, np.nextafter(cs, 0.0)])
    assert np.array_equal(gpu.sample(reg, locs, u), O.sample(n, psi, locs, u))
    n = 14
    reg = gpu.ArrayReg.from_vec(n, seeded_state(n, 82), "c64")
    z = reg.state_vec()  # the complex64 amplitudes, exactly
    p = O.probs(n, z)
    cs = np.cumsum(p)
    u = np.concatenate([cs, np.nextafter(cs, 0.0), np.random.default_rng(3).random(200)])
    assert np.array_equal(gpu.sample(reg, None, u), O.sample(n, z, None, u))


def test_measure_postprocess_many_buckets(gpu):
    """measure_with_postprocess on 13 of 14 qubits (8,192 outcomes): outcome bits
    and the collapsed state as the oracle's."""
    n = 14
    psi = seeded_state(n, 83)
    locs = list(range(13))
    for r in (0.0, 0.37, 0.999):
        reg = gpu.ArrayReg.from_vec(n, psi)
        bits = gpu.measure_with_postprocess(reg, locs, gpu.PostProcess.NoPostProcess, lambda: r)
        bits_o, state_o = O.measure(n, psi, locs, 0, 0, r)
        assert bits == bits_o
        assert rel_err(reg.state_vec(), state_o) < 1e-12


@pytest.mark.parametrize("post", [0, 1, 2])
def test_measure_postprocess_vs_oracle(gpu, post):
    n = 8
    psi = seeded_state(n, 30 + post)
    locs = [6, 1, 3]
    for r in (0.05, 0.5, 0.93):
        reg = gpu.ArrayReg.from_vec(n, psi)
        P = [gpu.PostProcess.NoPostProcess, gpu.PostProcess.ResetTo(5), gpu.PostProcess.RemoveMeasured][post]
        res = gpu.measure_with_postprocess(reg, locs, P, lambda: r)
        bits_o, state_o = O.measure(n, psi, locs, post, 5, r)
        if post == 2:
            bits, newreg = res
            assert rel_err(newreg.state_vec(), state_o) < 1e-12
        else:
            bits = res
            assert rel_err(reg.state_vec(), state_o) < 1e-12
        assert bits == bits_o


def test_measure_errors(gpu):
    reg = gpu.ArrayReg.from_vec(3, seeded_state(3, 0))
    with pytest.raises(gpu.YhError) as e:
        gpu.measure_with_postprocess(reg, [0, 0], gpu.PostProcess.NoPostProcess, lambda: 0.5)
    assert e.value.status == 7
    with pytest.raises(gpu.YhError) as e:
        gpu.measure_with_postprocess(reg, [0], gpu.PostProcess.ResetTo(2), lambda: 0.5)
    assert e.value.status == 13
    with pytest.raises(gpu.YhError) as e:
        gpu.apply_inplace(Circuit.qubits(4, [put([0], Gate.H)]), reg)
    assert e.value.status == 6


@pytest.mark.parametrize("fused", [0, 4, 5])
def test_qft12_config(gpu, fused):
    """12-qubit QFT with SWAPs on a seeded N(0,1) state: amplitudes, Z(0)Z(1), 1024 shots,
    gate by gate, as one LDS tile pass (the whole 12-qubit state is one tile) and as
    fused blocks."""
    n = 12
    psi = seeded_state(n, 2024)
    els = qft_with_swaps(n)
    reg = gpu.ArrayReg.from_vec(n, psi)
    gpu.apply_inplace(Circuit.qubits(n, els), reg, fused)
    want = O.apply(n, psi, lower(els))
    assert rel_err(reg.state_vec(), want) < TOL
    zz = OperatorPolynomial([1.0], [OperatorString([(0, Op.Z), (1, Op.Z)])])
    assert abs(gpu.expect_arrayreg(reg, zz) - O.expect(n, want, zz)) < 1e-12
    u = np.random.default_rng(1024).random(1024)
    assert np.array_equal(gpu.sample(reg, None, u), O.sample(n, reg.state_vec(), None, u))


# ---------------- tile passes (yh_apply_fused) ----------------
@pytest.mark.parametrize("n,seed", [(12, 0), (13, 1), (15, 2), (16, 3)])
def test_fused_random_circuits_vs_oracle(gpu, n, seed):
    """All dispatch classes, including gates too wide for a tile (run on their own)."""
    els = random_circuit(n, 250, 700 + seed, max_ctrls=3, custom_targets=(1, 2, 3, 4, 5, 6))
    psi = seeded_state(n, seed)
    print(f"[case] fused random n={n} seed={seed}", flush=True)
    got = run_gpu(gpu, n, psi, els, fused=4)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


@pytest.mark.parametrize("n,depth", [(14, 6), (20, 4)])
def test_fused_brickwork_vs_oracle(gpu, n, depth):
    els = brickwork(n, depth, 40 + n)
    psi = seeded_state(n, n)
    got = run_gpu(gpu, n, psi, els, fused=4)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


def test_fused_qft_vs_oracle(gpu):
    n = 16
    psi = seeded_state(n, 16)
    els = qft_with_swaps(n)
    got = run_gpu(gpu, n, psi, els, fused=4)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


def test_fused_complex64_error(gpu):
    n = 14
    els = brickwork(n, 6, 3)
    psi = seeded_state(n, 3)
    got = run_gpu(gpu, n, psi, els, "c64", fused=4)
    err = rel_err(got, O.apply(n, psi, lower(els)))
    print(f"complex64 tiled relative error n={n}: {err:.3e}")
    assert err < 2e-5


def test_fused_matches_gate_by_gate_26q(gpu):
    """Full-size cross-check between the two GPU paths (each pinned to the oracle
    at smaller sizes): depth-4 brickwork on 26 qubits."""
    n = 26
    els = brickwork(n, 4, 8)
    psi = seeded_state(n, 8)
    a = run_gpu(gpu, n, psi, els, fused=0)
    b = run_gpu(gpu, n, psi, els, fused=4)
    assert rel_err(b, a) < TOL


# ---------------- 5-qubit MFMA blocks (yh_apply_fused, max_fuse_qubits = 5) ----------------
@pytest.mark.parametrize("n,seed", [(5, 0), (6, 1), (9, 2), (11, 3), (12, 4), (13, 6), (15, 7), (16, 5)])
def test_block5_random_circuits_vs_oracle(gpu, n, seed):
    """All dispatch classes; gates on more than 5 qubits run on their own. n < 11
    leaves fewer than 64 columns per block (partial workgroup steps); from n = 13 on,
    runs of blocks are grouped into 12-bit LDS tile passes (k_tblocks), blocks
    moving past blocks on disjoint bits."""
    els = random_circuit(n, 200, 900 + seed, max_ctrls=3, custom_targets=(1, 2, 3, 4, 5, 6))
    psi = seeded_state(n, 50 + seed)
    print(f"[case] block5 random n={n} seed={seed}", flush=True)
    got = run_gpu(gpu, n, psi, els, fused=5)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


@pytest.mark.parametrize("n,depth", [(13, 6), (14, 6), (20, 4)])
def test_block5_brickwork_vs_oracle(gpu, n, depth):
    """Brickwork plans are mostly 16 x 16 blocks, so their passes take two tiles per
    workgroup step; n = 13 has a single such step (two tiles) per pass."""
    els = brickwork(n, depth, 60 + n)
    psi = seeded_state(n, 60 + n)
    print(f"[case] block5 brickwork n={n} depth={depth}", flush=True)
    got = run_gpu(gpu, n, psi, els, fused=5)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


def test_block5_qft_and_controlled_x(gpu):
    n = 16
    psi = seeded_state(n, 17)
    els = qft_with_swaps(n) + [PositionedGate(Gate.X, [7], [3, 9], [True, False]),
                               PositionedGate(Gate.X, [15], [0, 1, 2, 3, 4], [True] * 5)]
    print("[case] block5 qft16 + controlled X", flush=True)
    got = run_gpu(gpu, n, psi, els, fused=5)
    assert rel_err(got, O.apply(n, psi, lower(els))) < TOL


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_block_passes_with_16x16_blocks_vs_oracle(gpu, dtype):
    """A random 16-qubit circuit whose launch plan puts blocks acting on at most 4
    bits into multi-block passes (applied as 16 x 16 MFMA blocks): checked against
    the oracle in both precisions."""
    n = 16
    els = random_circuit(n, 300, 0)
    c = Circuit.qubits(n, els)
    st = (ctypes.c_int64 * 7)()
    A.check(A.lib().yh_block_pass_stats(n, c.ops.arr, c.ops.n, 0, st))
    assert st[3] > 0 and st[1] > 0  # 16 x 16 blocks in multi-block passes
    psi = seeded_state(n, 71)
    print(f"[case] block passes with 16x16 blocks {dtype}", flush=True)
    got = run_gpu(gpu, n, psi, els, dtype, fused=5)
    err = rel_err(got, O.apply(n, psi, lower(els)))
    assert err < (2e-5 if dtype == "c64" else TOL)


def test_block5_complex64_error(gpu):
    n = 14
    els = brickwork(n, 6, 5)
    psi = seeded_state(n, 5)
    print("[case] block5 complex64", flush=True)
    got = run_gpu(gpu, n, psi, els, "c64", fused=5)
    err = rel_err(got, O.apply(n, psi, lower(els)))
    print(f"complex64 block5 relative error n={n}: {err:.3e}")
    assert err < 2e-5


def test_block5_matches_gate_by_gate_26q(gpu):
    n = 26
    els = brickwork(n, 4, 9)
    psi = seeded_state(n, 9)
    print("[case] block5 vs gate-by-gate 26q", flush=True)
    a = run_gpu(gpu, n, psi, els, fused=0)
    b = run_gpu(gpu, n, psi, els, fused=5)
    assert rel_err(b, a) < TOL


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_low_target_bits_in_wave_match_pair_kernel(gpu, dtype):
    """A dense 1-qubit gate on memory bits 0..5 runs the in-wave exchange kernel,
    on bits >= 6 the pair kernel; both do the same operations per pair, so the gate
    on bit b equals, bit for bit, the gate on bit 8 of the state with bits b and 8
    exchanged (exchanged back afterwards)."""
    n = 10
    psi = seeded_state(n, 91)
    g = Gate.Rx(0.731)

    def swap_bits(v, a, b):
        t = np.asarray(v).reshape([2] * n)  # axis k <-> memory bit n - 1 - k
        return np.swapaxes(t, n - 1 - a, n - 1 - b).reshape(-1)

    for b in range(6):
        got = run_gpu(gpu, n, psi, [put([n - 1 - b], g)], dtype)
        via8 = run_gpu(gpu, n, swap_bits(psi, b, 8), [put([n - 1 - 8], g)], dtype)
        assert np.array_equal(got, swap_bits(via8, b, 8)), f"bit {b}"
