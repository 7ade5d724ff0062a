"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker:
against the reference's golden vectors (tests/golden/*.json, generated by Yao.jl),
its bitbasis index lists, closed forms, a brute-force dense embedding, and
hand-computed boundary cases for sampling and post-processing."""
import math

import numpy as np
import pytest

from conftest import load_golden, seeded_state
from fixture_circuits import (elements_from_case, parse_matrix, parse_state, qft_no_swaps, qft_with_swaps,
                              random_circuit)
from oracle import oracle as O
from yaohip.circuit import (Gate, NoiseChannel, OperatorPolynomial, OperatorString, Op, PositionedGate, channel,
                            control, lower, put)
from yaohip import _abi as A


# ---------------- bitbasis (bitbasis/src/iterate_control.rs:166-219) ----------------
def test_itercontrol_reference_lists():
    assert O.itercontrol(7, [0, 2, 3, 6], [1, 0, 1, 0]) == [9, 11, 25, 27, 41, 43, 57, 59]
    assert O.itercontrol(3, [], []) == list(range(8))
    assert O.itercontrol(2, [0, 1], [1, 0]) == [1]
    pos = [0, 2, 4, 6, 8, 10, 12, 14, 16, 18]
    cfg = [1, 0, 1, 0, 1, 0, 1, 0, 1, 0]
    vals = O.itercontrol(19, pos, cfg)
    assert len(vals) == 1 << 9
    for b in vals:
        for p, c in zip(pos, cfg):
            assert (b >> p) & 1 == c
    assert vals == sorted(vals)


# ---------------- gate matrices (tests/data/gates.json) ----------------
def test_gate_matrices_match_gates_json():
    from fixture_circuits import NAMED
    for e in load_golden("gates.json")["gates"]:
        g = NAMED[e["name"]](e["params"])
        ol = lower([put(list(range(g.num_sites())), g)])
        m = O.gate_matrix(ol.arr[0])
        want = np.array(e["matrix_re"]) + 1j * np.array(e["matrix_im"])
        assert np.max(np.abs(m - want)) < 1e-12, e["name"]


# ---------------- instructs (tests/data/instruct.json, Yao.jl) ----------------
def _instruct_cases():
    return load_golden("instruct.json")["cases"]


def test_instruct_direct_calls_match_julia():
    """The reference unit tests' direct instruct calls (src/unit_tests/instruct_qubit.rs:49-406)."""
    tested = 0
    for case in _instruct_cases():
        label = case["label"]
        if "regression" in label or "measure" in label or "locs" not in case:
            continue
        if "gate_matrix" not in case:
            continue
        n = case["nbits"]
        locs = case["locs"]
        m = parse_matrix(case["gate_matrix"])
        inp, want = parse_state(case["input_state"]), parse_state(case["output_state"])
        cl, cb = case.get("ctrl_locs", []), case.get("ctrl_bits", [])
        if m.shape == (2, 2) and len(locs) == 1:
            fn = "1q_controlled" if cl else "1q"
        elif m.shape == (4, 4) and len(locs) == 2:
            fn = "2q_controlled" if cl else "2q"
        else:
            continue
        got = O.instruct(fn, n, inp, locs, m, cl, cb)
        assert np.max(np.abs(got - want)) < 1e-10, label
        tested += 1
    assert tested >= 15


def test_instruct_named_gates_match_julia():
    """Named-gate cases go through dispatch (apply.rs:7-120); gates over several
    locs apply to each loc in turn."""
    from fixture_circuits import NAMED
    tested = 0
    for case in _instruct_cases():
        if "gate_name" not in case or case["gate_name"] in ("PSWAP", "CPHASE"):
            continue
        n = case["nbits"]
        g = NAMED[case["gate_name"]]([case.get("theta", 0.0)])
        cl = case.get("ctrl_locs", [])
        cb = [bool(b) for b in case.get("ctrl_bits", [])]
        if g.num_sites() == 2:
            els = [PositionedGate(g, case["locs"], cl, cb)]
        else:
            els = [PositionedGate(g, [l], cl, cb) for l in case["locs"]]
        got = O.apply(n, parse_state(case["input_state"]), lower(els))
        assert np.max(np.abs(got - parse_state(case["output_state"]))) < 1e-10, case["label"]
        tested += 1
    assert tested >= 20


def test_instruct_regression_20_random_2q():
    case = [c for c in _instruct_cases() if "regression" in c["label"]][0]
    m = parse_matrix(case["gate_matrix"])
    s = parse_state(case["input_state"])
    for pair in case["gate_pairs"]:
        s = O.instruct("2q", case["nbits"], s, pair, m)
    assert np.max(np.abs(s - parse_state(case["output_state"]))) < 1e-8


def test_instruct_probability_case():
    case = [c for c in _instruct_cases() if "measure" in c["label"]][0]
    p = O.probs(case["nbits"], parse_state(case["input_state"]))
    assert np.max(np.abs(p - np.array(case["probabilities"]))) < 1e-12


# ---------------- apply (tests/data/apply.json) ----------------
def test_apply_ground_truth():
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
        got = O.apply(n, inp, lower(elements_from_case(case)))
        want = np.array(case["output_state_re"]) + 1j * np.array(case["output_state_im"])
        assert np.max(np.abs(got - want)) < 1e-10, case["label"]
        tested += 1
    assert tested >= 60


# ---------------- brute-force dense embedding (yao gate order, F4) ----------------
def dense_unitary(n, pg):
    g = pg.gate
    if g.kind == A.GATE_CUSTOM:
        M = g.matrix
    else:
        M = O.gate_matrix(lower([put(list(range(g.num_sites())), g)]).arr[0])
    N = 1 << n
    U = np.zeros((N, N), dtype=complex)
    tb = [n - 1 - l for l in pg.target_locs]
    cbit = [n - 1 - l for l in pg.control_locs]
    for c in range(N):
        if not all(((c >> b) & 1) == int(v) for b, v in zip(cbit, pg.control_configs)):
            U[c, c] = 1
            continue
        j = sum(((c >> b) & 1) << k for k, b in enumerate(tb))
        base = c
        for b in tb:
            base &= ~(1 << b)
        for r in range(1 << len(tb)):
            out = base | sum(((r >> k) & 1) << b for k, b in enumerate(tb))
            U[out, c] += M[r, j]
    return U


@pytest.mark.parametrize("seed", range(6))
def test_apply_matches_dense_embedding(seed):
    n = 5 + seed % 2
    els = random_circuit(n, 40, 1000 + seed, max_targets=3, max_ctrls=2)
    psi = seeded_state(n, seed)
    want = psi.copy()
    for e in els:
        want = dense_unitary(n, e) @ want
    got = O.apply(n, psi, lower(els))
    assert np.max(np.abs(got - want)) < 1e-12


def test_qft_analytical():
    """QFT without SWAPs on |1> (tests/suites/benchmark_validation.rs:152-175)."""
    for n in range(4, 11):
        s = np.zeros(1 << n, dtype=complex)
        s[1] = 1
        got = O.apply(n, s, lower(qft_no_swaps(n)))
        N = 1 << n
        rev = np.array([int(format(k, f"0{n}b")[::-1], 2) for k in range(N)])
        want = np.exp(2j * math.pi * rev / N) / math.sqrt(N)
        assert np.max(np.abs(got - want)) < 1e-10


def test_qft_with_swaps_is_dft():
    """The textbook QFT with SWAPs is the unitary DFT with +i sign."""
    n = 6
    psi = seeded_state(n, 7)
    got = O.apply(n, psi, lower(qft_with_swaps(n)))
    want = np.fft.ifft(psi) * math.sqrt(1 << n)
    assert np.max(np.abs(got - want)) < 1e-12


# ---------------- measurement (tests/data/measure.json) ----------------
def test_measure_probabilities_ground_truth():
    for case in load_golden("measure.json")["cases"]:
        n = case["num_qubits"]
        s = np.zeros(1 << n, dtype=complex)
        s[0] = 1
        s = O.apply(n, s, lower(elements_from_case({"gates": case["gates"]})))
        assert np.max(np.abs(O.probs(n, s) - np.array(case["probabilities"]))) < 1e-10, case["label"]


def test_marginal_probs_bit_order():
    n = 4
    psi = seeded_state(n, 3)
    p = np.abs(psi) ** 2
    locs = [2, 0]
    got = O.probs(n, psi, locs)
    want = np.zeros(4)
    for b in range(16):
        bits = [(b >> (n - 1 - l)) & 1 for l in locs]   # loc l is bit n-1-l
        want[bits[0] * 2 + bits[1]] += p[b]             # locs[0] is the output MSB
    assert np.max(np.abs(got - want)) < 1e-15


def test_sample_boundaries():
    p = [0.25, 0.25, 0.5]
    assert O.sample_from_probs(p, 0.0) == 0
    assert O.sample_from_probs(p, 0.25) == 1     # strict r < cumsum
    assert O.sample_from_probs(p, 0.4999) == 1
    assert O.sample_from_probs(p, 0.5) == 2
    assert O.sample_from_probs(p, 1.0) == 2      # no cumsum exceeds r: last index
    assert O.sample_from_probs([0.0, 0.5, 0.0, 0.5], 0.0) == 1
    assert O.sample_from_probs([0.0, 0.5, 0.0, 0.5], 0.5) == 3


def test_measure_postprocess_closed_forms():
    ghz = np.zeros(8, dtype=complex)
    ghz[0] = ghz[7] = 1 / math.sqrt(2)
    bits, s = O.measure(3, ghz, [0], 0, 0, 0.2)
    assert bits == [0] and np.allclose(s, ghz)
    bits, s = O.measure(3, ghz, [0], 1, 0, 0.7)          # outcome 1, reset loc 0 to 0
    assert bits == [1]
    want = np.zeros(8, dtype=complex)
    want[0b011] = 1
    assert np.max(np.abs(s - want)) < 1e-15
    bits, s = O.measure(3, ghz, [0, 2], 2, 0, 0.7)       # remove measured qubits
    assert bits == [1, 1]
    want = np.zeros(2, dtype=complex)
    want[1] = 1
    assert np.max(np.abs(s - want)) < 1e-15


def test_measure_rejects_bad_locs():
    with pytest.raises(O.OracleError) as e:
        O.probs(3, seeded_state(3, 0), [0, 0])
    assert e.value.status == 7  # YH_ERR_MEASURE_LOC
    with pytest.raises(O.OracleError):
        O.probs(3, seeded_state(3, 0), [3])


# ---------------- noise (tests/data/noise.json) ----------------
KRAUS_FIXTURES = {
    "phase_amp_damping_0.3_0.2_0.4": NoiseChannel.PhaseAmplitudeDamping(0.3, 0.2, 0.4),
    "depolarizing_1_0.1": NoiseChannel.Depolarizing(1, 0.1),
    "amplitude_damping_0.3_0.0": NoiseChannel.AmplitudeDamping(0.3, 0.0),
    "thermal_relaxation_100_80_10_0.0": NoiseChannel.ThermalRelaxation(100.0, 80.0, 10.0, 0.0),
    "pauli_0.1_0.2_0.05": NoiseChannel.PauliChannel(0.1, 0.2, 0.05),
    "amplitude_damping_0.3_0.4": NoiseChannel.AmplitudeDamping(0.3, 0.4),
    "bit_flip_0.1": NoiseChannel.BitFlip(0.1),
    "phase_flip_0.2": NoiseChannel.PhaseFlip(0.2),
    "phase_damping_0.2": NoiseChannel.PhaseDamping(0.2),
    "phase_amp_damping_0.3_0.2_0.0": NoiseChannel.PhaseAmplitudeDamping(0.3, 0.2, 0.0),
}


def _mat(d):
    return np.array(d["re"]) + 1j * np.array(d["im"])


@pytest.mark.parametrize("name", sorted(KRAUS_FIXTURES))
def test_kraus_and_superop_match_julia(name):
    fx = load_golden("noise.json")["kraus"][name]
    ol = lower([channel([0], KRAUS_FIXTURES[name])])
    ks = O.kraus(ol)
    assert len(ks) == fx["num_kraus"]
    for k, want in zip(ks, fx["kraus"]):
        assert np.max(np.abs(k - _mat(want))) < 1e-10
    assert np.max(np.abs(O.superop(ol) - _mat(fx["superop"]))) < 1e-10
    # completeness sum K^dagger K = I (trace preservation)
    assert np.max(np.abs(sum(k.conj().T @ k for k in ks) - np.eye(2))) < 1e-14


def test_depolarizing_two_qubit_completeness():
    ks = O.kraus(lower([channel([0, 1], NoiseChannel.Depolarizing(2, 0.3))]))
    assert len(ks) == 16
    assert np.max(np.abs(sum(k.conj().T @ k for k in ks) - np.eye(4))) < 1e-14


def _plus():
    return np.array([1, 1], dtype=complex) / math.sqrt(2)


def test_dm_cases_match_julia():
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
        z = np.zeros(1 << n, dtype=complex)
        z[0] = 1
        rho = O.dm_apply(n, O.dm_from_reg(n, z), lower(els))
        want = _mat(fx[name]["rho"]).reshape(-1)
        assert np.max(np.abs(rho - want)) < 1e-10, name


def test_dm_expectations_match_julia():
    fx = load_golden("noise.json")["expectations"]
    z = np.array([1, 0], dtype=complex)
    cases = {
        "z_h_pure": ([put([0], Gate.H)], Op.Z),
        "x_h_depolarizing_0.1": ([put([0], Gate.H), channel([0], NoiseChannel.Depolarizing(1, 0.1))], Op.X),
        "z_h_phaseflip_0.1": ([put([0], Gate.H), channel([0], NoiseChannel.PhaseFlip(0.1))], Op.Z),
    }
    for name, (els, op) in cases.items():
        rho = O.dm_apply(1, O.dm_from_reg(1, z), lower(els))
        v = O.dm_expect(1, rho, OperatorPolynomial.single(0, op, 1.0))
        assert abs(v - complex(fx[name]["value_re"], fx[name]["value_im"])) < 1e-10, name


def test_dm_unitary_matches_pure_outer_product():
    n = 4
    psi = seeded_state(n, 11)
    els = random_circuit(n, 30, 77, max_targets=3, max_ctrls=1)
    rho = O.dm_apply(n, O.dm_from_reg(n, psi), lower(els))
    out = O.apply(n, psi, lower(els))
    assert np.max(np.abs(rho - np.outer(out, out.conj()).reshape(-1))) < 1e-12
    assert abs(O.dm_purity(n, rho) - 1) < 1e-12


def test_dm_trace_preserved_and_mixed_purity():
    n = 3
    els = [put([0], Gate.H), control([0], [1], Gate.X)]
    for q in range(n):
        els.append(channel([q], NoiseChannel.Depolarizing(1, 1.0)))
    z = np.zeros(8, dtype=complex)
    z[0] = 1
    rho = O.dm_apply(n, O.dm_from_reg(n, z), lower(els))
    assert abs(O.dm_trace(n, rho) - 1) < 1e-12
    # fully depolarized single qubits: the maximally mixed state, purity 1/2^n
    assert abs(O.dm_purity(n, rho) - 1 / 8) < 1e-12


def test_expect_dm_lsb_site_identity():
    """expect_dm reads sites as bit positions: <Z_a Z_b>_dm(psi) == <Z_{n-1-a} Z_{n-1-b}>_psi."""
    n = 5
    psi = seeded_state(n, 5)
    rho = O.dm_from_reg(n, psi)
    for a, b in [(0, 1), (1, 3), (0, 4)]:
        dm_v = O.dm_expect(n, rho, OperatorPolynomial([1.0], [OperatorString([(a, Op.Z), (b, Op.Z)])]))
        pure_v = O.expect(n, psi, OperatorPolynomial([1.0], [OperatorString([(n - 1 - a, Op.Z),
                                                                               (n - 1 - b, Op.Z)])]))
        assert abs(dm_v - pure_v) < 1e-12


def test_expect_pauli_closed_forms():
    """<Z> = cos(theta) after Rx/Ry(theta); <X> = sin(theta) after Ry; <Y> = -sin after Rx."""
    for th in [0.3, 1.1, -2.0]:
        z = np.array([1, 0], dtype=complex)
        ry = O.apply(1, z, lower([put([0], Gate.Ry(th))]))
        rx = O.apply(1, z, lower([put([0], Gate.Rx(th))]))
        assert abs(O.expect(1, ry, OperatorPolynomial.single(0, Op.Z, 1.0)) - math.cos(th)) < 1e-14
        assert abs(O.expect(1, ry, OperatorPolynomial.single(0, Op.X, 1.0)) - math.sin(th)) < 1e-14
        assert abs(O.expect(1, rx, OperatorPolynomial.single(0, Op.Y, 1.0)) + math.sin(th)) < 1e-14
    # projectors and ladder operators against the dense operator
    n = 3
    psi = seeded_state(n, 2)
    mats = {Op.X: [[0, 1], [1, 0]], Op.Y: [[0, -1j], [1j, 0]], Op.Z: [[1, 0], [0, -1]], Op.P0: [[1, 0], [0, 0]],
            Op.P1: [[0, 0], [0, 1]], Op.Pu: [[0, 1], [0, 0]], Op.Pd: [[0, 0], [1, 0]], Op.I: [[1, 0], [0, 1]]}
    for op in mats:
        for site in range(n):
            M = np.array([[1]])
            for q in range(n):
                M = np.kron(M, np.array(mats[op] if q == site else mats[Op.I]))
            want = np.vdot(psi, M @ psi)
            got = O.expect(n, psi, OperatorPolynomial.single(site, op, 1.0))
            assert abs(got - want) < 1e-14


# ---------------- Circuit::new validation (circuit.rs:150-272) ----------------
def test_validation_errors():
    ok = lower([control([0], [1], Gate.X)])
    assert O.validate(2, ok) == 0
    assert O.validate(2, lower([PositionedGate(Gate.X, [1], [0], [True, False])])) == 2
    assert O.validate(2, lower([put([2], Gate.H)])) == 3
    assert O.validate(2, lower([control([1], [1], Gate.X)])) == 4
    assert O.validate(2, lower([put([0], Gate.SWAP)])) == 5
    assert O.validate(2, lower([channel([0, 1], NoiseChannel.BitFlip(0.1))])) == 5
    assert O.validate(2, lower([channel([0, 0], NoiseChannel.Depolarizing(2, 0.1))])) == 4


# ---------------- constructors (register.rs, density_matrix.rs) ----------------
def test_deterministic_state_known_answers():
    """register.rs:188-194: 4 qubits, 16 amplitudes, norm 1, amplitudes 0 and 1
    differ; and the definition's closed forms: amplitude 0 is (1, 0) / norm, the
    phase of amplitude k is atan2(sin 0.2k, cos 0.1k)."""
    s = O.deterministic_state(4)
    assert s.shape == (16,) and abs(np.linalg.norm(s) - 1) < 1e-12 and abs(s[0] - s[1]) > 1e-6
    nrm = math.sqrt(sum(math.cos(0.1 * k) ** 2 + math.sin(0.2 * k) ** 2 for k in range(16)))
    assert abs(s[0] - 1 / nrm) < 1e-15
    assert abs(np.angle(s[7]) - math.atan2(math.sin(1.4), math.cos(0.7))) < 1e-14


def test_rand_state_from_uniforms():
    """rand_state (register.rs:35-45) with a fixed stream: constant uniforms 0.75 give
    every amplitude (0.25 + 0.25i) before normalisation, i.e. the uniform state times
    e^{i pi/4}; uniforms 0.5 in the imaginary slots give a real state."""
    n = 5
    s = O.rand_state(n, np.full(2 << n, 0.75))
    np.testing.assert_allclose(s, np.exp(1j * math.pi / 4) / math.sqrt(1 << n), rtol=0, atol=1e-15)
    u = np.random.default_rng(1).random(2 << n)
    u[1::2] = 0.5
    s = O.rand_state(n, u)
    assert np.all(s.imag == 0) and abs(np.linalg.norm(s) - 1) < 1e-14
    np.testing.assert_allclose(s.real * np.linalg.norm(u[0::2] - 0.5), u[0::2] - 0.5, rtol=0, atol=1e-15)


def test_mixed_known_answers():
    """density_matrix.rs:304-310 (purity 1/2 of an equal mixture of |0>, |1>),
    :322-330 (entropy ln 2), tests/suites/measure.rs:610-619 (diagonal 0.25, 0.75)."""
    e0, e1 = np.array([1, 0], dtype=complex), np.array([0, 1], dtype=complex)
    r = O.dm_mixed(1, [0.5, 0.5], [e0, e1])
    assert abs(O.dm_purity(1, r) - 0.5) < 1e-12
    assert abs(O.dm_entropy(1, r.reshape(2, 2)) - math.log(2)) < 1e-8
    np.testing.assert_allclose(O.dm_probs(1, O.dm_mixed(1, [0.25, 0.75], [e0, e1])), [0.25, 0.75], atol=1e-15)
    # a one-register mixture with weight 1 is from_reg, bit for bit
    psi = seeded_state(4, 3)
    assert np.array_equal(O.dm_mixed(4, [1.0], [psi]), O.dm_from_reg(4, psi))


def test_density_matrix_from_reg_known_answers():
    """density_matrix_from_reg (density_matrix.rs:266-272) = partial_tr of from_reg:
    keeping every bit is from_reg; one qubit of a Bell pair is maximally mixed
    (:334-341 of the partial-trace test); for a product state the reduced matrix
    of the kept bits is the product of their factors."""
    psi = seeded_state(3, 5)
    full = O.density_matrix_from_reg(3, psi, [0, 1, 2]).reshape(-1)
    assert np.array_equal(full, O.dm_from_reg(3, psi))
    bell = np.array([1, 0, 0, 1], dtype=complex) / math.sqrt(2)
    red = O.density_matrix_from_reg(2, bell, [0])
    np.testing.assert_allclose(red, np.eye(2) / 2, atol=1e-15)
    a, b = seeded_state(1, 6), seeded_state(2, 7)
    prod = np.kron(a, b)  # bits 2 (a) above bits 1, 0 (b)
    red = O.density_matrix_from_reg(3, prod, [0, 1])
    np.testing.assert_allclose(red, np.outer(b, b.conj()), atol=1e-15)
