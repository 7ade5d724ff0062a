"""Parity at the benchmark workloads' full sizes (the reference's style: whole-state
comparison on every case, tests/suites/apply.rs:700-733).

- 28-qubit random brickwork, depth 20, complex128 (4 GiB), from deterministic_state:
  the fused MFMA block path against the per-gate path on the full circuit (1e-12
  relative), and both against the oracle on the circuit's first two layers
  (1e-12; the oracle runs on the host's cores, see orc_set_threads).
- The same circuit in complex64: error of every reported result against the FP64
  state, per result (amplitudes, fidelity, <Z0 Z1>, marginal probabilities).
- 14-qubit noisy density matrix, depth 10 (2^28 entries): the superoperator block
  path against the per-gate fused path (entries 1e-12 relative, trace 1e-12, all 13
  <Z_i Z_i+1>), and the 10-qubit depth-10 circuit against the oracle on all three
  density-matrix paths.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from yaohip.circuit import Circuit, Op, OperatorPolynomial, OperatorString, lower
from yaohip.workloads import brickwork, brickwork_layers, noisy_brickwork

pytestmark = pytest.mark.gpu

N28, DEPTH28, SEED28 = 28, 20, 20  # the benchmark's circuit


def threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def max_rel(a, b, chunk=1 << 24):
    """max |a - b| / max |b| over two large complex arrays, in chunks."""
    num, den = 0.0, 0.0
    for i in range(0, a.size, chunk):
        x, y = a[i:i + chunk], b[i:i + chunk]
        num = max(num, float(np.max(np.abs(x - y))))
        den = max(den, float(np.max(np.abs(y))))
    return num / den


def zz(i):
    return OperatorPolynomial([1.0], [OperatorString([(i, Op.Z), (i + 1, Op.Z)])])


def test_brickwork28_depth20_fused_vs_gate_by_gate(gpu):
    els = brickwork(N28, DEPTH28, SEED28)
    circ = Circuit.qubits(N28, els)
    print(f"[case] brickwork 28q depth 20 c128: {len(els)} gates, blocks vs gate by gate", flush=True)
    a = gpu.ArrayReg.deterministic_state(N28)
    b = a.clone()
    gpu.apply_inplace(circ, a, 5)
    gpu.apply_inplace(circ, b, 0)
    assert abs(a.norm() - 1) < 1e-12 and abs(b.norm() - 1) < 1e-12
    sa = a.state_vec()
    del a
    sb = b.state_vec()
    err = max_rel(sa, sb)
    print(f"[case] fused vs gate by gate: max rel err {err:.3e}", flush=True)
    assert err < 1e-12


def test_brickwork28_first_two_layers_vs_oracle(gpu):
    els = brickwork_layers(N28, DEPTH28, SEED28, 2)
    print(f"[case] brickwork 28q first 2 layers ({len(els)} gates) vs oracle, {threads()} threads", flush=True)
    reg = gpu.ArrayReg.deterministic_state(N28)
    psi = reg.state_vec()
    gbg = reg.clone()
    gpu.apply_inplace(Circuit.qubits(N28, els), reg, 5)
    gpu.apply_inplace(Circuit.qubits(N28, els), gbg, 0)
    got = reg.state_vec()
    del reg
    got0 = gbg.state_vec()
    del gbg
    prev = O.set_threads(threads())
    try:
        want = O.apply(N28, psi, lower(els))
    finally:
        O.set_threads(prev)
    del psi
    e5, e0 = max_rel(got, want), max_rel(got0, want)
    print(f"[case] vs oracle: blocks {e5:.3e}, gate by gate {e0:.3e}", flush=True)
    assert e5 < 1e-12 and e0 < 1e-12


def test_brickwork28_complex64_errors_per_result(gpu):
    """complex64 amplitudes with FP32 MFMA blocks, FP64 reductions: every result's
    error against the complex128 path (itself within 1e-12 of the oracle)."""
    els = brickwork(N28, DEPTH28, SEED28)
    circ = Circuit.qubits(N28, els)
    print("[case] brickwork 28q depth 20 c64 vs c128, per result", flush=True)
    r64 = gpu.ArrayReg.deterministic_state(N28, "c64")
    gpu.apply_inplace(circ, r64, 5)
    r128 = gpu.ArrayReg.deterministic_state(N28)
    gpu.apply_inplace(circ, r128, 5)
    res = {}
    for i in (0, 13, 26):
        res[f"zz{i}"] = abs(gpu.expect_arrayreg(r64, zz(i)) - gpu.expect_arrayreg(r128, zz(i)))
    for locs in ([0, 1], [0, 1, 2, 3, 4, 5, 6, 7], [27, 3, 14]):
        res[f"marg{locs}"] = float(np.max(np.abs(gpu.probs(r64, locs) - gpu.probs(r128, locs))))
    s128 = r128.state_vec()
    del r128
    s64 = r64.state_vec()
    del r64
    res["amp_max_rel"] = max_rel(s64, s128)
    n64, n128 = np.vdot(s64, s64).real, np.vdot(s128, s128).real
    res["norm_dev"] = abs(n64 - 1.0)
    res["fidelity_loss"] = 1.0 - abs(np.vdot(s128, s64)) ** 2 / (n64 * n128)
    print(f"[case] complex64 errors: {res}", flush=True)
    assert res["amp_max_rel"] < 1e-5
    assert 0 <= res["fidelity_loss"] < 1e-9 and res["norm_dev"] < 1e-6
    assert all(res[k] < 1e-6 for k in res if k.startswith("zz") or k.startswith("marg"))


@pytest.mark.parametrize("fused", [0, 4, 5])
def test_qft12_complex64_errors_per_result(gpu, fused):
    """The 12-qubit QFT with SWAPs on a seeded state in complex64 (gate by gate, one LDS
    tile pass, FP32 MFMA blocks) against the oracle (FP64): every reported result's
    error, per result."""
    from conftest import seeded_state
    from yaohip.workloads import qft_with_swaps
    n = 12
    psi = seeded_state(n, 2024)
    els = qft_with_swaps(n)
    want = O.apply(n, psi, lower(els))
    reg = gpu.ArrayReg.from_vec(n, psi, "c64")
    gpu.apply_inplace(Circuit.qubits(n, els), reg, fused)
    got = reg.state_vec()
    res = {"amp_max_rel": max_rel(got, want),
           "fidelity_loss": 1.0 - abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)}
    for i in (0, 5, 10):
        res[f"zz{i}"] = abs(gpu.expect_arrayreg(reg, zz(i)) - O.expect(n, want, zz(i)).real)
    for locs in ([0, 1], list(range(12)), [11, 3, 7]):
        res[f"marg{locs if len(locs) < 12 else 'all'}"] = float(np.max(np.abs(gpu.probs(reg, locs) - O.probs(n, want, locs))))
    print(f"[case] qft12 complex64 fused={fused} errors: {res}", flush=True)
    assert res["amp_max_rel"] < 1e-6
    assert all(v < 1e-6 for k, v in res.items() if k != "amp_max_rel")


def test_noisy_dm14_depth10_blocks_vs_per_gate(gpu):
    n = 14
    els = noisy_brickwork(n, 10, 3)
    circ = Circuit.qubits(n, els)
    print(f"[case] noisy dm 14q depth 10: {len(els)} elements, blocks vs per-gate fused", flush=True)
    a = gpu.DensityMatrix.zero_state(n)
    a.apply(circ, fused="blocks")
    b = gpu.DensityMatrix.zero_state(n)
    b.apply(circ, fused=True)
    ta, tb = a.trace(), b.trace()
    assert abs(ta - 1) < 1e-12 and abs(tb - 1) < 1e-12
    assert abs(a.purity() - b.purity()) < 1e-12
    dz = [abs(gpu.expect_dm(a, zz(i)) - gpu.expect_dm(b, zz(i))) for i in range(n - 1)]
    assert max(dz) < 1e-12
    ra = a.state_data()
    del a
    err = max_rel(ra, b.state_data())
    print(f"[case] dm14 entries max rel err {err:.3e}, trace {ta}, max |dZZ| {max(dz):.2e}", flush=True)
    assert err < 1e-12


def test_noisy_dm10_depth10_vs_oracle(gpu):
    n = 10
    els = noisy_brickwork(n, 10, 3)
    circ = Circuit.qubits(n, els)
    print(f"[case] noisy dm 10q depth 10 ({len(els)} elements) vs oracle, {threads()} threads", flush=True)
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    prev = O.set_threads(threads())
    try:
        want = O.dm_apply(n, O.dm_from_reg(n, z), lower(els))
    finally:
        O.set_threads(prev)
    wzz = [O.dm_expect(n, want, zz(i)) for i in range(n - 1)]
    for mode in (False, True, "blocks"):
        dm = gpu.DensityMatrix.zero_state(n)
        dm.apply(circ, fused=mode)
        err = max_rel(dm.state_data(), want)
        print(f"[case] dm10 mode={mode}: max rel err {err:.3e}", flush=True)
        assert err < 1e-12
        assert abs(dm.trace() - 1) < 1e-12
        assert abs(dm.purity() - O.dm_purity(n, want)) < 1e-12
        assert max(abs(gpu.expect_dm(dm, zz(i)) - wzz[i]) for i in range(n - 1)) < 1e-12
