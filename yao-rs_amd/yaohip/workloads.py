"""Circuit builders of the benchmark workloads, as yao-rs element lists.

- qft_with_swaps: the QFT with the final bit-reversal SWAPs (examples/qft.rs:30-54,
  the same circuit as benches/apply.rs:18-43); 12 qubits is the 84-gate circuit of
  the plain CPU-runnable workload.
- qft_no_swaps: easybuild::qft_circuit (easybuild.rs:55-70).
- brickwork: random brickwork, per layer Rx, Ry, Rz (angles U[0, 2 pi)) on every
  qubit, then CNOT(2i -> 2i+1) in even layers and CNOT(2i+1 -> 2i+2) in odd ones
  (28 qubits, depth 20: 1,950 gates). yao-rs has no such builder; this one fixes
  the gate order and the angle stream (numpy's PCG64 from a seed).
- noisy_brickwork: the brickwork on a density matrix with Depolarizing(p) then
  AmplitudeDamping(gamma) after every gate on each touched qubit.
"""
import math

import numpy as np

from .circuit import Gate, NoiseChannel, PositionedGate, channel, control, put


def qft_with_swaps(n):
    """QFT with the final bit-reversal SWAPs (examples/qft.rs:30-54)."""
    els = qft_no_swaps(n)
    for i in range(n // 2):
        els.append(PositionedGate(Gate.SWAP, [i, n - 1 - i]))
    return els


def qft_no_swaps(n):
    """easybuild::qft_circuit (easybuild.rs:55-70): H on loc i, then
    Phase(2 pi / 2^(j+1)) on loc i controlled by loc i + j."""
    els = []
    for i in range(n):
        els.append(put([i], Gate.H))
        for j in range(1, n - i):
            els.append(control([i + j], [i], Gate.Phase(2.0 * math.pi / float(1 << (j + 1)))))
    return els


def brickwork(n, depth, seed):
    """Random brickwork: per layer Rx, Ry, Rz (angles U[0, 2pi)) on every qubit, then
    CNOT(2i -> 2i+1) in even layers and CNOT(2i+1 -> 2i+2) in odd layers."""
    rng = np.random.default_rng(seed)
    els = []
    for layer in range(depth):
        for q in range(n):
            a = rng.uniform(0.0, 2.0 * math.pi, size=3)
            els.append(put([q], Gate.Rx(a[0])))
            els.append(put([q], Gate.Ry(a[1])))
            els.append(put([q], Gate.Rz(a[2])))
        start = 0 if layer % 2 == 0 else 1
        for c in range(start, n - 1, 2):
            els.append(control([c], [c + 1], Gate.X))
    return els


def brickwork_layers(n, depth, seed, layers):
    """The first `layers` layers of brickwork(n, depth, seed): the same gates and
    angles as the full circuit's prefix."""
    per_layer = [3 * n + len(range(l % 2, n - 1, 2)) for l in range(depth)]
    return brickwork(n, depth, seed)[: sum(per_layer[:layers])]


def noisy_brickwork(n, depth, seed, p=1e-3, gamma=1e-3):
    """Brickwork on a density matrix: after every gate, on each touched qubit, a
    Depolarizing(p) then AmplitudeDamping(gamma) channel."""
    out = []
    for e in brickwork(n, depth, seed):
        out.append(e)
        for q in list(e.target_locs) + list(e.control_locs):
            out.append(channel([q], NoiseChannel.Depolarizing(1, p)))
            out.append(channel([q], NoiseChannel.AmplitudeDamping(gamma, 0.0)))
    return out
