"""Builders shared by the tests: circuits from the reference's JSON fixtures
(mirroring tests/common/mod.rs:154-240 of yao-rs) and random circuits of every
dispatch class. The benchmark workloads live in yaohip.workloads and are
re-exported here."""
import math

import numpy as np

from yaohip.circuit import Gate, PositionedGate, put
from yaohip.workloads import (brickwork, brickwork_layers, noisy_brickwork, qft_no_swaps,  # noqa: F401
                              qft_with_swaps)

NAMED = {
    "X": lambda p: Gate.X, "Y": lambda p: Gate.Y, "Z": lambda p: Gate.Z, "H": lambda p: Gate.H,
    "S": lambda p: Gate.S, "T": lambda p: Gate.T, "SWAP": lambda p: Gate.SWAP,
    "SqrtX": lambda p: Gate.SqrtX, "SqrtY": lambda p: Gate.SqrtY, "SqrtW": lambda p: Gate.SqrtW,
    "ISWAP": lambda p: Gate.ISWAP, "Phase": lambda p: Gate.Phase(p[0]), "Rx": lambda p: Gate.Rx(p[0]),
    "Ry": lambda p: Gate.Ry(p[0]), "Rz": lambda p: Gate.Rz(p[0]), "FSim": lambda p: Gate.FSim(p[0], p[1]),
}


def gate_from_spec(spec):
    """gate_from_spec (tests/common/mod.rs:154-188)."""
    if spec["name"] == "Custom":
        m = np.array(spec["matrix_re"]) + 1j * np.array(spec["matrix_im"])
        diag = all(i == j or abs(m[i, j]) < 1e-15 for i in range(len(m)) for j in range(len(m)))
        return Gate.Custom(m, diag, spec.get("label", ""))
    return NAMED[spec["name"]](spec["params"])


def elements_from_case(case):
    """circuit_from_case (tests/common/mod.rs:216-240)."""
    els = []
    for spec in case["gates"]:
        g = gate_from_spec(spec)
        if not spec["controls"]:
            els.append(put(spec["targets"], g))
        else:
            cfg = spec["control_configs"] or [True] * len(spec["controls"])
            els.append(PositionedGate(g, spec["targets"], spec["controls"], cfg))
    return els


def parse_state(pairs):
    return np.array([complex(a, b) for a, b in pairs], dtype=np.complex128)


def parse_matrix(rows):
    return np.array([[complex(a, b) for a, b in row] for row in rows], dtype=np.complex128)


def random_circuit(n, ngates, seed, max_targets=3, max_ctrls=2, custom_targets=(1, 2, 3)):
    """Random gates of every dispatch class: named gates, controlled (active high and
    low), diagonal and dense Custom matrices on 1..max_targets targets, SWAPs."""
    rng = np.random.default_rng(seed)
    kinds = ["X", "Y", "Z", "H", "S", "T", "Phase", "Rx", "Ry", "Rz", "SqrtX", "SqrtY", "SqrtW",
             "CustomDense", "CustomDiag"]
    if n >= 2:
        kinds += ["SWAP", "ISWAP", "FSim"]
    els = []
    for _ in range(ngates):
        k = kinds[rng.integers(len(kinds))]
        if k in ("SWAP", "ISWAP", "FSim"):
            nt = 2
        elif k in ("CustomDense", "CustomDiag"):
            nt = int(rng.choice([t for t in custom_targets if t <= n]))
        else:
            nt = 1
        nc = int(rng.integers(0, max_ctrls + 1)) if n - nt > 0 else 0
        nc = min(nc, n - nt)
        locs = rng.permutation(n)[: nt + nc]
        tl, cl = [int(x) for x in locs[:nt]], [int(x) for x in locs[nt:]]
        cfg = [bool(rng.integers(2)) for _ in cl]
        th = rng.uniform(-math.pi, math.pi, size=2)
        if k == "CustomDense":
            d = 1 << nt
            a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
            q, _ = np.linalg.qr(a)
            g = Gate.Custom(q, False)
        elif k == "CustomDiag":
            d = 1 << nt
            g = Gate.Custom(np.diag(np.exp(1j * rng.uniform(-math.pi, math.pi, size=d))), True)
        else:
            g = NAMED[k](th)
        els.append(PositionedGate(g, tl, cl, cfg))
    return els
