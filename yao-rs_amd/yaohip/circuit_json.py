"""Circuit JSON: the circuit format of yao-rs's CLI and visualisation
(src/json.rs:7-38 for the schema, :115-138 circuit_to_json, :141-246
circuit_from_json).

{"num_qubits": n, "elements": [...]} where an element is either
{"type": "gate", "gate": name, "params"?: [f64], "targets": [loc],
 "controls"?: [loc], "control_configs"?: [bool], "label"?: str,
 "matrix"?: [[[re, im], ...], ...], "is_diagonal"?: bool}
or {"type": "label", "text": str, "loc": loc}.

Reading follows circuit_from_json: the same gate names (X also as CNOT and CX),
the same error messages (JSON parse error, Unknown gate type, missing or short
params, Custom matrix checks, Circuit validation error), control_configs
defaulting to all true. Writing follows circuit_to_json: channels are dropped,
optional fields are left out when absent, two-space indentation.
"""
import json

import numpy as np

from . import _abi as A
from .circuit import Annotation, Circuit, CircuitError, Gate, PositionedGate

_NO_PARAM = {"X": A.GATE_X, "CNOT": A.GATE_X, "CX": A.GATE_X, "Y": A.GATE_Y, "Z": A.GATE_Z, "H": A.GATE_H,
             "S": A.GATE_S, "T": A.GATE_T, "SWAP": A.GATE_SWAP, "SqrtX": A.GATE_SQRTX, "SqrtY": A.GATE_SQRTY,
             "SqrtW": A.GATE_SQRTW, "ISWAP": A.GATE_ISWAP}
_ONE_PARAM = {"Phase": Gate.Phase, "Rx": Gate.Rx, "Ry": Gate.Ry, "Rz": Gate.Rz}


class CircuitJsonError(ValueError):
    """The Err(String) of circuit_from_json; str(e) is the reference's message."""


def _gate_from_json(gj):
    name = gj.get("gate")
    params = gj.get("params")
    if name in _NO_PARAM:
        gate = Gate(_NO_PARAM[name])
    elif name in _ONE_PARAM:
        if params is None:
            raise CircuitJsonError(f"{name} gate requires params")
        if len(params) == 0:
            raise CircuitJsonError(f"{name} gate requires at least 1 parameter")
        gate = _ONE_PARAM[name](params[0])
    elif name == "FSim":
        if params is None:
            raise CircuitJsonError("FSim gate requires params")
        if len(params) < 2:
            raise CircuitJsonError("FSim gate requires 2 parameters")
        gate = Gate.FSim(params[0], params[1])
    elif name == "Custom":
        rows = gj.get("matrix")
        if rows is None:
            raise CircuitJsonError("Custom gate requires matrix")
        if len(rows) == 0:
            raise CircuitJsonError("Custom gate matrix cannot be empty")
        ncols = len(rows[0])
        if any(len(r) != ncols for r in rows):
            raise CircuitJsonError("Custom gate matrix rows must have equal length")
        m = np.array([[complex(re, im) for re, im in r] for r in rows], dtype=np.complex128).reshape(len(rows), ncols)
        gate = Gate.Custom(m, bool(gj.get("is_diagonal") or False), gj.get("label") or "")
    else:
        raise CircuitJsonError(f"Unknown gate type: {name}")
    controls = list(gj.get("controls") or [])
    configs = gj.get("control_configs")
    configs = [True] * len(controls) if configs is None else [bool(c) for c in configs]
    return PositionedGate(gate, list(gj["targets"]), controls, configs)


def circuit_from_json(text):
    """circuit_from_json (json.rs:229-246): a validated Circuit, or CircuitJsonError."""
    try:
        doc = json.loads(text)
        nq = int(doc["num_qubits"])
        raw = list(doc["elements"])
    except (ValueError, KeyError, TypeError) as e:
        raise CircuitJsonError(f"JSON parse error: {e}") from None
    elements = []
    for ej in raw:
        kind = ej.get("type") if isinstance(ej, dict) else None
        if kind == "gate":
            elements.append(_gate_from_json(ej))
        elif kind == "label":
            elements.append(Annotation(ej["text"], int(ej["loc"])))
        else:
            raise CircuitJsonError(f"JSON parse error: unknown variant {kind!r}, expected `gate` or `label`")
    try:
        return Circuit.qubits(nq, elements)
    except CircuitError as e:
        raise CircuitJsonError(f"Circuit validation error: {e}") from None


def _gate_to_json(pg):
    g = pg.gate
    out = {"type": "gate", "gate": A.GATE_NAMES[g.kind]}
    if g.kind in (A.GATE_PHASE, A.GATE_RX, A.GATE_RY, A.GATE_RZ, A.GATE_FSIM):
        out["params"] = [float(p) for p in g.params]
    out["targets"] = list(pg.target_locs)
    if pg.control_locs:
        out["controls"] = list(pg.control_locs)
    if pg.control_configs:
        out["control_configs"] = [bool(c) for c in pg.control_configs]
    if g.kind == A.GATE_CUSTOM:
        out["label"] = g.label
        out["matrix"] = [[[float(z.real), float(z.imag)] for z in row] for row in g.matrix]
        out["is_diagonal"] = bool(g.is_diagonal)
    return out


def circuit_to_json(circuit):
    """circuit_to_json (json.rs:115-138): gates and labels; channels are dropped."""
    elements = []
    for e in circuit.elements:
        if isinstance(e, PositionedGate):
            elements.append(_gate_to_json(e))
        elif isinstance(e, Annotation):
            elements.append({"type": "label", "text": e.label, "loc": e.loc})
    return json.dumps({"num_qubits": circuit.nbits, "elements": elements}, indent=2)
