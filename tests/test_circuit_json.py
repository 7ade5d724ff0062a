"""Circuit JSON (src/json.rs) in the host mirror, and the reference CLI's example
circuits as end-to-end fixtures: tests/golden/cli_examples holds the circuits and
the probabilities / expectation the reference CLI computed for them
(docs/src/examples/generated/{circuits,results}, copied verbatim). Here the oracle
is pinned to those outputs; the GPU runs the same circuits in
test_gpu_circuit_json.py. Error cases follow src/unit_tests/json.rs."""
import glob
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from yaohip.circuit import (Annotation, Circuit, Gate, Op, OperatorPolynomial, OperatorString, PositionedGate,
                            channel, control, label, lower, put, NoiseChannel)
from yaohip.circuit_json import CircuitJsonError, circuit_from_json, circuit_to_json

EXAMPLES = os.path.join(GOLDEN, "cli_examples")
NAMES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(EXAMPLES, "*.json"))
               if not p.endswith(("-probs.json", "-expect.json")))


def example(name):
    with open(os.path.join(EXAMPLES, name + ".json")) as f:
        circ = circuit_from_json(f.read())
    res = None
    for suffix in ("-probs.json", "-expect.json"):
        p = os.path.join(EXAMPLES, name + suffix)
        if os.path.exists(p):
            with open(p) as f:
                res = json.load(f)
    return circ, res


def parse_operator(text):
    """The CLI's operator strings used by the examples, e.g. "Z(0)Z(1)"."""
    ops = []
    for part in text.replace(" ", "").split(")")[:-1]:
        name, site = part.split("(")
        ops.append((int(site), getattr(Op, name)))
    return OperatorPolynomial([1.0], [OperatorString(ops)])


def test_every_example_has_a_result():
    assert len(NAMES) == 11
    for name in NAMES:
        assert example(name)[1] is not None, name


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_reference_cli_outputs(name):
    circ, res = example(name)
    n = circ.nbits
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    psi = O.apply(n, z, lower(circ.elements))
    if "probabilities" in res:
        assert res["num_qubits"] == n and res["locs"] is None
        np.testing.assert_allclose(O.probs(n, psi), res["probabilities"], rtol=0, atol=1e-13)
    else:
        v = O.expect(n, psi, parse_operator(res["operator"]))
        want = complex(res["expectation_value"]["re"], res["expectation_value"]["im"])
        assert abs(v - want) < 1e-13


@pytest.mark.parametrize("name", NAMES)
def test_round_trip_keeps_every_element(name):
    circ, _ = example(name)
    back = circuit_from_json(circuit_to_json(circ))
    assert back.nbits == circ.nbits and len(back.elements) == len(circ.elements)
    for a, b in zip(circ.elements, back.elements):
        assert type(a) is type(b)
        if isinstance(a, PositionedGate):
            assert (a.gate.kind, a.gate.params, a.target_locs, a.control_locs, a.control_configs) == \
                   (b.gate.kind, b.gate.params, b.target_locs, b.control_locs, b.control_configs)


@pytest.mark.parametrize("name", ["visualization-example-qft", "visualization-example-basic"])
def test_visualization_examples_parse_and_round_trip(name):
    """The circuits of the reference's visualisation page (visualization/example-*.json,
    copied as *.circuit): they parse, round-trip, and the oracle keeps their norm."""
    with open(os.path.join(EXAMPLES, name + ".circuit")) as f:
        text = f.read()
    c = circuit_from_json(text)
    assert json.loads(circuit_to_json(c)) == json.loads(circuit_to_json(circuit_from_json(circuit_to_json(c))))
    n = c.nbits
    z = np.zeros(1 << n, dtype=complex)
    z[0] = 1
    assert abs(np.linalg.norm(O.apply(n, z, lower(c.elements))) - 1) < 1e-14


def test_structure_and_optional_fields():
    c = Circuit.qubits(2, [put([0], Gate.H), control([0], [1], Gate.X), label(1, "test"),
                           put([1], Gate.Rx(0.5)), channel([0], NoiseChannel.BitFlip(0.1))])
    doc = json.loads(circuit_to_json(c))
    assert doc["num_qubits"] == 2
    e = doc["elements"]
    assert len(e) == 4  # the channel is dropped
    assert e[0] == {"type": "gate", "gate": "H", "targets": [0]}
    assert e[1] == {"type": "gate", "gate": "X", "targets": [1], "controls": [0], "control_configs": [True]}
    assert e[2] == {"type": "label", "text": "test", "loc": 1}
    assert e[3] == {"type": "gate", "gate": "Rx", "params": [0.5], "targets": [1]}
    assert list(e[1].keys()) == ["type", "gate", "targets", "controls", "control_configs"]


def test_custom_and_fsim_round_trip():
    m = np.array([[1, 0], [0, np.exp(0.3j)]])
    c = Circuit.qubits(2, [put([0], Gate.Custom(m, True, "DiagGate")), put([0, 1], Gate.FSim(1.0, 0.5))])
    back = circuit_from_json(circuit_to_json(c))
    g0, g1 = back.elements[0].gate, back.elements[1].gate
    assert g0.label == "DiagGate" and g0.is_diagonal
    np.testing.assert_array_equal(g0.matrix, m)
    assert g1.params == (1.0, 0.5)


def test_aliases_and_default_control_configs():
    c = circuit_from_json('{"num_qubits": 3, "elements": [{"type": "gate", "gate": "CNOT", "targets": [2], '
                          '"controls": [0, 1]}, {"type": "gate", "gate": "CX", "targets": [0], "controls": [1]}]}')
    assert [e.gate.kind for e in c.elements] == [Gate(0).kind] * 2
    assert c.elements[0].control_configs == [True, True]


@pytest.mark.parametrize("text,msg", [
    ("not valid json{{{", "JSON parse error"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "FooBar", "targets": [0]}]}',
     "Unknown gate type: FooBar"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Phase", "targets": [0]}]}',
     "Phase gate requires params"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Phase", "targets": [0], "params": []}]}',
     "Phase gate requires at least 1 parameter"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Rx", "targets": [0]}]}', "Rx gate requires params"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Ry", "targets": [0]}]}', "Ry gate requires params"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Rz", "targets": [0]}]}', "Rz gate requires params"),
    ('{"num_qubits": 2, "elements": [{"type": "gate", "gate": "FSim", "targets": [0, 1]}]}',
     "FSim gate requires params"),
    ('{"num_qubits": 2, "elements": [{"type": "gate", "gate": "FSim", "targets": [0, 1], "params": [1.0]}]}',
     "FSim gate requires 2 parameters"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Custom", "targets": [0]}]}',
     "Custom gate requires matrix"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Custom", "targets": [0], "matrix": []}]}',
     "Custom gate matrix cannot be empty"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "Custom", "targets": [0], '
     '"matrix": [[[1,0],[0,0]], [[0,0]]]}]}', "Custom gate matrix rows must have equal length"),
    ('{"num_qubits": 1, "elements": [{"type": "gate", "gate": "H", "targets": [5]}]}', "Circuit validation error"),
])
def test_reference_error_messages(text, msg):
    with pytest.raises(CircuitJsonError) as e:
        circuit_from_json(text)
    assert msg in str(e.value)
