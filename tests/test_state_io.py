"""yao-state-v1 header handling on the host (no device needed): files from an
independent writer are accepted, and every rejection of read_state_from_reader
(yao-cli/src/state_io.rs:63-88) is reproduced with its message."""
import pytest

import yaohip
from state_files import header_bytes, state_bytes
from yaohip import _abi as A


def write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def test_accepts_reference_encoding(tmp_path):
    p = write(tmp_path, "a.bin", state_bytes(2, [0.5, 0.5j, 0.5, -0.5j]))
    assert yaohip.read_state_header(p) == 2
    p = write(tmp_path, "b.bin", header_bytes(3, extra={"note": {"x": [1, 2]}}) + bytes(128))
    assert yaohip.read_state_header(p) == 3
    spaced = b'{ "dtype" : "complex128", "num_elements" : 4 , "dims" : [ 2 , 2 ], "num_qubits": 2,\n'
    spaced = spaced[:-1] + b' "format": "yao-state-v1" }\n'
    p = write(tmp_path, "c.bin", spaced + bytes(64))
    assert yaohip.read_state_header(p) == 2


@pytest.mark.parametrize("kw,needle", [
    (dict(num_qubits=2, num_elements=999), "num_elements=999"),
    (dict(num_qubits=2, fmt="yao-state-v2"), "Unknown state format: yao-state-v2"),
    (dict(num_qubits=2, dtype="complex64"), "Unsupported dtype: complex64"),
    (dict(num_qubits=2, dims=[3, 2], num_elements=6), "Only qubit (d=2) states are supported"),
    (dict(num_qubits=3, dims=[2, 2], num_elements=4), "2^num_qubits"),
])
def test_rejections(tmp_path, kw, needle):
    p = write(tmp_path, "bad.bin", header_bytes(**kw) + bytes(64))
    with pytest.raises(A.YhError) as e:
        yaohip.read_state_header(p)
    assert e.value.status == 16 and needle in e.value.message


def test_io_errors(tmp_path):
    with pytest.raises(A.YhError) as e:
        yaohip.read_state_header(str(tmp_path / "missing.bin"))
    assert e.value.status == 15
    for bad in (b"not json\n", b'{"format":"yao-state-v1"}\n', b""):
        p = write(tmp_path, "h.bin", bad)
        with pytest.raises(A.YhError):
            yaohip.read_state_header(p)


def test_deeply_nested_unknown_field(tmp_path):
    """An unknown field nested far deeper than serde_json's recursion limit (128) is
    rejected as a format error, not by overflowing the host stack."""
    deep = b'{"format": "yao-state-v1", "note": ' + b"[" * 200000 + b"]" * 200000 + b"}\n"
    p = write(tmp_path, "deep.bin", deep)
    with pytest.raises(A.YhError) as e:
        yaohip.read_state_header(p)
    assert e.value.status == 16
    ok = header_bytes(2, extra={"note": [[[[1]]]]}) + bytes(64)
    assert yaohip.read_state_header(write(tmp_path, "ok.bin", ok)) == 2
