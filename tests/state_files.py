"""An independent writer of yao-state-v1 files (yao-cli/src/state_io.rs:18-47):
serde's compact JSON of the header struct, in field order, a newline, then
little-endian (re, im) f64 pairs. Used by the state-file tests as the reference
encoding."""
import json
import struct

import numpy as np


def header_bytes(num_qubits, dims=None, num_elements=None, fmt="yao-state-v1", dtype="complex128", extra=None):
    dims = [2] * num_qubits if dims is None else dims
    h = {"format": fmt, "num_qubits": num_qubits, "dims": dims,
         "num_elements": (1 << num_qubits) if num_elements is None else num_elements, "dtype": dtype}
    if extra:
        h.update(extra)
    return json.dumps(h, separators=(",", ":")).encode() + b"\n"


def state_bytes(num_qubits, amps):
    amps = np.asarray(amps, dtype=np.complex128)
    body = b"".join(struct.pack("<dd", a.real, a.imag) for a in amps) if amps.size <= 4096 else \
        np.ascontiguousarray(amps).view("<f8").tobytes()
    return header_bytes(num_qubits) + body
