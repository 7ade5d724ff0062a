"""yaohip — MI355X (gfx950) backend for yao-rs's dense simulation path.

The names mirror yao-rs's public API (src/lib.rs:25-48). All simulation runs in
libyaohip.so through the C ABI of include/yaohip.h.
"""
from . import _abi
from .circuit import (Annotation, Circuit, CircuitError, Gate, NoiseChannel, Op, OperatorPolynomial,
                      OperatorString, PositionedChannel, PositionedGate, channel, control, label, lower,
                      put, tfim_hamiltonian)
from .circuit_json import CircuitJsonError, circuit_from_json, circuit_to_json
from .register import (ArrayReg, DensityMatrix, PostProcess, apply, apply_inplace, channel_kraus,
                       density_matrix_from_reg, expect_arrayreg, expect_dm, expect_grad, gate_matrix, measure_with_postprocess, probs,
                       read_state, read_state_header, sample, write_state)

YhError = _abi.YhError

__all__ = [
    "Annotation", "ArrayReg", "Circuit", "CircuitError", "CircuitJsonError", "DensityMatrix", "Gate", "NoiseChannel", "Op",
    "OperatorPolynomial", "OperatorString", "PositionedChannel", "PositionedGate", "PostProcess", "YhError",
    "apply", "apply_inplace", "channel", "channel_kraus", "circuit_from_json", "density_matrix_from_reg", "circuit_to_json", "control", "expect_arrayreg", "expect_dm", "expect_grad",
    "gate_matrix", "label", "lower", "measure_with_postprocess", "probs", "put", "read_state", "read_state_header",
    "sample", "tfim_hamiltonian", "write_state",
]
