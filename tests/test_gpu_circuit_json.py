"""The reference CLI's example circuits (tests/golden/cli_examples, read through the
circuit JSON mirror) on the GPU: probabilities and the QAOA expectation against
the outputs the reference CLI wrote, gate by gate and through the fused block
path. Bar: 1e-12."""
import numpy as np
import pytest

from test_circuit_json import NAMES, example, parse_operator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused", [0, 5])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_reference_cli_outputs(gpu, name, fused):
    circ, res = example(name)
    reg = gpu.ArrayReg.zero_state(circ.nbits)
    gpu.apply_inplace(circ, reg, fused)
    if "probabilities" in res:
        np.testing.assert_allclose(gpu.probs(reg), res["probabilities"], rtol=0, atol=1e-12)
    else:
        v = gpu.expect_arrayreg(reg, parse_operator(res["operator"]))
        assert abs(v - complex(res["expectation_value"]["re"], res["expectation_value"]["im"])) < 1e-12
