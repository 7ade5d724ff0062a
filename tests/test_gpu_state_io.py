"""yao-state-v1 files to and from device registers (yh_state_write / yh_state_read):
the written bytes equal the independent reference encoding, round trips are
bit-exact (complex128) through the chunked pinned-buffer path, complex64
registers widen and narrow exactly as a cast does, and the reference's own
state_io tests pass."""
import numpy as np
import pytest

from conftest import seeded_state
from state_files import header_bytes, state_bytes

pytestmark = pytest.mark.gpu


def test_write_matches_reference_encoding(gpu, tmp_path):
    n = 10
    psi = seeded_state(n, 1)
    reg = gpu.ArrayReg.from_vec(n, psi)
    p = tmp_path / "s.bin"
    gpu.write_state(reg, str(p))
    assert p.read_bytes() == state_bytes(n, psi)
    back = gpu.read_state(str(p))
    assert np.array_equal(back.state_vec(), psi)


def test_reference_state_io_cases(gpu, tmp_path):
    # test_state_round_trip_file / _bytes (yao-cli/src/state_io.rs:122-163)
    v = np.array([0.5, 0.5j, 0.5, -0.5j])
    p = tmp_path / "t.bin"
    gpu.write_state(gpu.ArrayReg.from_vec(2, v), str(p))
    r = gpu.read_state(str(p))
    assert r.nqubits() == 2 and np.max(np.abs(r.state_vec() - v)) < 1e-12
    gpu.write_state(gpu.ArrayReg.zero_state(3), str(p))
    r = gpu.read_state(str(p))
    z = np.zeros(8)
    z[0] = 1
    assert r.nqubits() == 3 and np.array_equal(r.state_vec(), z)


def test_large_state_chunked_round_trip(gpu, tmp_path):
    """26 qubits = 1 GiB: 16 chunks of 64 MiB through the two pinned buffers."""
    n = 26
    rng = np.random.default_rng(26)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    reg = gpu.ArrayReg.from_vec(n, psi)
    p = tmp_path / "big.bin"
    gpu.write_state(reg, str(p))
    raw = np.fromfile(str(p), dtype=np.uint8)
    hdr = header_bytes(n)
    assert raw[:len(hdr)].tobytes() == hdr
    body = raw[len(hdr):].view("<f8").view(np.complex128)
    assert np.array_equal(body, psi)
    back = gpu.read_state(str(p))
    assert np.array_equal(back.state_vec(), psi)


def test_complex64_registers(gpu, tmp_path):
    n = 12
    psi = seeded_state(n, 5)
    p = tmp_path / "c.bin"
    p.write_bytes(state_bytes(n, psi))
    r64 = gpu.read_state(str(p), "c64")
    want = psi.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(r64.state_vec(), want)
    q = tmp_path / "d.bin"
    gpu.write_state(r64, str(q))
    assert q.read_bytes() == state_bytes(n, want)


def test_truncated_and_bad_files(gpu, tmp_path):
    import yaohip
    p = tmp_path / "short.bin"
    p.write_bytes(state_bytes(4, np.ones(16))[:-8])
    with pytest.raises(yaohip.YhError) as e:
        gpu.read_state(str(p))
    assert "Failed to read state data" in e.value.message
    p.write_bytes(header_bytes(2, num_elements=999) + bytes(64))
    with pytest.raises(yaohip.YhError) as e:
        gpu.read_state(str(p))
    assert "num_elements=999" in e.value.message
