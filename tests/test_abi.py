"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/yaohip.h declares, and its host-side logic (validation, gate matrices,
Kraus operators) agrees with the oracle. No device call is made."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from fixture_circuits import NAMED, random_circuit
from oracle import oracle as O
from yaohip import _abi as A
from yaohip.circuit import (Circuit, CircuitError, Gate, NoiseChannel, PositionedGate, channel, control, lower,
                            put)
from test_oracle import KRAUS_FIXTURES


def declared_symbols():
    with open(os.path.join(ROOT, "include", "yaohip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?(yh_[a-z0-9_]+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = A.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s
    # the ctypes signature table covers the header exactly
    assert sorted(n for n, _, _ in A.SIGNATURES) == syms


def test_struct_layout_matches_header():
    # int32 x 4 + 12 + 16 + 16 int32, 4 doubles, pointer, 4 int32
    assert ctypes.sizeof(A.YhOp) == 4 * 4 + 4 * (12 + 16 + 16) + 8 * 4 + 8 + 4 * 4


def test_gate_matrices_bitwise_equal_to_oracle():
    L = A.lib()
    for name, mk in NAMED.items():
        for th in ([0.0], [0.7], [-2.5], [1e-10], [np.pi]):
            g = mk(th + [0.3])
            ol = lower([put(list(range(g.num_sites())), g)])
            out = np.zeros(32)
            dim = ctypes.c_int()
            A.check(L.yh_gate_matrix(ctypes.byref(ol.arr[0]), A.dptr(out), ctypes.byref(dim)))
            d = dim.value
            mine = A.deinterleave(out[:2 * d * d]).reshape(d, d)
            assert np.array_equal(mine, O.gate_matrix(ol.arr[0])), name


@pytest.mark.parametrize("name", sorted(KRAUS_FIXTURES))
def test_kraus_bitwise_equal_to_oracle(name):
    from yaohip.register import channel_kraus
    ks = channel_kraus(KRAUS_FIXTURES[name])
    ref = O.kraus(lower([channel([0], KRAUS_FIXTURES[name])]))
    assert len(ks) == len(ref)
    for a, b in zip(ks, ref):
        assert np.array_equal(a, b)


def test_kraus_multi_qubit_depolarizing_and_custom():
    from yaohip.register import channel_kraus
    ch = NoiseChannel.Depolarizing(2, 0.2)
    ks = channel_kraus(ch, [0, 1])
    ref = O.kraus(lower([channel([0, 1], ch)]))
    assert len(ks) == 16 and all(np.array_equal(a, b) for a, b in zip(ks, ref))
    k0 = np.diag([1.0, np.sqrt(0.5)])
    k1 = np.array([[0, np.sqrt(0.5)], [0, 0]])
    ks = channel_kraus(NoiseChannel.Custom([k0, k1]))
    assert np.array_equal(ks[0], k0) and np.array_equal(ks[1], k1)


def test_channel_parameter_validation():
    from yaohip.register import channel_kraus
    for bad in [NoiseChannel.BitFlip(1.5), NoiseChannel.PauliChannel(0.5, 0.5, 0.5),
                NoiseChannel.AmplitudeDamping(-0.1), NoiseChannel.ThermalRelaxation(1.0, 3.0, 1.0),
                NoiseChannel.Reset(0.7, 0.6)]:
        with pytest.raises(A.YhError) as e:
            channel_kraus(bad)
        assert e.value.status == 8


@pytest.mark.parametrize("seed", range(4))
def test_validation_agrees_with_oracle_on_random_circuits(seed):
    n = 6
    els = random_circuit(n, 30, seed)
    ol = lower(els)
    assert A.lib().yh_validate(n, ol.arr, ol.n) == O.validate(n, ol) == 0
    rng = np.random.default_rng(seed)
    for _ in range(50):
        e = els[rng.integers(len(els))]
        bad = PositionedGate(e.gate, list(e.target_locs), list(e.control_locs), list(e.control_configs))
        kind = rng.integers(4)
        if kind == 0:
            bad.target_locs[0] = n + int(rng.integers(3))
        elif kind == 1 and bad.control_locs:
            bad.control_locs[0] = bad.target_locs[0]
        elif kind == 2:
            bad.control_configs = bad.control_configs + [True]
        else:
            bad.target_locs = bad.target_locs + [x for x in range(n) if x not in bad.target_locs
                                                 and x not in bad.control_locs][:1]
        ob = lower([bad])
        assert A.lib().yh_validate(n, ob.arr, ob.n) == O.validate(n, ob)


def test_circuit_error_variants():
    with pytest.raises(CircuitError) as e:
        Circuit.qubits(2, [PositionedGate(Gate.X, [1], [0], [True, True])])
    assert e.value.variant == "ControlConfigLengthMismatch"
    with pytest.raises(CircuitError) as e:
        Circuit.qubits(2, [put([5], Gate.H)])
    assert e.value.variant == "LocOutOfRange"
    with pytest.raises(CircuitError) as e:
        Circuit.qubits(2, [control([0], [0], Gate.X)])
    assert e.value.variant == "OverlappingLocs"
    with pytest.raises(CircuitError) as e:
        Circuit.qubits(2, [put([0, 1], Gate.H)])
    assert e.value.variant == "MatrixSizeMismatch"


def test_no_device_here_fails_loudly():
    """Without a GPU the library reports NO_DEVICE instead of computing on the CPU."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    st = A.lib().yh_reg_create(4, A.C128, ctypes.byref(h))
    assert st == 14


def test_tile_plan_covers_every_gate():
    """Host-only planner: every gate of the 28-qubit depth-20 brickwork lands in a
    tile pass, with far fewer passes than gates."""
    from fixture_circuits import brickwork
    els = brickwork(28, 20, 20)
    ol = lower(els)
    out = np.zeros(4, dtype=np.int64)
    A.check(A.lib().yh_plan_stats(28, ol.arr, ol.n, A.i64ptr(out)))
    passes, groups, tiled, single = [int(x) for x in out]
    assert tiled == len(els) and single == 0
    assert passes <= 60 and groups >= passes


def test_block_plan_covers_every_gate():
    """Host-only block planner on the 28-qubit depth-20 brickwork: every gate lands
    in a block; far fewer blocks than gates (blocks of 4 or 5 bits). Wide gates run
    alone."""
    from fixture_circuits import brickwork
    els = brickwork(28, 20, 20)
    ol = lower(els)
    out = np.zeros(3, dtype=np.int64)
    A.check(A.lib().yh_block_plan_stats(28, ol.arr, ol.n, A.i64ptr(out)))
    blocks, inblk, single = [int(x) for x in out]
    assert inblk == len(els) and single == 0
    assert blocks <= 80
    wide = [PositionedGate(Gate.Custom(np.eye(8), False), [0, 1, 2], [3, 4, 5], [True, True, False])]
    ol = lower(els[:10] + wide)
    A.check(A.lib().yh_block_plan_stats(28, ol.arr, ol.n, A.i64ptr(out)))
    assert int(out[2]) == 1 and int(out[1]) == 10


def test_block_plan_bits():
    """yh_block_plan lists each block's 5 distinct memory bits, in plan order, and
    -1 rows for gates run alone; every gate's bits lie inside some block."""
    from fixture_circuits import brickwork
    n = 28
    els = brickwork(n, 4, 5)
    wide = PositionedGate(Gate.Custom(np.eye(8), False), [0, 1, 2], [3, 4, 5], [True, True, False])
    ol = lower(els + [wide])
    ns = ctypes.c_size_t()
    A.check(A.lib().yh_block_plan(n, ol.arr, ol.n, None, 0, ctypes.byref(ns)))
    bits = np.zeros(5 * ns.value, dtype=np.int32)
    A.check(A.lib().yh_block_plan(n, ol.arr, ol.n, A.i32ptr(bits), ns.value, ctypes.byref(ns)))
    rows = bits.reshape(-1, 5)
    singles = [r for r in rows if (r == -1).all()]
    blocks = [set(int(b) for b in r) for r in rows if (r >= 0).all()]
    assert len(singles) == 1 and len(singles) + len(blocks) == len(rows)
    assert all(len(b) == 5 and max(b) < n for b in blocks)
    for e in els:  # memory bit of loc l is n - 1 - l
        g = {n - 1 - l for l in list(e.target_locs) + list(e.control_locs)}
        assert any(g <= b for b in blocks)


def test_circuit_num_params_matches_python_mirror():
    from test_ad_oracle import random_param_circuit
    els = random_param_circuit(5, 60, 3)
    c = Circuit.qubits(5, els)
    n = ctypes.c_size_t()
    A.check(A.lib().yh_circuit_num_params(c.ops.arr, c.ops.n, ctypes.byref(n)))
    assert n.value == c.num_params() == len(c.parameters())


def _pass_stats(n, els, dm):
    c = Circuit.qubits(n, els)
    out = (ctypes.c_int64 * 7)()
    A.check(A.lib().yh_block_pass_stats(n, c.ops.arr, c.ops.n, dm, out))
    return list(out)


def _modelled(stats):
    """The planner's launch-time model, in 32 x 32 block-times: 0.46 per multi-block
    pass, 1 per 32 x 32 and 0.55 per 16 x 16 block in it, 1.8 per block alone, 1 per
    gate alone."""
    launches, multi, in_multi, small, din, dout, alone = stats
    return 0.46 * multi + (in_multi - small) + 0.55 * small + 1.8 * (launches - multi) + alone


def test_block_pass_stats_brickwork28():
    """The 28-qubit depth-20 brickwork's launch plan: every block is in exactly one
    pass; the plan is mostly 16 x 16 blocks (their time per gate is lower), in at
    most 20 HBM passes that read (first block) and write (last block) HBM directly;
    its modelled time is below that of the plan of 48 blocks of 5 bits in 23
    passes (57.8 block-times)."""
    from fixture_circuits import brickwork
    els = brickwork(28, 20, 20)
    st = _pass_stats(28, els, 0)
    launches, multi, in_multi, small, din, dout, alone = st
    plan = (ctypes.c_int64 * 3)()
    c = Circuit.qubits(28, els)
    A.check(A.lib().yh_block_plan_stats(28, c.ops.arr, c.ops.n, plan))
    assert alone == 0
    assert (launches - multi) + in_multi == plan[0]  # blocks alone + blocks in passes
    assert launches <= 20 and small >= in_multi // 2
    assert din >= multi - 2 and dout >= multi // 2
    assert _modelled(st) < 50


def test_block_pass_stats_density_matrix():
    """Noisy brickwork on a density matrix: every superoperator block acts on the
    row and column bits of at most two qubits, so every block in a multi-block
    pass is a 16 x 16 block, and its identity row lands above memory bits 0..3,
    so every pass reads HBM directly."""
    from fixture_circuits import noisy_brickwork
    launches, multi, in_multi, small, din, dout, alone = _pass_stats(14, noisy_brickwork(14, 10, 3), 1)
    assert alone == 0 and launches == multi and multi <= 20
    assert small == in_multi == 65
    assert din == multi and dout >= multi // 2


def test_block_pass_stats_invariants():
    """Random 16-qubit circuits: counts are consistent with the block plan."""
    from fixture_circuits import random_circuit
    for seed in range(3):
        els = random_circuit(16, 300, seed)
        launches, multi, in_multi, small, din, dout, alone = _pass_stats(16, els, 0)
        plan = (ctypes.c_int64 * 3)()
        c = Circuit.qubits(16, els)
        A.check(A.lib().yh_block_plan_stats(16, c.ops.arr, c.ops.n, plan))
        assert (launches - multi) + in_multi == plan[0]
        assert alone == plan[2]
        assert 0 < small < in_multi and din <= multi and dout <= multi
        assert in_multi <= 5 * multi


def test_block_plan_cache_keys_on_structure():
    """Plans are cached by the whole circuit structure: circuits with the same qubit
    and element counts but different positions get their own plans, and a cached
    plan is returned only for its own circuit (angles do not matter)."""
    from fixture_circuits import brickwork
    n = 16
    a = brickwork(n, 3, 1)
    b = [put([n - 1 - e.target_locs[0]], e.gate) if not e.control_locs else
         control([n - 1 - e.control_locs[0]], [n - 1 - e.target_locs[0]], e.gate) for e in a]
    a2 = brickwork(n, 3, 2)  # same structure, other angles

    def plan(els):
        ol = lower(els)
        ns = ctypes.c_size_t()
        A.check(A.lib().yh_block_plan(n, ol.arr, ol.n, None, 0, ctypes.byref(ns)))
        bits = np.zeros(5 * ns.value, dtype=np.int32)
        A.check(A.lib().yh_block_plan(n, ol.arr, ol.n, A.i32ptr(bits), ns.value, ctypes.byref(ns)))
        return bits.reshape(-1, 5)

    pa, pb = plan(a), plan(b)
    assert not np.array_equal(pa, pb)
    assert np.array_equal(plan(a), pa) and np.array_equal(plan(a2), pa) and np.array_equal(plan(b), pb)
    # every gate of b lies inside a block of b's plan
    blocks = [set(int(x) for x in r) for r in pb if (r >= 0).all()]
    for e in b:
        g = {n - 1 - l for l in list(e.target_locs) + list(e.control_locs)}
        assert any(g <= s for s in blocks)
